"""Stream-order rule of the native sources (CPU, reads the sources as text).

A synchronous host-to-device hipMemcpy from pageable memory may return before its DMA has
landed, and hipMemcpy / hipMemset run on the null stream, which the library's non-blocking
streams do not wait for. So every such call must be followed, before any kernel on another
stream can read the memory, by hipStreamSynchronize(nullptr). A search whose stage builds read
the matched-filter taps too early under a loaded server stacked empty frames; this keeps the
rule for every such call site."""
import re
from pathlib import Path

CSRC = Path(__file__).resolve().parent.parent / "fainttrail_amd" / "csrc"
WINDOW = 16  # lines after the call within which the null stream is synchronised


def _sites():
    for src in sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.hpp"))):
        lines = src.read_text().splitlines()
        for i, line in enumerate(lines):
            code = line.split("//")[0]
            call = re.search(r"\bhip(Memcpy|Memset)\(", code)
            if not call:
                continue
            stmt = " ".join(x.split("//")[0] for x in lines[i:i + 3])
            if call.group(1) == "Memcpy" and "hipMemcpyHostToDevice" not in stmt:
                continue  # device-to-host copies have landed when they return
            yield src.name, i + 1, lines[i + 1:i + 1 + WINDOW]


def test_null_stream_copies_and_clears_are_synchronised():
    sites = list(_sites())
    assert sites, "no synchronous copies or clears found: the scan is broken"
    missing = [(f, n) for f, n, after in sites if not any("hipStreamSynchronize(nullptr)" in x for x in after)]
    assert not missing, f"null-stream copy or clear without a null-stream synchronisation after it: {missing}"
