"""The allocation-free FP64 evaluation (ftk_stack_fp64_noalloc, through the C-ABI)
equals ftk_stack_fp64 bit for bit: two calls in a row, the second on a smaller input
(the shared scratch keeps the first call's partials, which the second overwrites), a
row range that needs several strips of the 64 MiB scratch, and a shape whose single
row of partials does not fit is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _bind(L):
    vp = C.c_void_p
    L.ftk_stack_fp64_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_size_t, C.c_int, vp, vp, C.c_int, C.c_int,
                                         C.c_int, vp, vp, vp, C.c_char_p, C.c_size_t]
    return L


def _case(seed, n, h, w, nv):
    rng = np.random.default_rng(seed)
    frames = po.test_normals(seed, (n, h, w))
    frames[rng.integers(0, n), rng.integers(0, h), rng.integers(0, w)] = np.nan
    span = 12
    dx = rng.integers(-span, span + 1, (nv, n)).astype(np.int32)
    dy = rng.integers(-span, span + 1, (nv, n)).astype(np.int32)
    return frames, dx, dy


def _noalloc(frames_d, dx, dy, min_count, y0, y1):
    import torch
    import fainttrail_amd as ft
    L = _bind(ft.lib())
    n, h, w = frames_d.shape
    ddx = torch.from_numpy(np.ascontiguousarray(dx)).cuda()
    ddy = torch.from_numpy(np.ascontiguousarray(dy)).cuda()
    out = torch.full((y1 - y0, w), -7.0, dtype=torch.float64, device="cuda")
    best = torch.full((y1 - y0, w), -7, dtype=torch.int32, device="cuda")
    e = C.create_string_buffer(512)
    s = torch.cuda.current_stream()
    rc = L.ftk_stack_fp64_noalloc(n, w, h, frames_d.data_ptr(), frames_d.stride(0), dx.shape[0], ddx.data_ptr(),
                                  ddy.data_ptr(), min_count, y0, y1, out.data_ptr(), best.data_ptr(), s.cuda_stream, e,
                                  len(e))
    torch.cuda.synchronize()
    return rc, e.value.decode(), out, best


@pytest.mark.parametrize("first,second", [((1, 9, 120, 200, 700, 0, 120), (2, 5, 40, 64, 150, 3, 31)),
                                          ((3, 12, 300, 1024, 6000, 0, 300), (4, 6, 90, 700, 300, 10, 80))])
def test_noalloc_fp64_equals_allocating_entry_point(hip, first, second):
    import torch
    for seed, n, h, w, nv, y0, y1 in (first, second):
        frames, dx, dy = _case(seed, n, h, w, nv)
        min_count = max(1, (2 * n) // 3)
        fd = torch.from_numpy(frames).cuda()
        chunks = -(-nv // 128)
        strips = -(-(y1 - y0) // ((64 << 20) // (w * chunks * 12)))
        print(f"fp64 no-alloc: seed {seed}, {n} x {w} x {h}, {nv} motions, rows {y0}-{y1}, {strips} strip(s)",
              flush=True)
        want_m, want_b = hip.stack_fp64(fd, dx, dy, min_count, y0, y1)
        rc, msg, got_m, got_b = _noalloc(fd, dx, dy, min_count, y0, y1)
        assert rc == 0, msg
        assert torch.equal(got_b, want_b)
        assert np.array_equal(got_m.cpu().numpy().view(np.uint64), want_m.cpu().numpy().view(np.uint64))
        assert int((got_b >= 0).sum()) > 0


def test_noalloc_fp64_refuses_a_row_beyond_the_scratch(hip):
    import torch
    n, h, w, nv = 2, 1, 60000, 12800  # one row of partials: 60000 x 100 x 12 bytes > 64 MiB
    frames = torch.zeros((n, h, w), dtype=torch.float32, device="cuda")
    dx = np.zeros((nv, n), np.int32)
    rc, msg, _, _ = _noalloc(frames, dx, dx, 1, 0, h)
    assert rc == 1 and "scratch" in msg, msg
