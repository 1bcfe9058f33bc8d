#!/bin/sh
# The reference-side integration of the "hip" backend as a built-in GPU backend, the
# way the reference builds its CUDA backend in (FAINTTRAIL_HAVE_CUDA): writes edited
# copies of the reference's src/backend.cpp and cli/main.cpp into OUT (a temporary
# build directory outside this repository; they are compiled from there and not kept).
# The edits, guarded by FAINTTRAIL_HAVE_HIP:
#   src/backend.cpp  backend_names() lists "hip" first when a device is present (so
#                    "auto" picks it, as it picks "cuda" first, :46-52); make_backend()
#                    takes "hip" and "hip:key=value,..." (the cuda:... option syntax) and
#                    gives "auto" / "" the HIP backend when a device is present (:54-87).
#   cli/main.cpp     check-backend runs its fourth check (row bands under a 2 MB cap,
#                    :735-740) for "hip" as it does for "cuda".
# The helpers hip_backend_available() / make_hip_backend() are in
# fainttrail_amd/integration/backend_hip.cpp.
# Usage: hip_integration.sh REFERENCE_DIR OUT_DIR
set -eu
REF=$1
OUT=$2
mkdir -p "$OUT"
sed \
  -e '/^std::vector<std::string> backend_names() {/i\
#ifdef FAINTTRAIL_HAVE_HIP\
namespace detail {\
bool hip_backend_available(std::string* why);\
std::unique_ptr<SearchBackend> make_hip_backend(const std::string& options);\
}  // namespace detail\
#endif' \
  -e '/names.push_back("cuda");  \/\/ GPU first/i\
#ifdef FAINTTRAIL_HAVE_HIP\
    if (detail::hip_backend_available(nullptr)) names.push_back("hip");\
#endif' \
  -e '/^std::unique_ptr<SearchBackend> make_backend(const std::string& name) {/a\
#ifdef FAINTTRAIL_HAVE_HIP\
    if (name == "hip" || name.rfind("hip:", 0) == 0) return detail::make_hip_backend(name.size() > 4 ? name.substr(4) : std::string());\
    if ((name.empty() || name == "auto") && detail::hip_backend_available(nullptr)) return detail::make_hip_backend(std::string());\
#endif' \
  "$REF/src/backend.cpp" > "$OUT/backend.cpp"
sed \
  -e 's/if (cand->name() == "cuda") {/if (cand->name() == "cuda" || cand->name() == "hip") {/' \
  -e 's/make_backend("cuda:max_mb=2,launch_adds=2e6")/make_backend(cand->name() + ":max_mb=2,launch_adds=2e6")/' \
  "$REF/cli/main.cpp" > "$OUT/main.cpp"
# every edit must have applied
for pat in 'names.push_back("hip")' 'detail::make_hip_backend(std::string())' 'namespace detail {'; do
  grep -q "$pat" "$OUT/backend.cpp" || { echo "hip_integration.sh: backend.cpp edit did not apply ($pat)" >&2; exit 1; }
done
grep -q 'cand->name() == "hip"' "$OUT/main.cpp" || { echo "hip_integration.sh: main.cpp edit did not apply" >&2; exit 1; }
grep -q 'make_backend(cand->name() + ":max_mb' "$OUT/main.cpp" || { echo "hip_integration.sh: main.cpp edit did not apply" >&2; exit 1; }
