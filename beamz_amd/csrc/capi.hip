// C-ABI plumbing: per-thread error strings, version and device query.
#include "common.h"

namespace bz {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
int hip_fail(const char* where, hipError_t e) {
    g_error = std::string(where) + ": " + hipGetErrorString(e);
    return BZ_EHIP;
}
const char* last_error() { return g_error.c_str(); }
}  // namespace bz

extern "C" const char* bz_last_error(void) { return bz::last_error(); }

extern "C" int bz_version(void) { return 100; }

extern "C" int bz_abi_sizes(int* out, int n) {
    BZ_REQUIRE(out && n >= 0, "bz_abi_sizes: null pointer");
    const int sizes[3] = {(int)sizeof(bz_sim2d), (int)sizeof(bz_adj2d), (int)sizeof(bz_box_source)};
    for (int q = 0; q < n && q < 3; ++q) out[q] = sizes[q];
    return BZ_OK;
}

extern "C" int bz_device_count(int* count) {
    BZ_REQUIRE(count, "bz_device_count: null pointer");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return bz::hip_fail("bz_device_count", e);
    }
    return BZ_OK;
}
