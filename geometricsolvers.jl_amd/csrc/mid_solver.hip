// libgsolve_hip, mid-size systems: one wave per system, each lane's factor row in VGPRs — the
// hand-over flags and the launch of the plain forms (the device code is mid_kernel.hpp).
#include <mutex>

#include "mid_kernel.hpp"

namespace gs {

namespace {

// The per-system hand-over flags of the two kernels, one byte per system, kept per device until
// gs_shutdown (the entry point synchronises before it returns, so no launch is in flight when the
// buffer grows).
struct RedoFlags {
    uint8_t* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_redo_mutex;
RedoFlags g_redo[64];

int redo_flags(size_t bytes, uint8_t** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(GS_EHIP, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_redo_mutex);
    RedoFlags& w = g_redo[dev];
    if (w.bytes < bytes) {
        if (w.ptr) (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
        if (hipMalloc(reinterpret_cast<void**>(&w.ptr), bytes) != hipSuccess)
            return set_error(GS_EHIP, "hipMalloc of the wave-per-system flags failed");
        w.bytes = bytes;
    }
    *out = w.ptr;
    return GS_OK;
}

}  // namespace

// Free the flags of every device (gs_shutdown; no call may be in flight).
void release_mid() {
    std::lock_guard<std::mutex> lock(g_redo_mutex);
    for (int d = 0; d < 64; ++d) {
        if (!g_redo[d].ptr) continue;
        if (hipSetDevice(d) == hipSuccess) (void)hipFree(g_redo[d].ptr);
        g_redo[d] = RedoFlags{};
    }
}

int launch_fpu_gl3(const MidLaunch& L0, hipStream_t st) {
    if (L0.nsys > 0x7fffffffLL) return set_error(GS_EINVAL, "a wave-per-system batch holds at most 2^31-1 systems");
    MidLaunch L = L0;
    if (!L.redo) {
        const int rc = redo_flags(size_t(L.nsys), &L.redo);
        if (rc != GS_OK) return rc;
    }
    return L.reuse ? launch_fpu_gl3_reuse(L, st) : launch_forms<false>(L, st);
}

}  // namespace gs
