// acp_case_get_macros_noalloc: the density and velocity read-back of
// acp_case_get_macros without any device allocation.
//
// acp_case_get_macros stages the macroscopic fields in a device buffer it
// allocates and frees on every call; hipFree waits on every stream of the
// device.  This variant stages them in `scratch`, a file-scope device array
// of kScratchNodes nodes (rho and three velocity components in float64 per
// node: 128 MiB), which every GPU the library runs on holds from the first
// kernel launch of the library until the process ends.  The array's address
// on a device is taken once, with that device current, through
// hipGetSymbolAddress and kept per device; host code never passes the array's
// own name to a kernel or a copy.
//
// The kernels write every element the call later reads (rho[l] and u[.][l]
// for every node l < N), so nothing needs clearing between calls.  All calls
// on one GPU share its array: calls that overlap in time on one GPU (from
// several host threads, or on several streams) race and must be serialised
// by the caller.
#include <mutex>
#include <string>

#include "engine.hpp"

namespace {

constexpr int64_t kScratchNodes = (int64_t)1 << 22;
__device__ __align__(256) double scratch[4 * kScratchNodes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (current when called).
int scratch_on(int dev, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return acp::hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ACP_OK;
}

// Makes `dev` current for the lifetime of the object and restores the
// caller's device afterwards.
struct DeviceScope {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceScope(int dev) {
        err = hipGetDevice(&prev);
        if (err == hipSuccess && prev != dev) err = hipSetDevice(dev);
    }
    ~DeviceScope() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" int acp_case_get_macros_noalloc(acp_case* c, double* rho, double* u) {
    if (!c) return acp::fail(ACP_E_ARG, "null case handle");
    if (c->pack) return acp::fail(ACP_E_STATE, "the case is stepped by a pack: acp_pack_remove it first");
    if (!rho || !u) return acp::fail(ACP_E_ARG, "null output array");
    const int64_t N = c->g.N;
    if (N > kScratchNodes)
        return acp::fail(ACP_E_ARG, "acp_case_get_macros_noalloc: " + std::to_string(N) +
                                        " nodes exceed its scratch buffer of " + std::to_string(kScratchNodes) +
                                        " nodes; use acp_case_get_macros");
    hipDevice_t dev = 0;
    hipError_t e = hipStreamGetDevice(c->stream, &dev);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_case_get_macros_noalloc: stream device");
    if (dev < 0 || dev >= kMaxDevices) return acp::fail(ACP_E_ARG, "acp_case_get_macros_noalloc: device index");
    DeviceScope scope(dev);
    if (scope.err != hipSuccess) return acp::hip_fail(scope.err, "acp_case_get_macros_noalloc: set device");
    double* buf = nullptr;
    if (int rc = scratch_on(dev, &buf)) return rc;
    e = c->precision == ACP_PREC_F64
            ? acp::launch_macros_out64(c->dk, c->A, c->g, buf, buf + N, c->stream)
            : acp::launch_macros_out_shift(c->precision, c->dks, c->SA, c->g, buf, buf + N, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rho, buf, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(u, buf + N, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_case_get_macros_noalloc");
    return ACP_OK;
}
