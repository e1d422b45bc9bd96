// Device-copy ceilings for the population layouts (acp_copy_probe).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "acp.h"
#include "engine.hpp"

namespace acp {

// The copies use the bulk kernels' access rule: the non-temporal hint when the
// two arrays exceed the last-level cache (NT, plane.hpp).

// one element of each of P planes per thread, loads first, then stores
template <int P, bool NT>
__global__ __launch_bounds__(512) void copy_planes(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 512 + threadIdx.x;
    if (i >= n) return;
    double v[P];
#pragma unroll
    for (int s = 0; s < P; ++s) v[s] = ld_pop<NT>(a + s * n + i);
#pragma unroll
    for (int s = 0; s < P; ++s) st_pop<NT>(b + s * n + i, v[s]);
}

// 19 populations of one node per thread in the block layout (plane.hpp):
// loads first, then stores
template <bool NT>
__global__ __launch_bounds__(512) void copy_blocks(const double* __restrict__ a, double* __restrict__ b,
                                                   const PopLayout L) {
    const int64_t i = (int64_t)blockIdx.x * 512 + threadIdx.x;
    if (i >= L.N) return;
    double v[19];
#pragma unroll
    for (int s = 0; s < 19; ++s) v[s] = ld_pop<NT>(a + pidx(L, s, i));
#pragma unroll
    for (int s = 0; s < 19; ++s) st_pop<NT>(b + pidx(L, s, i), v[s]);
}

__global__ __launch_bounds__(512) void fill_ramp(double* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 512 + threadIdx.x;
    if (i < n) a[i] = (double)(i & 1023);
}

}  // namespace acp

extern "C" int acp_copy_probe(int32_t planes, int64_t plane_elems, int64_t block_nodes, int32_t reps, void* stream,
                              double* ms_per_copy) {
    acp::clear_error();
    if ((planes != 1 && planes != 19) || plane_elems <= 0 || reps <= 0 || !ms_per_copy || block_nodes < 0 ||
        (block_nodes & (block_nodes - 1)) != 0 || (block_nodes > 0 && planes != 19))
        return acp::fail(ACP_E_ARG, "acp_copy_probe: planes 1 or 19, sizes positive, blocks a power of two (19 planes)");
    const acp::PopLayout L = acp::make_layout(plane_elems, block_nodes);
    const int64_t total = planes == 19 ? L.elems : plane_elems;
    const unsigned blocks = (unsigned)((plane_elems + 511) / 512);
    if ((plane_elems + 511) / 512 > 0x7fffffff) return acp::fail(ACP_E_ARG, "acp_copy_probe: plane too large");
    hipStream_t s = (hipStream_t)stream;
    double *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ACP_OK;
    hipError_t e = hipMalloc((void**)&a, sizeof(double) * total);
    if (e == hipSuccess) e = hipMalloc((void**)&b, sizeof(double) * total);
    if (e == hipErrorOutOfMemory) rc = acp::fail(ACP_E_OOM, "acp_copy_probe: out of device memory");
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(acp::fill_ramp, dim3((unsigned)((total + 511) / 512)), dim3(512), 0, s, a, total);
        e = hipGetLastError();
    }
    const bool nt = 2 * total * (int64_t)sizeof(double) > acp::kCacheBytes;
    auto launch = [&]() {
        if (planes == 1) {
            if (nt) hipLaunchKernelGGL((acp::copy_planes<1, true>), dim3(blocks), dim3(512), 0, s, a, b, plane_elems);
            else hipLaunchKernelGGL((acp::copy_planes<1, false>), dim3(blocks), dim3(512), 0, s, a, b, plane_elems);
        } else if (L.block) {
            if (nt) hipLaunchKernelGGL(acp::copy_blocks<true>, dim3(blocks), dim3(512), 0, s, a, b, L);
            else hipLaunchKernelGGL(acp::copy_blocks<false>, dim3(blocks), dim3(512), 0, s, a, b, L);
        } else {
            if (nt) hipLaunchKernelGGL((acp::copy_planes<19, true>), dim3(blocks), dim3(512), 0, s, a, b, plane_elems);
            else hipLaunchKernelGGL((acp::copy_planes<19, false>), dim3(blocks), dim3(512), 0, s, a, b, plane_elems);
        }
        return hipGetLastError();
    };
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) *ms_per_copy = (double)ms / reps;
    else if (rc == ACP_OK) rc = acp::fail(ACP_E_HIP, hipGetErrorString(e));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
