// tfs_real32.hip -- FP32 host arrays at the C-ABI.
//
// The reference's plugin surface hands over arrays of its build's Real, which
// is float as shipped (config.hpp:36,52; mobility_solver.hpp:57-63).  The
// *_f32 entry points take such arrays: the floats cross PCIe (half the bytes
// of FP64) into a device staging buffer and one kernel widens them into the
// handle's FP64 arrays; results are narrowed on the device and copied back as
// floats.  Widening is exact, so an FP32 caller's inputs reach the kernels
// unchanged; outputs are FP64 results rounded once to FP32.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tfs_api_internal.h"

namespace tfs {
namespace {

__global__ __launch_bounds__(256) void widen_kernel(const float* __restrict__ src, double* __restrict__ dst,
                                                    int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = (double)src[i];
}

__global__ __launch_bounds__(256) void narrow_kernel(const double* __restrict__ src, float* __restrict__ dst,
                                                     int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = (float)src[i];
}

unsigned grid_of(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

int ensure_stage(F32Stage& st, int64_t n, hipStream_t s, const char* where) {
    if ((size_t)n <= st.cap) return TFS_OK;
    if (st.p) {
        // the old buffer may still feed queued copies
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, where);
        (void)hipFree(st.p);
    }
    st.p = nullptr;
    st.cap = 0;
    if (hipMalloc(&st.p, sizeof(float) * (size_t)n) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, std::string(where) + ": FP32 staging buffer");
    st.cap = (size_t)n;
    return TFS_OK;
}

}  // namespace

void free_stage(F32Stage& st) {
    if (st.p) (void)hipFree(st.p);
    st = F32Stage{};
}

int widen_in(F32Stage& st, const float* src, double* dst, int64_t n, bool host, hipStream_t s, const char* where) {
    if (n <= 0) return TFS_OK;
    const float* dsrc = src;
    if (host) {
        int rc = ensure_stage(st, n, s, where);
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(st.p, src, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, where);
        dsrc = st.p;
    }
    hipLaunchKernelGGL(widen_kernel, dim3(grid_of(n)), dim3(256), 0, s, dsrc, dst, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? TFS_OK : hip_fail(e, where);
}

int narrow_out(F32Stage& st, const double* src, float* dst, int64_t n, bool host, hipStream_t s, const char* where) {
    if (n <= 0) return TFS_OK;
    float* ddst = dst;
    if (host) {
        int rc = ensure_stage(st, n, s, where);
        if (rc) return rc;
        ddst = st.p;
    }
    hipLaunchKernelGGL(narrow_kernel, dim3(grid_of(n)), dim3(256), 0, s, src, ddst, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && host)
        e = hipMemcpyAsync(dst, st.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? TFS_OK : hip_fail(e, where);
}

}  // namespace tfs
