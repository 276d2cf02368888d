// dense_kernels.hip -- batched Gauss-Jordan inverse of small dense blocks.
#include "dense_kernels.h"
#include "wave_ops.h"

namespace tfs {
namespace {

// In-place Gauss-Jordan inverse of every row-major block, partial pivoting,
// one workgroup per block.  Step k: pivot search over rows k..bs-1, row swap,
// scale of the pivot row (its pivot entry becomes 1/d), elimination of column k
// from every other row (column k becomes -A[i][k]/d); the row swaps are undone
// as column swaps in reverse order at the end.  status[b] = 1 if a pivot is zero
// or not finite.
__global__ __launch_bounds__(256) void block_gauss_jordan_kernel(double* __restrict__ A,
                                                                 const int64_t* __restrict__ mat_off,
                                                                 const int32_t* __restrict__ sizes,
                                                                 int32_t* __restrict__ piv_all,
                                                                 const int32_t* __restrict__ piv_off,
                                                                 int32_t* __restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double red_v[256];
    __shared__ int red_i[256];
    __shared__ int bad;
    const int b = blockIdx.x;
    const int bs = sizes[b];
    double* M = A + mat_off[b];
    int32_t* piv = piv_all + piv_off[b];
    double* colk = lds;        // column k of the other rows
    double* rowk = lds + bs;   // the scaled pivot row
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < bs; ++k) {
        // pivot search
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < bs; i += blockDim.x) {
            const double v = fabs(M[(int64_t)i * bs + k]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (tid < w) {
                const double o = red_v[tid + w];
                const int oi = red_i[tid + w];
                if (o > red_v[tid] || (o == red_v[tid] && oi < red_i[tid])) {
                    red_v[tid] = o;
                    red_i[tid] = oi;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        // row swap k <-> p, then the pivot row is scaled into LDS
        if (p != k)
            for (int j = tid; j < bs; j += blockDim.x) {
                const double t = M[(int64_t)k * bs + j];
                M[(int64_t)k * bs + j] = M[(int64_t)p * bs + j];
                M[(int64_t)p * bs + j] = t;
            }
        if (tid == 0) piv[k] = p;
        __syncthreads();
        const double d = M[(int64_t)k * bs + k];
        if (!(fabs(d) > 0.0) || !isfinite(d)) {
            if (tid == 0) bad = 1;
            break;   // uniform across the workgroup: every thread read the same d
        }
        const double inv = 1.0 / d;
        for (int j = tid; j < bs; j += blockDim.x) {
            const double v = j == k ? inv : M[(int64_t)k * bs + j] * inv;
            rowk[j] = v;
            M[(int64_t)k * bs + j] = v;
        }
        for (int i = tid; i < bs; i += blockDim.x) colk[i] = i == k ? 0.0 : M[(int64_t)i * bs + k];
        __syncthreads();
        for (int64_t e = tid; e < (int64_t)bs * bs; e += blockDim.x) {
            const int i = (int)(e / bs), j = (int)(e % bs);
            if (i == k) continue;
            const double f = colk[i];
            M[e] = j == k ? -f * rowk[k] : M[e] - f * rowk[j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) {
        if (tid == 0) status[b] = 1;
        return;
    }
    for (int k = bs - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < bs; i += blockDim.x) {
                const double t = M[(int64_t)i * bs + k];
                M[(int64_t)i * bs + k] = M[(int64_t)i * bs + p];
                M[(int64_t)i * bs + p] = t;
            }
        __syncthreads();
    }
    if (tid == 0) status[b] = 0;
}

// In-place inverse of symmetric positive definite row-major blocks of at most
// 64 rows by Gauss-Jordan elimination without pivoting (SPD: every pivot is
// positive), one wave per block, four blocks per workgroup.  Lane j keeps
// column j in registers; the pivot row and column come from the owning lanes
// by readlane, so no LDS and no barrier is needed.  status[b] = 1 if a pivot
// is not positive and finite (the block is then left partly eliminated).
__global__ __launch_bounds__(256) void spd_inverse_wave_kernel(double* __restrict__ A,
                                                               const int64_t* __restrict__ off,
                                                               const int32_t* __restrict__ sizes, int nblocks,
                                                               int32_t* __restrict__ status) {
    constexpr int W = 64;
    const int b = blockIdx.x * 4 + (int)(threadIdx.x / W);
    const int lane = threadIdx.x & (W - 1);
    if (b >= nblocks) return;
    const int n = sizes[b];
    double* M = A + off[b];
    const bool mine = lane < n;
    double col[W];
#pragma unroll
    for (int i = 0; i < W; ++i) col[i] = (i < n && mine) ? M[(int64_t)i * n + lane] : 0.0;
    int bad = 0;
    for (int k = 0; k < n; ++k) {
        double akj = 0.0;
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (i == k) akj = col[i];
        const double d = lane_bcast(akj, k);
        if (!(d > 0.0) || !__builtin_isfinite(d)) {
            bad = 1;
            break;  // uniform: every lane read the same d
        }
        const double inv = 1.0 / d;
        const double rk = lane == k ? inv : akj * inv;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            if (i < n) {
                const double f = lane_bcast(col[i], k);
                col[i] = i == k ? rk : (lane == k ? -f * inv : col[i] - f * rk);
            }
        }
    }
    if (!bad && mine) {
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (i < n) M[(int64_t)i * n + lane] = col[i];
    }
    if (lane == 0) status[b] = bad;
}

}  // namespace

hipError_t k_batched_inverse_spd(double* A, const int64_t* off, const int32_t* sizes, int nblocks, int max_size,
                                 int32_t* status, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (max_size > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spd_inverse_wave_kernel, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, stream, A, off,
                       sizes, nblocks, status);
    return hipGetLastError();
}

hipError_t k_batched_inverse(double* A, const int64_t* off, const int32_t* sizes, int nblocks, int max_size,
                             int32_t* piv, const int32_t* piv_off, int32_t* status, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    const size_t lds = sizeof(double) * 2 * (size_t)max_size;
    hipLaunchKernelGGL(block_gauss_jordan_kernel, dim3((unsigned)nblocks), dim3(256), lds, stream, A, off, sizes, piv,
                       piv_off, status);
    return hipGetLastError();
}

}  // namespace tfs
