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
// positive), one workgroup per block: the block sits in LDS, and each of the
// n steps stages the pivot row and column, then every thread updates its
// elements (column = lane, rows wave, wave + 4, ...: 16 of a 64 x 64 block,
// no index division).  status[b] = 1 if a pivot is not positive and finite
// (the block is then left as loaded).
__global__ __launch_bounds__(256) void spd_inverse_lds_kernel(double* __restrict__ A,
                                                              const int64_t* __restrict__ off,
                                                              const int32_t* __restrict__ sizes,
                                                              int32_t* __restrict__ status) {
    constexpr int W = 64;
    __shared__ double S[W][W + 1];
    __shared__ double rowk[W], colk[W];
    __shared__ int bad;
    const int b = blockIdx.x;
    const int n = sizes[b];
    double* M = A + off[b];
    const int tid = threadIdx.x;
    for (int e = tid; e < n * n; e += 256) S[e / n][e % n] = M[e];
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double d = S[k][k];
        if (!(d > 0.0) || !__builtin_isfinite(d)) {
            if (tid == 0) bad = 1;
            break;  // uniform: every thread read the same d
        }
        const double inv = 1.0 / d;
        // the pivot row scaled (its new values) and the pivot column (old values)
        if (tid < n) rowk[tid] = tid == k ? inv : S[k][tid] * inv;
        else if (tid >= 128 && tid - 128 < n) colk[tid - 128] = S[tid - 128][k];
        __syncthreads();
        const int j = tid & (W - 1);
        if (j < n) {
            const double rk = rowk[j];
            for (int i = tid / W; i < n; i += 256 / W) {
                const double f = colk[i];
                S[i][j] = i == k ? rk : (j == k ? -f * inv : S[i][j] - f * rk);
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (!bad)
        for (int e = tid; e < n * n; e += 256) M[e] = S[e / n][e % n];
    if (tid == 0) status[b] = bad;
}

// ---- blocked Gauss-Jordan on 64 x 64 tiles ------------------------------------
constexpr int T = kDenseTile;

// C (64 x 64 at c, ld ldc) = alpha * Aa * Bb (+ C if accumulate), the tiles
// staged in LDS; 256 threads, 4 x 4 outputs each.  Aa, Bb may alias C: both
// are read completely before C is written.
__device__ void tile_mm(const double* __restrict__ a, int64_t lda, const double* __restrict__ b, int64_t ldb,
                        double* c, int64_t ldc, double alpha, bool accumulate) {
    __shared__ double As[T][T + 1];
    __shared__ double Bs[T][T + 1];
    for (int e = threadIdx.x; e < T * T; e += 256) {
        const int r = e / T, q = e % T;
        As[r][q] = a[r * lda + q];
        Bs[r][q] = b[r * ldb + q];
    }
    __syncthreads();
    const int r0 = (threadIdx.x / 16) * 4, c0 = (threadIdx.x % 16) * 4;
    double acc[4][4] = {};
    for (int k = 0; k < T; ++k) {
        double av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) av[i] = As[r0 + i][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = Bs[k][c0 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double* dst = c + (int64_t)(r0 + i) * ldc + c0 + j;
            *dst = accumulate ? *dst + alpha * acc[i][j] : alpha * acc[i][j];
        }
}

__global__ __launch_bounds__(256) void gj_copy_pivot_kernel(const double* __restrict__ A, int n_pad, int k,
                                                           double* __restrict__ P) {
    for (int e = threadIdx.x; e < T * T; e += 256) {
        const int r = e / T, q = e % T;
        P[e] = A[(int64_t)(k * T + r) * n_pad + k * T + q];
    }
}

// Row panel: A_kj := P A_kj for every column tile j != k.
__global__ __launch_bounds__(256) void gj_row_panel_kernel(double* A, int n_pad, int k, const double* P) {
    const int j = blockIdx.x < (unsigned)k ? blockIdx.x : blockIdx.x + 1;
    double* tile = A + (int64_t)k * T * n_pad + (int64_t)j * T;
    tile_mm(P, T, tile, n_pad, tile, n_pad, 1.0, false);
}

// Trailing update: A_ij -= A_ik A_kj (A_kj already P A_kj) for i, j != k.
__global__ __launch_bounds__(256) void gj_update_kernel(double* A, int n_pad, int k) {
    const int nt = n_pad / T - 1;
    const int bi = blockIdx.x / nt, bj = blockIdx.x % nt;
    const int i = bi < k ? bi : bi + 1, j = bj < k ? bj : bj + 1;
    tile_mm(A + (int64_t)i * T * n_pad + (int64_t)k * T, n_pad, A + (int64_t)k * T * n_pad + (int64_t)j * T, n_pad,
            A + (int64_t)i * T * n_pad + (int64_t)j * T, n_pad, -1.0, true);
}

// Column panel: A_ik := -A_ik P for i != k, and A_kk := P.
__global__ __launch_bounds__(256) void gj_col_panel_kernel(double* A, int n_pad, int k, const double* P) {
    const int nt = n_pad / T;
    if ((int)blockIdx.x == nt - 1) {
        for (int e = threadIdx.x; e < T * T; e += 256)
            A[(int64_t)(k * T + e / T) * n_pad + k * T + e % T] = P[e];
        return;
    }
    const int i = blockIdx.x < (unsigned)k ? blockIdx.x : blockIdx.x + 1;
    double* tile = A + (int64_t)i * T * n_pad + (int64_t)k * T;
    tile_mm(tile, n_pad, P, T, tile, n_pad, -1.0, false);
}

__global__ __launch_bounds__(256) void gj_status_kernel(const int32_t* __restrict__ st, int32_t* __restrict__ out, int k) {
    if (threadIdx.x == 0) out[k] = st[0];
}

// y = A x for the symmetric n_pad x n_pad A (column major) from its lower
// triangle, which is read once: workgroup b takes tile (I, J), I >= J, of
// 64 x 64 (b = I (I + 1) / 2 + J) and writes its row sums A_IJ x_J to
// partial[(I nt + J) 64 + .] and, off the diagonal, its column sums
// A_IJ^T x_I to partial[(J nt + I) 64 + .], so that every (R, K) slot of the
// nt x nt x 64 partials is written once.  Wave g holds rows I*64 + lane and
// columns J*64 + 16 g .. + 16; the column sums go down the wave by recursive
// halving (each lane keeps half the columns per exchange), 17 shuffles for
// 16 columns.
__global__ __launch_bounds__(256) void symv_tiles_kernel(const double* __restrict__ A, int n_pad,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ partial) {
    __shared__ double rows[4][64];
    const int nt = n_pad / 64;
    const int b = (int)blockIdx.x;
    int I = (int)((__builtin_sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= b) ++I;
    while (I * (I + 1) / 2 > b) --I;
    const int J = b - I * (I + 1) / 2;
    const int lane = (int)threadIdx.x & 63, g = (int)threadIdx.x >> 6;
    const int r = I * 64 + lane;
    const int c0 = J * 64 + g * 16;
    const double* col = A + (int64_t)c0 * n_pad + r;
    double v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = col[(int64_t)k * n_pad];
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        s0 += v[k] * x[c0 + k];
        s1 += v[k + 1] * x[c0 + k + 1];
    }
    rows[g][lane] = s0 + s1;
    if (I != J) {  // uniform over the workgroup
        const double xr = x[r];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] *= xr;
        // halving: bit 5 picks columns 8-15, bit 4 the upper four, bit 3 two, bit 2 one
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool hi = lane & 32;
            const double send = hi ? v[k] : v[k + 8];
            v[k] = (hi ? v[k + 8] : v[k]) + __shfl_xor(send, 32, 64);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool hi = lane & 16;
            const double send = hi ? v[k] : v[k + 4];
            v[k] = (hi ? v[k + 4] : v[k]) + __shfl_xor(send, 16, 64);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool hi = lane & 8;
            const double send = hi ? v[k] : v[k + 2];
            v[k] = (hi ? v[k + 2] : v[k]) + __shfl_xor(send, 8, 64);
        }
        {
            const bool hi = lane & 4;
            const double send = hi ? v[0] : v[1];
            v[0] = (hi ? v[1] : v[0]) + __shfl_xor(send, 4, 64);
        }
        v[0] += __shfl_xor(v[0], 2, 64);
        v[0] += __shfl_xor(v[0], 1, 64);
        if ((lane & 3) == 0) {
            const int c = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
            partial[((int64_t)J * nt + I) * 64 + g * 16 + c] = v[0];
        }
    }
    __syncthreads();
    if (g == 0) partial[((int64_t)I * nt + J) * 64 + lane] = (rows[0][lane] + rows[1][lane]) + (rows[2][lane] + rows[3][lane]);
}

// y[R 64 + l] = sum over K of partial[(R nt + K) 64 + l]: workgroup R, wave
// w summing K in [w q, (w + 1) q) (q = ceil(nt / 4)), the four quarters added
// in order through LDS.
__global__ __launch_bounds__(256) void symv_tiles_sum_kernel(const double* __restrict__ partial, int n_pad,
                                                            double* __restrict__ y) {
    __shared__ double quarter[4][64];
    const int nt = n_pad / 64, R = (int)blockIdx.x;
    const int l = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int q = (nt + 3) / 4;
    const int k0 = w * q, k1 = k0 + q < nt ? k0 + q : nt;
    const double* p = partial + (int64_t)R * nt * 64 + l;
    double a0 = 0.0, a1 = 0.0;
    int K = k0;
    for (; K + 2 <= k1; K += 2) {
        a0 += p[(int64_t)K * 64];
        a1 += p[(int64_t)(K + 1) * 64];
    }
    if (K < k1) a0 += p[(int64_t)K * 64];
    quarter[w][l] = a0 + a1;
    __syncthreads();
    if (w == 0) y[R * 64 + l] = (quarter[0][l] + quarter[1][l]) + (quarter[2][l] + quarter[3][l]);
}

}  // namespace

hipError_t k_batched_inverse_spd(double* A, const int64_t* off, const int32_t* sizes, int nblocks, int max_size,
                                 int32_t* status, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (max_size > 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spd_inverse_lds_kernel, dim3((unsigned)nblocks), dim3(256), 0, stream, A, off, sizes, status);
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

hipError_t k_spd_inverse_blocked(double* A, int n_pad, const DenseWork& w, hipStream_t stream) {
    const int nt = n_pad / T;
    for (int k = 0; k < nt; ++k) {
        hipLaunchKernelGGL(gj_copy_pivot_kernel, dim3(1), dim3(256), 0, stream, A, n_pad, k, w.pivot);
        hipError_t e = k_batched_inverse(w.pivot, w.off, w.size, 1, T, w.piv, w.piv_off, w.status + nt, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(gj_status_kernel, dim3(1), dim3(64), 0, stream, w.status + nt, w.status, k);
        if (nt > 1) {
            hipLaunchKernelGGL(gj_row_panel_kernel, dim3((unsigned)(nt - 1)), dim3(256), 0, stream, A, n_pad, k,
                               (const double*)w.pivot);
            hipLaunchKernelGGL(gj_update_kernel, dim3((unsigned)((nt - 1) * (nt - 1))), dim3(256), 0, stream, A,
                               n_pad, k);
        }
        hipLaunchKernelGGL(gj_col_panel_kernel, dim3((unsigned)nt), dim3(256), 0, stream, A, n_pad, k,
                           (const double*)w.pivot);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t k_symv_dense(const double* A, int n_pad, const double* x, double* y, double* partial, int chunks,
                        hipStream_t stream) {
    const int nt = n_pad / 64;
    if (n_pad % 64 || chunks != nt) return hipErrorInvalidValue;
    hipLaunchKernelGGL(symv_tiles_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, stream, A, n_pad, x,
                       partial);
    hipLaunchKernelGGL(symv_tiles_sum_kernel, dim3((unsigned)nt), dim3(256), 0, stream, (const double*)partial,
                       n_pad, y);
    return hipGetLastError();
}

}  // namespace tfs
