// tfs_broyden_api.hip -- device-resident Broyden updates behind include/tfs_broyden.h.
//
// The reference keeps C and D as host `matrix` columns and forms every dot
// product and column update in host loops (broyden_solver.cpp:26-129).  Here
// C and D live in HBM (column-major, FP64); the sums over previous iterations
// are one transposed GEMV (error . D_i for all i) and one GEMV (sum h_i C_i),
// shared with the GMRES kernels, and J0^{-1} is applied as block-diagonal dense
// blocks, one workgroup per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/tfs_broyden.h"
#include "solver_kernels.h"
#include "dense_kernels.h"
#include "tfs_api_internal.h"

using namespace tfs;

namespace {

// y[rows of block b] = alpha * A_b x[rows of block b]; A_b is stored column-major
// (the transpose of the caller's row-major block) so that the threads of a
// workgroup, one per row, read consecutive addresses.
__global__ __launch_bounds__(128) void block_gemv_kernel(const double* __restrict__ A,
                                                         const int64_t* __restrict__ mat_off,
                                                         const int32_t* __restrict__ row_off,
                                                         const int32_t* __restrict__ sizes,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         double alpha) {
    const int b = blockIdx.x;
    const int bs = sizes[b];
    const double* Ab = A + mat_off[b];
    const int r0 = row_off[b];
    for (int r = threadIdx.x; r < bs; r += blockDim.x) {
        // four independent chains keep more loads in flight
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int c = 0;
        for (; c + 4 <= bs; c += 4) {
            s0 += Ab[(int64_t)c * bs + r] * x[r0 + c];
            s1 += Ab[(int64_t)(c + 1) * bs + r] * x[r0 + c + 1];
            s2 += Ab[(int64_t)(c + 2) * bs + r] * x[r0 + c + 2];
            s3 += Ab[(int64_t)(c + 3) * bs + r] * x[r0 + c + 3];
        }
        for (; c < bs; ++c) s0 += Ab[(int64_t)c * bs + r] * x[r0 + c];
        y[r0 + r] = alpha * ((s0 + s1) + (s2 + s3));
    }
}

// dst_b = src_b^T for every block.
__global__ __launch_bounds__(256) void block_transpose_kernel(const double* __restrict__ src,
                                                              double* __restrict__ dst,
                                                              const int64_t* __restrict__ mat_off,
                                                              const int32_t* __restrict__ sizes) {
    const int b = blockIdx.x;
    const int bs = sizes[b];
    const double* S = src + mat_off[b];
    double* T = dst + mat_off[b];
    for (int64_t e = threadIdx.x; e < (int64_t)bs * bs; e += blockDim.x) {
        const int r = (int)(e / bs), c = (int)(e % bs);
        T[(int64_t)c * bs + r] = S[e];
    }
}

// Free bodies, once per J0: per filament W = J_i^{-1} A_i (6N x 6) and
// BJ = B_i J_i^{-1} (6 x 6N), J_i^{-1} stored column-major at mat_off.  One
// workgroup per filament.
__global__ __launch_bounds__(256) void coupling_prep_kernel(const double* __restrict__ J,
                                                            const int64_t* __restrict__ mat_off, int fil_per_swim,
                                                            int blocks_per_swim, int n6,
                                                            const double* __restrict__ A,
                                                            const double* __restrict__ B, double* __restrict__ W,
                                                            double* __restrict__ BJ) {
    const int f = blockIdx.x;
    const int s = f / fil_per_swim, i = f % fil_per_swim;
    const double* Ji = J + mat_off[(int64_t)s * blocks_per_swim + i];
    const int64_t o = (int64_t)f * 6 * n6;
    for (int r = threadIdx.x; r < n6; r += blockDim.x) {
        double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < n6; ++c) {
            const double jv = Ji[(int64_t)c * n6 + r];
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] += jv * A[o + (int64_t)c * 6 + k];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) W[o + (int64_t)r * 6 + k] = w[k];
    }
    for (int c = threadIdx.x; c < n6; c += blockDim.x) {
        double bj[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < n6; ++r) {
            const double jv = Ji[(int64_t)c * n6 + r];
#pragma unroll
            for (int k = 0; k < 6; ++k) bj[k] += B[o + (int64_t)k * n6 + r] * jv;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) BJ[o + (int64_t)k * n6 + c] = bj[k];
    }
}

// S = S0 - sum_i BJ_i A_i, one workgroup per swimmer, thread (k, j) < 36.
__global__ __launch_bounds__(64) void coupling_schur_kernel(const double* __restrict__ S0,
                                                            const double* __restrict__ BJ,
                                                            const double* __restrict__ A, int fil_per_swim, int n6,
                                                            double* __restrict__ S) {
    const int s = blockIdx.x, t = threadIdx.x;
    if (t >= 36) return;
    const int k = t / 6, j = t % 6;
    double acc = S0[36 * s + t];
    for (int i = 0; i < fil_per_swim; ++i) {
        const int64_t o = ((int64_t)s * fil_per_swim + i) * 6 * n6;
        for (int c = 0; c < n6; ++c) acc -= BJ[o + (int64_t)k * n6 + c] * A[o + (int64_t)c * 6 + j];
    }
    S[36 * s + t] = acc;
}

// After the block-diagonal product y = alpha diag(J_i^{-1}, S^{-1}, fac I ...)
// x: the body rows become alpha S^{-1} (x_b - sum_i BJ_i x_i - fac sum_m (x_m,
// d_m x x_m)) and the filament rows y_i - W_i y_b (swimmer.cpp:1012-1066); with
// blob blocks (the Broyden-for-everything form, swimmer.cpp:931-1010) the blob
// rows y_m - fac dt_fac (y_b + y_b,rot x d_m).  One workgroup per swimmer.
struct BlobCoupling {
    int nblob = 0;               // blobs per swimmer (0: none)
    const double* disp = nullptr;  // 3 per blob: Q_init ref
    double fac = 0.0, dt_fac = 0.0;
};
__global__ __launch_bounds__(256) void coupling_apply_kernel(const double* __restrict__ W,
                                                             const double* __restrict__ BJ,
                                                             const double* __restrict__ Sinv, int fil_per_swim,
                                                             int n6, BlobCoupling bc, const double* __restrict__ x,
                                                             double* __restrict__ y, double alpha) {
    const int s = blockIdx.x;
    const int64_t rows = (int64_t)fil_per_swim * n6;
    const int64_t r0 = (int64_t)s * (rows + 6 + 3LL * bc.nblob);
    const int64_t o = (int64_t)s * fil_per_swim * 6 * n6;
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t e = threadIdx.x; e < rows; e += blockDim.x) {
        const int64_t i = e / n6, c = e % n6;
        const double xv = x[r0 + e];
#pragma unroll
        for (int k = 0; k < 6; ++k) g[k] += BJ[o + i * 6 * n6 + (int64_t)k * n6 + c] * xv;
    }
    for (int64_t m = threadIdx.x; m < bc.nblob; m += blockDim.x) {
        const double* xb = x + r0 + rows + 6 + 3 * m;
        const double* d = bc.disp + 3 * ((int64_t)s * bc.nblob + m);
        g[0] += bc.fac * xb[0];
        g[1] += bc.fac * xb[1];
        g[2] += bc.fac * xb[2];
        g[3] += bc.fac * (d[1] * xb[2] - d[2] * xb[1]);
        g[4] += bc.fac * (d[2] * xb[0] - d[0] * xb[2]);
        g[5] += bc.fac * (d[0] * xb[1] - d[1] * xb[0]);
    }
    __shared__ double sm[6][256];
    __shared__ double ob[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] = g[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) a += Sinv[36 * s + 6 * k + j] * sm[j][0];
        const double v = y[r0 + rows + k] - alpha * a;
        ob[k] = v;
        y[r0 + rows + k] = v;
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < rows; e += blockDim.x) {
        const int64_t i = e / n6, r = e % n6;
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) a += W[o + i * 6 * n6 + r * 6 + k] * ob[k];
        y[r0 + e] -= a;
    }
    const double fd = bc.fac * bc.dt_fac;
    for (int64_t m = threadIdx.x; m < bc.nblob; m += blockDim.x) {
        double* yb = y + r0 + rows + 6 + 3 * m;
        const double* d = bc.disp + 3 * ((int64_t)s * bc.nblob + m);
        yb[0] -= fd * (ob[0] + (ob[4] * d[2] - ob[5] * d[1]));
        yb[1] -= fd * (ob[1] + (ob[5] * d[0] - ob[3] * d[2]));
        yb[2] -= fd * (ob[2] + (ob[3] * d[1] - ob[4] * d[0]));
    }
}

// As block_gemv_kernel for blocks of 64 rows or more: 1024 threads per
// block, eight column groups of 128 rows each, the groups added in a fixed
// order through LDS.
__global__ __launch_bounds__(1024) void block_gemv_wide_kernel(const double* __restrict__ A,
                                                               const int64_t* __restrict__ mat_off,
                                                               const int32_t* __restrict__ row_off,
                                                               const int32_t* __restrict__ sizes,
                                                               const double* __restrict__ x, double* __restrict__ y,
                                                               double alpha) {
    __shared__ double part[8][128];
    const int b = blockIdx.x;
    const int bs = sizes[b];
    const double* Ab = A + mat_off[b];
    const int r0 = row_off[b];
    const int rl = threadIdx.x & 127, q = threadIdx.x >> 7;
    const int cq = (bs + 7) / 8;
    const int c0 = q * cq, c1 = c0 + cq < bs ? c0 + cq : bs;
    for (int rb = 0; rb < bs; rb += 128) {
        const int r = rb + rl;
        double s0 = 0.0, s1 = 0.0;
        if (r < bs) {
            int c = c0;
            for (; c + 8 <= c1; c += 8) {  // eight loads in flight per batch
                double a[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] = Ab[(int64_t)(c + u) * bs + r];
#pragma unroll
                for (int u = 0; u < 8; u += 2) {
                    s0 += a[u] * x[r0 + c + u];
                    s1 += a[u + 1] * x[r0 + c + u + 1];
                }
            }
            for (; c < c1; ++c) s0 += Ab[(int64_t)c * bs + r] * x[r0 + c];
        }
        part[q][rl] = s0 + s1;
        __syncthreads();
        if (q == 0 && r < bs)
            y[r0 + r] = alpha * (((part[0][rl] + part[1][rl]) + (part[2][rl] + part[3][rl])) +
                                 ((part[4][rl] + part[5][rl]) + (part[6][rl] + part[7][rl])));
        __syncthreads();
    }
}

// h[c] = D(:, c) . e for c < k: workgroup c, its four waves over rows
// 256 i + 64 w + lane (four accumulation chains), a fixed shuffle tree per
// wave and the waves added in order.
__global__ __launch_bounds__(256) void lowrank_dots_kernel(const double* __restrict__ D, int64_t ld, int k,
                                                           const double* __restrict__ e, int n,
                                                           double* __restrict__ h,
                                                           const int32_t* __restrict__ d_k) {
    __shared__ double wsum[4];
    const int c = (int)blockIdx.x;
    if (d_k && c >= (*d_k < k ? *d_k : k)) return;  // device count: k is the largest
    const double* d = D + (int64_t)c * ld;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int r = (int)threadIdx.x;
    for (; r + 768 < n; r += 1024) {
        s0 += d[r] * e[r];
        s1 += d[r + 256] * e[r + 256];
        s2 += d[r + 512] * e[r + 512];
        s3 += d[r + 768] * e[r + 768];
    }
    for (; r < n; r += 256) s0 += d[r] * e[r];
    double s = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) h[c] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

// out[r] -= sum_{c < k} C(r, c) h[c] (four accumulation chains).
__global__ __launch_bounds__(256) void lowrank_gemv_kernel(const double* __restrict__ C, int64_t ld, int k,
                                                           const double* __restrict__ h, double* __restrict__ out,
                                                           int n, const int32_t* __restrict__ d_k) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    if (d_k) k = *d_k < k ? *d_k : k;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = 0;
    for (; c + 8 <= k; c += 8) {  // eight loads in flight per batch
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = C[(int64_t)(c + u) * ld + r];
        s0 += a[0] * h[c] + a[4] * h[c + 4];
        s1 += a[1] * h[c + 1] + a[5] * h[c + 5];
        s2 += a[2] * h[c + 2] + a[6] * h[c + 6];
        s3 += a[3] * h[c + 3] + a[7] * h[c + 7];
    }
    for (; c < k; ++c) s0 += C[(int64_t)c * ld + r] * h[c];
    out[r] -= (s0 + s1) + (s2 + s3);
}

// Device-counted end of iteration: when the residual was too large
// (gate[0] > 1) and iteration it = *d_iter < max_iter - 1, C(:, it) and
// D(:, it) = tC, tD over ||tD|| (every workgroup sums the G partials of
// ||tD||^2 itself, in order) and the new residual becomes the engine's error
// (slot0 = slot1).
__global__ __launch_bounds__(256) void commit_update_kernel(const double* __restrict__ tC,
                                                            const double* __restrict__ tD, double* __restrict__ C,
                                                            double* __restrict__ D, int64_t ld, int n,
                                                            double* __restrict__ slot0,
                                                            const double* __restrict__ slot1,
                                                            const double* __restrict__ partial, int G,
                                                            const double* __restrict__ gate,
                                                            const int32_t* __restrict__ d_iter, int max_iter,
                                                            const int32_t* __restrict__ pending) {
    __shared__ double sh[4];
    __shared__ double inv;
    const int it = *d_iter;
    if ((pending && !*pending) || !(gate[0] > 1.0) || it >= max_iter - 1) return;  // uniform
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) s += partial[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double nrm = __builtin_sqrt((sh[0] + sh[1]) + (sh[2] + sh[3]));
        inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    }
    __syncthreads();
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) {
        C[(int64_t)it * ld + r] = tC[r] * inv;
        D[(int64_t)it * ld + r] = tD[r] * inv;
        slot0[r] = slot1[r];
    }
}

// The end of iteration for up to kEndSmall unknowns in one workgroup:
// tD = slot1 - slot0 and its norm, the commit of C(:, it) = tC / ||tD||,
// D(:, it) = tD / ||tD|| and slot0 = slot1.  Device-counted (d_iter set): the
// commit is gated by the residual's flag and the count, the count advances,
// and nothing runs unless an iteration is pending; otherwise iteration it_fixed
// commits unconditionally.  The norm sums 1024 per-thread partials in a fixed
// tree.
__global__ __launch_bounds__(1024) void end_small_kernel(const double* __restrict__ tC, double* __restrict__ tD,
                                                         double* __restrict__ C, double* __restrict__ D,
                                                         int64_t ld, int n, double* __restrict__ slot0,
                                                         const double* __restrict__ slot1,
                                                         const double* __restrict__ gate, int32_t* __restrict__ d_iter,
                                                         int it_fixed, int max_iter, int32_t* __restrict__ pending) {
    __shared__ double sh[16];
    __shared__ double inv;
    const int tid = (int)threadIdx.x;
    if (pending && !*pending) return;  // uniform
    const int it = d_iter ? *d_iter : it_fixed;
    const bool commit = d_iter ? gate[0] > 1.0 && it < max_iter - 1 : true;
    double s = 0.0;
    for (int r = tid; r < n; r += 1024) {
        const double d = slot1[r] - slot0[r];
        tD[r] = d;
        s += d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((tid & 63) == 0) sh[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += sh[w];
        const double nrm = __builtin_sqrt(t);
        inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    }
    __syncthreads();
    if (commit) {
        for (int r = tid; r < n; r += 1024) {
            C[(int64_t)it * ld + r] = tC[r] * inv;
            D[(int64_t)it * ld + r] = tD[r] * inv;
            slot0[r] = slot1[r];
        }
    }
    __syncthreads();  // every thread has read *d_iter
    if (tid == 0 && d_iter) {
        *d_iter = it + 1;
        if (pending) *pending = 0;
    }
}

__global__ void set_count_kernel(int32_t* __restrict__ c, int32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

// One more iteration, if one is pending (which it then consumes).
__global__ void advance_count_kernel(int32_t* __restrict__ c, int32_t* __restrict__ pending) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || (pending && !*pending)) return;
    *c += 1;
    if (pending) *pending = 0;
}

// D_k = s1 - s0 and partial[b] = sum over workgroup b's rows of D_k^2.
__global__ __launch_bounds__(256) void diff_sumsq_kernel(const double* __restrict__ s1, const double* __restrict__ s0,
                                                         double* __restrict__ Dk, int n,
                                                         double* __restrict__ partial) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += 256 * gridDim.x) {
        const double d = s1[r] - s0[r];
        Dk[r] = d;
        s += d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// C_k and D_k divided by ||D_k|| = sqrt(sum_b partial[b]) (every workgroup
// sums the G partials itself, in the same order); scal[0] = ||D_k||.
__global__ __launch_bounds__(256) void scale_pair_kernel(double* __restrict__ Ck, double* __restrict__ Dk, int n,
                                                         const double* __restrict__ partial, int G,
                                                         double* __restrict__ scal) {
    __shared__ double sh[4];
    __shared__ double inv;
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) s += partial[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double nrm = __builtin_sqrt((sh[0] + sh[1]) + (sh[2] + sh[3]));
        inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        if (blockIdx.x == 0) scal[0] = nrm;
    }
    __syncthreads();
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) {
        Ck[r] *= inv;
        Dk[r] *= inv;
    }
}

// out = dexpinv(u, v) = v - u x v / 2 - fac u x (u x v),
// fac = ((theta/2) cot(theta/2) - 1) / theta^2, -1/12 below theta = 1e-5
// (quaternion.cpp:8, 610-633).
__device__ inline void dexpinv(double* out, const double* u, const double* v) {
#pragma clang fp contract(off)
    const double t2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double th = sqrt(t2);
    const double fac = th < 1e-5 ? -1.0 / 12.0 : (0.5 * th * cos(0.5 * th) / sin(0.5 * th) - 1.0) / t2;
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    out[0] = v[0] - 0.5 * c[0] - fac * (u[1] * c[2] - u[2] * c[1]);
    out[1] = v[1] - 0.5 * c[1] - fac * (u[2] * c[0] - u[0] * c[2]);
    out[2] = v[2] - 0.5 * c[2] - fac * (u[0] * c[1] - u[1] * c[0]);
}

struct ResidualArgs {
    int nswim, nfil, nseg, bdf2;
    double dt, pos_tol, lie_tol;
    const double *x, *xm1, *xm2, *u, *um1, *v;
    const double *bx, *bxm1, *bxm2, *bu, *bum1, *bv;
    int nblob;                        // blobs per swimmer (Broyden-for-everything form) or 0
    const double *xb, *vb;            // 3 per blob
    // free bodies in the Broyden-for-everything form: the body rows are the
    // force and torque balance (written by body_balance_kernel) and the blob
    // rows take the velocities the body's positions imply
    int free_bfe;
    const double *fs, *fb;            // 6 per segment, 3 per blob
};

// out = dexp(u, v) = v + alpha u x v + beta u x (u x v) (quaternion.cpp:574-608).
__device__ inline void dexp(double* out, const double* u, const double* v) {
#pragma clang fp contract(off)
    const double t2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double t = sqrt(t2);
    double alpha, beta;
    if (t < 1e-5) {
        const double t4 = t2 * t2;
        beta = (1.0 / 3.0 + t2 / 72.0) / (2.0 + t2 / 6.0 + t4 / 72.0);
        alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - t2 / 12.0);
    } else {
        alpha = sin(0.5 * t);
        alpha *= 2.0 * alpha / t2;
        beta = (t - sin(t)) / (t * t2);
    }
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const double cc[3] = {u[1] * c[2] - u[2] * c[1], u[2] * c[0] - u[0] * c[2], u[0] * c[1] - u[1] * c[0]};
    for (int k = 0; k < 3; ++k) out[k] = v[k] + alpha * c[k] + beta * cc[k];
}

// The body rows of a free body in the Broyden-for-everything form
// (mobility_solver.cpp:2676-2681, 2837-2847): -swimmer.f (its segments' forces
// and their torques about the body's centre, swimmer.cpp:469-499) plus the blob
// forces and their torques.  One workgroup per swimmer.
__global__ __launch_bounds__(256) void body_balance_kernel(ResidualArgs a, double* __restrict__ err) {
    const int s = blockIdx.x;
    const int64_t nsw = (int64_t)a.nfil * a.nseg;
    const int64_t per = 6 * (nsw + 1) + 3LL * a.nblob;
    const double* X = a.bx + 3 * s;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = threadIdx.x; i < nsw + a.nblob; i += 256) {
        const double *f, *x;
        double T[3] = {0.0, 0.0, 0.0};
        if (i < nsw) {
            const int64_t g = s * nsw + i;
            f = a.fs + 6 * g;
            x = a.x + 3 * g;
            T[0] = f[3];
            T[1] = f[4];
            T[2] = f[5];
        } else {
            const int64_t g = s * (int64_t)a.nblob + (i - nsw);
            f = a.fb + 3 * g;
            x = a.xb + 3 * g;
        }
        const double r[3] = {x[0] - X[0], x[1] - X[1], x[2] - X[2]};
        acc[0] += f[0];
        acc[1] += f[1];
        acc[2] += f[2];
        acc[3] += T[0] + (r[1] * f[2] - r[2] * f[1]);
        acc[4] += T[1] + (r[2] * f[0] - r[0] * f[2]);
        acc[5] += T[2] + (r[0] * f[1] - r[1] * f[0]);
    }
    __shared__ double sm[6][256];
    for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) err[s * per + 6 * nsw + threadIdx.x] = sm[threadIdx.x][0];
}

// One thread per segment (then one per body): 3 position and 3 Lie-algebra
// residuals in the reference's storage order (mobility_solver.cpp:2661-2946),
// evaluated without contraction in the reference's operation order.
// partial[2*block] = max |e|/tolerance, partial[2*block+1] = 1 if any residual
// of the block is not finite.
__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a, double* __restrict__ err,
                                                       double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double sm[256];
    __shared__ double sn[256];
    const int64_t nsegs = (int64_t)a.nswim * a.nfil * a.nseg;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const double c1 = -4.0 / 3.0, c2 = 1.0 / 3.0, c3 = -2.0 / 3.0;
    // per-swimmer stride of the Broyden vector
    const int64_t per = 6 * ((int64_t)a.nfil * a.nseg + 1) + 3LL * a.nblob;
    double ratio = 0.0, nonfinite = 0.0;
    if (g >= nsegs + a.nswim && g < nsegs + a.nswim + (int64_t)a.nswim * a.nblob) {
        // blob rows (mobility_solver.cpp:2812-2835): V + Omega x (x_blob - X) - v_blob
        const int64_t m = g - nsegs - a.nswim;
        const int64_t n = m / a.nblob, k = m % a.nblob;
        const double* V = a.bv + 6 * n;
        double Vf[6];
        if (a.free_bfe) {
            // the body's velocities from its positions (mobility_solver.cpp:2786-2806)
            const double *x = a.bx + 3 * n, *xm1 = a.bxm1 + 3 * n, *u = a.bu + 3 * n;
            if (a.bdf2) {
                const double *xm2 = a.bxm2 + 3 * n, *um1 = a.bum1 + 3 * n;
                double w[3];
                for (int c = 0; c < 3; ++c) {
                    Vf[c] = -(x[c] + c1 * xm1[c] + c2 * xm2[c]) / (c3 * a.dt);
                    w[c] = (c2 * um1[c] - u[c]) / (c3 * a.dt);
                }
                dexp(Vf + 3, u, w);
            } else {
                for (int c = 0; c < 3; ++c) {
                    Vf[c] = (x[c] - xm1[c]) / a.dt;
                    Vf[3 + c] = u[c] / a.dt;
                }
            }
            V = Vf;
        }
        double d[3];
        for (int c = 0; c < 3; ++c) d[c] = a.xb[3 * m + c] - a.bx[3 * n + c];
        const double w[3] = {V[4] * d[2] - V[5] * d[1], V[5] * d[0] - V[3] * d[2], V[3] * d[1] - V[4] * d[0]};
        const int64_t rid = n * per + 6 * ((int64_t)a.nfil * a.nseg + 1) + 3 * k;
        for (int c = 0; c < 3; ++c) {
            const double e = V[c] + w[c] - a.vb[3 * m + c];
            err[rid + c] = e;
            if (!isfinite(e)) nonfinite = 1.0;
            ratio = fmax(ratio, fabs(e) / a.lie_tol);
        }
    } else if (g < nsegs + a.nswim) {
        const bool body = g >= nsegs;
        int64_t pid, lid;
        const double *x, *xm1, *xm2, *u, *um1, *vel, *om;
        if (!body) {
            const int64_t n = g / ((int64_t)a.nfil * a.nseg);
            const int64_t r = g % ((int64_t)a.nfil * a.nseg);
            const int64_t i = r / a.nseg, j = r % a.nseg;
            pid = n * per + 6 * i * a.nseg + 3 * j;
            lid = pid + 3 * a.nseg;
            x = a.x + 3 * g;
            xm1 = a.xm1 + 3 * g;
            xm2 = a.xm2 ? a.xm2 + 3 * g : nullptr;
            u = a.u + 3 * g;
            um1 = a.um1 ? a.um1 + 3 * g : nullptr;
            vel = a.v + 6 * g;
            om = a.v + 6 * g + 3;
        } else if (a.free_bfe) {
            // force and torque balance, formed by body_balance_kernel (tolerance TOL)
            const int64_t n = g - nsegs;
            pid = n * per + 6 * (int64_t)a.nfil * a.nseg;
            for (int d = 0; d < 6; ++d) {
                const double e = err[pid + d];
                if (!isfinite(e)) nonfinite = 1.0;
                ratio = fmax(ratio, fabs(e) / a.lie_tol);
            }
            x = nullptr;
        } else {
            const int64_t n = g - nsegs;
            pid = n * per + 6 * (int64_t)a.nfil * a.nseg;
            lid = pid + 3;
            x = a.bx + 3 * n;
            xm1 = a.bxm1 + 3 * n;
            xm2 = a.bxm2 ? a.bxm2 + 3 * n : nullptr;
            u = a.bu + 3 * n;
            um1 = a.bum1 ? a.bum1 + 3 * n : nullptr;
            vel = a.bv + 6 * n;
            om = a.bv + 6 * n + 3;
        }
        double W[3];
        if (x) dexpinv(W, u, om);
        const double c3dt = c3 * a.dt;
        for (int d = 0; d < 3 && x; ++d) {
            double ep, el;
            if (a.bdf2) {
                ep = x[d] + c1 * xm1[d] + c2 * xm2[d] + c3dt * vel[d];
                el = u[d] - c2 * um1[d] + c3dt * W[d];
            } else {
                ep = x[d] - xm1[d] - a.dt * vel[d];
                el = u[d] - a.dt * W[d];
            }
            err[pid + d] = ep;
            err[lid + d] = el;
            if (!isfinite(ep) || !isfinite(el)) nonfinite = 1.0;
            ratio = fmax(ratio, fmax(fabs(ep) / a.pos_tol, fabs(el) / a.lie_tol));
        }
    }
    sm[threadIdx.x] = ratio;
    sn[threadIdx.x] = nonfinite;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
            sn[threadIdx.x] = fmax(sn[threadIdx.x], sn[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sm[0];
        partial[2 * blockIdx.x + 1] = sn[0];
    }
}

// out[0] = max ratio, out[1] = nonfinite flag over nb partial pairs (one block).
// (book: see tfs::broyden_set_bookkeeping, may be null)
__global__ __launch_bounds__(256) void residual_finish_kernel(const double* __restrict__ partial, int nb,
                                                              double* __restrict__ out, int32_t* __restrict__ book,
                                                              int book_max) {
    __shared__ double sm[256];
    __shared__ double sn[256];
    double m = 0.0, f = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        m = fmax(m, partial[2 * b]);
        f = fmax(f, partial[2 * b + 1]);
    }
    sm[threadIdx.x] = m;
    sn[threadIdx.x] = f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
            sn[threadIdx.x] = fmax(sn[threadIdx.x], sn[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0];
        out[1] = sn[0];
        // one more iteration, unless the iterations had stopped; stop once the
        // residual is small enough, not finite, or book_max iterations have run
        if (book && !book[2]) {
            const bool nonfinite = sn[0] != 0.0;
            book[0] += 1;
            book[1] = nonfinite ? 0 : 1;
            if (nonfinite) book[2] = 2;
            else if (!(sm[0] > 1.0) || book[0] >= book_max) book[2] = 1;
        }
    }
}

int check(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }

int64_t round32(int64_t l) { return ((l + 31) / 32) * 32; }

}  // namespace

struct tfs_broyden {
    tfs_broyden_desc desc{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int n = 0, max_iter = 0, iter = 0, nblocks = 0, G = 1, RB = 1;
    int latest = 0;  // slot holding the latest residual handed to the engine
    int64_t ld = 0, jac_len = 0;
    double* arena = nullptr;
    double *C = nullptr, *D = nullptr, *slot[2] = {nullptr, nullptr}, *t = nullptr, *h = nullptr,
           *partial = nullptr, *scal = nullptr, *J = nullptr, *Jstage = nullptr;
    int64_t* mat_off = nullptr;
    int32_t* row_off = nullptr;
    int32_t* sizes = nullptr;
    int32_t* piv = nullptr;
    int32_t* status = nullptr;
    int max_block = 0;
    // device-counted iterations (tfs::broyden_dev_*): the iteration count on
    // the device, the next C and D columns before their commit
    int32_t* d_iter = nullptr;
    double *tC = nullptr, *tD = nullptr;
    int32_t* book = nullptr;  // iteration bookkeeping after each residual (tfs::broyden_set_bookkeeping)
    int book_max = 0;
    double* host_scal = nullptr;
    std::vector<int32_t> host_status;
    std::vector<int32_t> host_sizes;
    std::vector<int64_t> host_moff;
    // free-body coupling (tfs_broyden_set_body_coupling[_blobs])
    bool coupled = false;
    int cpl_swim = 0, cpl_fil = 0, cpl_n6 = 0;
    double* cpl_arena = nullptr;
    double *W = nullptr, *BJ = nullptr, *Sinv = nullptr, *Sbuf = nullptr, *cpl_in = nullptr;
    BlobCoupling cpl_blobs;
    double* cpl_disp = nullptr;  // device, 3 per blob
    int64_t cpl_disp_len = 0;
};

namespace {

bool host_mem(const tfs_broyden* b) { return b->desc.memory == TFS_MEMORY_HOST; }

// Up to this many unknowns the low-rank sums take one wave per stored column.
constexpr int kWaveDotsMax = 65536;
// Up to this many unknowns the device-counted end of iteration is one workgroup.
constexpr int kEndSmall = 8192;

hipMemcpyKind to_device(const tfs_broyden* b) {
    return host_mem(b) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
}

int apply_jinv(tfs_broyden* b, const double* x, double* y, double alpha) {
    if (b->max_block >= 64)
        hipLaunchKernelGGL(block_gemv_wide_kernel, dim3(b->nblocks), dim3(1024), 0, b->stream, b->J, b->mat_off,
                           b->row_off, b->sizes, x, y, alpha);
    else
        hipLaunchKernelGGL(block_gemv_kernel, dim3(b->nblocks), dim3(128), 0, b->stream, b->J, b->mat_off,
                           b->row_off, b->sizes, x, y, alpha);
    int rc = check(hipGetLastError(), "J0^{-1}");
    if (!rc && b->coupled) {
        hipLaunchKernelGGL(coupling_apply_kernel, dim3(b->cpl_swim), dim3(256), 0, b->stream, (const double*)b->W,
                           (const double*)b->BJ, (const double*)b->Sinv, b->cpl_fil, b->cpl_n6, b->cpl_blobs, x, y,
                           alpha);
        rc = check(hipGetLastError(), "J0^{-1} body coupling");
    }
    return rc;
}

// out = -c J0^{-1} e - C(:, 0:iter) (D(:, 0:iter)^T e)
int inverse_update(tfs_broyden* b, const double* e, double* out) {
    int rc = apply_jinv(b, e, out, -b->desc.confidence);
    const int k = std::min(b->iter, b->max_iter - 1);
    if (rc || k == 0) return rc;
    if (b->n <= kWaveDotsMax) {
        hipLaunchKernelGGL(lowrank_dots_kernel, dim3((unsigned)k), dim3(256), 0, b->stream,
                           (const double*)b->D, (int64_t)b->ld, k, e, b->n, b->h, (const int32_t*)nullptr);
        hipLaunchKernelGGL(lowrank_gemv_kernel, dim3((unsigned)((b->n + 255) / 256)), dim3(256), 0, b->stream,
                           (const double*)b->C, (int64_t)b->ld, k, (const double*)b->h, out, b->n,
                           (const int32_t*)nullptr);
        return check(hipGetLastError(), "low-rank update");
    }
    rc = check(k_dot_cols(b->D, b->ld, k, e, b->n, b->partial, b->G, b->stream), "D^T e");
    if (!rc) rc = check(k_reduce_cols(b->partial, b->G, k, b->h, nullptr, b->stream), "reduce");
    if (!rc) rc = check(k_gemv(b->C, b->ld, k, b->h, out, b->n, true, b->stream), "C h");
    return rc;
}

int copy_in(tfs_broyden* b, double* dst, const double* src, int64_t count, const char* where) {
    hipError_t e = hipMemcpyAsync(dst, src, count * 8, to_device(b), b->stream);
    if (e == hipSuccess && host_mem(b)) e = hipStreamSynchronize(b->stream);
    return check(e, where);
}

}  // namespace

extern "C" {

int tfs_broyden_create(const tfs_broyden_desc* desc, tfs_broyden** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_create: null argument");
    *out = nullptr;
    const tfs_broyden_desc& D = *desc;
    if (D.n < 1 || D.max_iter < 1 || D.num_blocks < 1 || !D.block_sizes || !(D.confidence > 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_create: bad sizes or confidence");
    if (D.memory != TFS_MEMORY_HOST && D.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_create: unknown memory space");
    std::vector<int64_t> moff(D.num_blocks);
    std::vector<int32_t> roff(D.num_blocks), sz(D.block_sizes, D.block_sizes + D.num_blocks);
    int64_t rows = 0, mats = 0;
    int max_block = 0;
    for (int i = 0; i < D.num_blocks; ++i) {
        if (sz[i] < 1 || sz[i] > 2048)
            return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_create: block sizes must lie in [1, 2048]");
        roff[i] = (int32_t)rows;
        moff[i] = mats;
        rows += sz[i];
        mats += (int64_t)sz[i] * sz[i];
        max_block = std::max(max_block, (int)sz[i]);
    }
    if (rows != D.n) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_create: block sizes do not sum to n");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TFS_ERR_NO_DEVICE, "tfs_broyden_create: no HIP device visible");
    tfs_broyden* b = new tfs_broyden();
    b->desc = D;
    b->desc.block_sizes = nullptr;
    b->n = D.n;
    b->max_iter = D.max_iter;
    b->nblocks = D.num_blocks;
    b->max_block = max_block;
    b->ld = round32(D.n);
    b->jac_len = mats;
    b->G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (D.n + 255) / 256));
    b->RB = (int)((D.n / 3 + 256) / 256 + 1);   // residual blocks: one thread per segment, body or blob
    b->host_status.resize(D.num_blocks);
    b->host_sizes = sz;
    b->host_moff = moff;
    hipError_t e = hipSuccess;
    if (D.stream) {
        b->stream = (hipStream_t)D.stream;
    } else {
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        b->own_stream = e == hipSuccess;
    }
    const int64_t parts[] = {b->ld * D.max_iter, b->ld * D.max_iter, b->ld, b->ld, b->ld, D.max_iter + 1,
                             std::max<int64_t>((int64_t)b->G * (D.max_iter + 1), 2LL * b->RB), 8, mats, mats,
                             b->ld, b->ld};
    double** dst[] = {&b->C, &b->D, &b->slot[0], &b->slot[1], &b->t, &b->h, &b->partial, &b->scal, &b->J,
                      &b->Jstage, &b->tC, &b->tD};
    int64_t total = 0;
    for (int64_t l : parts) total += round32(l);
    if (e == hipSuccess) e = hipMalloc(&b->arena, total * 8);
    if (e == hipSuccess) e = hipMalloc(&b->mat_off, sizeof(int64_t) * D.num_blocks);
    if (e == hipSuccess) e = hipMalloc(&b->row_off, sizeof(int32_t) * D.num_blocks);
    if (e == hipSuccess) e = hipMalloc(&b->sizes, sizeof(int32_t) * D.num_blocks);
    if (e == hipSuccess) e = hipMalloc(&b->piv, sizeof(int32_t) * D.n);
    if (e == hipSuccess) e = hipMalloc(&b->status, sizeof(int32_t) * D.num_blocks);
    if (e == hipSuccess) e = hipMalloc(&b->d_iter, sizeof(int32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->host_scal, 64, hipHostMallocDefault);
    if (e == hipSuccess) {
        double* p = b->arena;
        for (size_t i = 0; i < sizeof(parts) / sizeof(parts[0]); ++i) {
            *dst[i] = p;
            p += round32(parts[i]);
        }
        e = hipMemsetAsync(b->arena, 0, total * 8, b->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->mat_off, moff.data(), sizeof(int64_t) * D.num_blocks, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->row_off, roff.data(), sizeof(int32_t) * D.num_blocks, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->sizes, sz.data(), sizeof(int32_t) * D.num_blocks, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        tfs_broyden_destroy(b);
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_create: device memory");
    }
    *out = b;
    return TFS_OK;
}

int tfs_broyden_destroy(tfs_broyden* b) {
    if (!b) return TFS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    (void)hipFree(b->arena);
    (void)hipFree(b->mat_off);
    (void)hipFree(b->row_off);
    (void)hipFree(b->sizes);
    (void)hipFree(b->piv);
    (void)hipFree(b->status);
    if (b->d_iter) (void)hipFree(b->d_iter);
    if (b->cpl_arena) (void)hipFree(b->cpl_arena);
    if (b->cpl_disp) (void)hipFree(b->cpl_disp);
    if (b->host_scal) (void)hipHostFree(b->host_scal);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TFS_OK;
}

int tfs_broyden_set_jacobian_inverse(tfs_broyden* b, const double* blocks) {
    if (!b || !blocks) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_set_jacobian_inverse: null argument");
    b->coupled = false;
    int rc = copy_in(b, b->Jstage, blocks, b->jac_len, "tfs_broyden_set_jacobian_inverse");
    if (rc) return rc;
    hipLaunchKernelGGL(block_transpose_kernel, dim3(b->nblocks), dim3(256), 0, b->stream, b->Jstage, b->J,
                       b->mat_off, b->sizes);
    return check(hipGetLastError(), "tfs_broyden_set_jacobian_inverse");
}

int tfs_broyden_set_jacobian(tfs_broyden* b, const double* blocks) {
    if (!b || !blocks) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_set_jacobian: null argument");
    b->coupled = false;
    int rc = copy_in(b, b->Jstage, blocks, b->jac_len, "tfs_broyden_set_jacobian");
    if (rc) return rc;
    const size_t lds = sizeof(double) * 2 * b->max_block;
    rc = check(k_batched_inverse(b->Jstage, b->mat_off, b->sizes, b->nblocks, b->max_block, b->piv, b->row_off,
                                 b->status, b->stream),
               "Gauss-Jordan");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(b->host_status.data(), b->status, sizeof(int32_t) * b->nblocks,
                                  hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_broyden_set_jacobian");
    for (int i = 0; i < b->nblocks; ++i)
        if (b->host_status[i])
            return fail(TFS_ERR_NON_FINITE,
                        "tfs_broyden_set_jacobian: block " + std::to_string(i) + " is singular or not finite");
    hipLaunchKernelGGL(block_transpose_kernel, dim3(b->nblocks), dim3(256), 0, b->stream, b->Jstage, b->J,
                       b->mat_off, b->sizes);
    return check(hipGetLastError(), "tfs_broyden_set_jacobian");
}

}  // extern "C"

namespace {

int set_coupling(tfs_broyden* b, int32_t num_swim, int32_t fil_per_swim, int32_t nblob, const double* A,
                 const double* B, const double* S0, const double* disp, double fac, double dt_fac) {
    if (!b || !A || !B || !S0 || (nblob > 0 && !disp))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_set_body_coupling: null argument");
    const int per_swim = fil_per_swim + 1 + nblob;
    if (num_swim < 1 || fil_per_swim < 1 || nblob < 0 || (int64_t)num_swim * per_swim != b->nblocks)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_broyden_set_body_coupling: block count is not num_swim * (fil_per_swim + 1 + blobs)");
    const int n6 = b->host_sizes[0];
    for (int s = 0; s < num_swim; ++s)
        for (int i = 0; i < per_swim; ++i)
            if (b->host_sizes[(size_t)s * per_swim + i] != (i < fil_per_swim ? n6 : i == fil_per_swim ? 6 : 3))
                return fail(TFS_ERR_INVALID_ARGUMENT,
                            "tfs_broyden_set_body_coupling: blocks are not fil_per_swim equal blocks, a 6 x 6 and "
                            "3 x 3 blob blocks per swimmer");
    const int64_t lab = (int64_t)num_swim * fil_per_swim * 6 * n6;
    if (b->cpl_arena && (b->cpl_swim != num_swim || b->cpl_fil != fil_per_swim)) {
        (void)hipFree(b->cpl_arena);
        b->cpl_arena = nullptr;
    }
    if (!b->cpl_arena) {
        const int64_t parts[] = {lab, lab, 36LL * num_swim, 36LL * num_swim, 2 * lab + 36LL * num_swim};
        double** dst[] = {&b->W, &b->BJ, &b->Sinv, &b->Sbuf, &b->cpl_in};
        int64_t total = 0;
        for (int64_t l : parts) total += round32(l);
        if (hipMalloc(&b->cpl_arena, total * 8) != hipSuccess)
            return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_set_body_coupling: device memory");
        double* p = b->cpl_arena;
        for (size_t i = 0; i < sizeof(parts) / sizeof(parts[0]); ++i) {
            *dst[i] = p;
            p += round32(parts[i]);
        }
        b->cpl_swim = num_swim;
        b->cpl_fil = fil_per_swim;
    }
    b->cpl_n6 = n6;
    b->coupled = false;
    const double *dA = A, *dB = B, *dS0 = S0;
    int rc = TFS_OK;
    if (host_mem(b)) {
        rc = copy_in(b, b->cpl_in, A, lab, "tfs_broyden_set_body_coupling");
        if (!rc) rc = copy_in(b, b->cpl_in + lab, B, lab, "tfs_broyden_set_body_coupling");
        if (!rc) rc = copy_in(b, b->cpl_in + 2 * lab, S0, 36LL * num_swim, "tfs_broyden_set_body_coupling");
        dA = b->cpl_in;
        dB = b->cpl_in + lab;
        dS0 = b->cpl_in + 2 * lab;
    }
    if (rc) return rc;
    b->cpl_blobs = BlobCoupling{};
    if (!rc && nblob > 0) {
        const int64_t len = 3LL * num_swim * nblob;
        if (b->cpl_disp_len < len) {
            if (b->cpl_disp) (void)hipFree(b->cpl_disp);
            b->cpl_disp = nullptr;
            b->cpl_disp_len = 0;
            if (hipMalloc(&b->cpl_disp, len * 8) != hipSuccess)
                return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_set_body_coupling: device memory");
            b->cpl_disp_len = len;
        }
        rc = copy_in(b, b->cpl_disp, disp, len, "tfs_broyden_set_body_coupling");
        b->cpl_blobs.nblob = nblob;
        b->cpl_blobs.disp = b->cpl_disp;
        b->cpl_blobs.fac = fac;
        b->cpl_blobs.dt_fac = dt_fac;
    }
    if (rc) return rc;
    hipLaunchKernelGGL(coupling_prep_kernel, dim3((unsigned)(num_swim * fil_per_swim)), dim3(256), 0, b->stream,
                       (const double*)b->J, (const int64_t*)b->mat_off, fil_per_swim, per_swim, n6, dA, dB, b->W,
                       b->BJ);
    rc = check(hipGetLastError(), "body coupling");
    if (!rc) {
        hipLaunchKernelGGL(coupling_schur_kernel, dim3((unsigned)num_swim), dim3(64), 0, b->stream, dS0,
                           (const double*)b->BJ, dA, fil_per_swim, n6, b->Sbuf);
        rc = check(hipGetLastError(), "body Schur complement");
    }
    std::vector<double> S(36 * (size_t)num_swim), Si(36 * (size_t)num_swim);
    if (!rc) rc = check(hipMemcpyAsync(S.data(), b->Sbuf, 288LL * num_swim, hipMemcpyDeviceToHost, b->stream), "Schur");
    if (!rc) rc = check(hipStreamSynchronize(b->stream), "Schur");
    if (rc) return rc;
    // invert each 6 x 6 (Gauss-Jordan, partial pivoting)
    for (int s = 0; s < num_swim; ++s) {
        double a[6][12];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 12; ++c) a[r][c] = c < 6 ? S[36 * (size_t)s + 6 * r + c] : (c - 6 == r ? 1.0 : 0.0);
        for (int k = 0; k < 6; ++k) {
            int piv = k;
            for (int r = k + 1; r < 6; ++r)
                if (std::fabs(a[r][k]) > std::fabs(a[piv][k])) piv = r;
            if (!(std::fabs(a[piv][k]) > 0.0) || !std::isfinite(a[piv][k]))
                return fail(TFS_ERR_NON_FINITE, "tfs_broyden_set_body_coupling: singular body Schur complement");
            if (piv != k)
                for (int c = 0; c < 12; ++c) std::swap(a[k][c], a[piv][c]);
            const double inv_p = 1.0 / a[k][k];
            for (int c = 0; c < 12; ++c) a[k][c] *= inv_p;
            for (int r = 0; r < 6; ++r)
                if (r != k) {
                    const double f = a[r][k];
                    for (int c = 0; c < 12; ++c) a[r][c] -= f * a[k][c];
                }
        }
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) Si[36 * (size_t)s + 6 * r + c] = a[r][6 + c];
    }
    rc = check(hipMemcpyAsync(b->Sinv, Si.data(), 288LL * num_swim, hipMemcpyHostToDevice, b->stream), "S^{-1}");
    // the body blocks of J0^{-1}: S^{-1}, column-major like every block
    std::vector<double> colmajor(36 * (size_t)num_swim);
    for (int s = 0; s < num_swim; ++s)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) colmajor[36 * (size_t)s + 6 * c + r] = Si[36 * (size_t)s + 6 * r + c];
    for (int s = 0; s < num_swim && !rc; ++s) {
        const size_t blk = (size_t)s * per_swim + fil_per_swim;
        rc = check(hipMemcpyAsync(b->J + b->host_moff[blk], colmajor.data() + 36 * (size_t)s, 288,
                                  hipMemcpyHostToDevice, b->stream),
                   "S^{-1} block");
    }
    if (!rc) rc = check(hipStreamSynchronize(b->stream), "tfs_broyden_set_body_coupling");
    if (!rc) b->coupled = true;
    return rc;
}

}  // namespace

extern "C" {

int tfs_broyden_set_body_coupling(tfs_broyden* b, int32_t num_swim, int32_t fil_per_swim, const double* A,
                                  const double* B, const double* S0) {
    return set_coupling(b, num_swim, fil_per_swim, 0, A, B, S0, nullptr, 0.0, 0.0);
}

int tfs_broyden_set_body_coupling_blobs(tfs_broyden* b, int32_t num_swim, int32_t fil_per_swim,
                                        int32_t blobs_per_swim, const double* A, const double* B, const double* S0,
                                        const double* blob_disp, double blob_mob_fac, double dt_fac) {
    if (blobs_per_swim < 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_set_body_coupling_blobs: blobs_per_swim < 1");
    return set_coupling(b, num_swim, fil_per_swim, blobs_per_swim, A, B, S0, blob_disp, blob_mob_fac, dt_fac);
}

int tfs_broyden_start(tfs_broyden* b, const double* error) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_start: null handle");
    b->iter = 0;
    b->latest = 0;
    if (!error) return TFS_OK;
    return copy_in(b, b->slot[0], error, b->n, "tfs_broyden_start");
}

int tfs_broyden_find_update(tfs_broyden* b, double* update) {
    if (!b || !update) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_find_update: null argument");
    double* out = host_mem(b) ? b->t : update;
    int rc = inverse_update(b, b->slot[0], out);
    if (rc) return rc;
    if (host_mem(b)) {
        hipError_t e = hipMemcpyAsync(update, out, (size_t)b->n * 8, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        return check(e, "tfs_broyden_find_update");
    }
    // device vectors: ordered on the caller's stream, or complete on return when
    // the engine runs on a stream of its own
    return b->own_stream ? check(hipStreamSynchronize(b->stream), "tfs_broyden_find_update") : TFS_OK;
}

int tfs_broyden_end_of_iter(tfs_broyden* b, const double* new_error, int32_t error_is_too_large) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_end_of_iter: null handle");
    const int n = b->n;
    if (error_is_too_large && b->iter < b->max_iter - 1) {
        int rc = new_error ? copy_in(b, b->slot[1], new_error, n, "tfs_broyden_end_of_iter") : TFS_OK;
        double* Ck = b->C + (int64_t)b->iter * b->ld;
        double* Dk = b->D + (int64_t)b->iter * b->ld;
        if (!rc && n <= kEndSmall) {
            // one workgroup, as the device-counted form: slot 0 takes the new
            // error by a copy
            rc = inverse_update(b, b->slot[1], b->tC);
            if (!rc) {
                hipLaunchKernelGGL(end_small_kernel, dim3(1), dim3(1024), 0, b->stream, (const double*)b->tC, b->tD,
                                   b->C, b->D, (int64_t)b->ld, n, b->slot[0], (const double*)b->slot[1],
                                   (const double*)nullptr, (int32_t*)nullptr, b->iter, b->max_iter,
                                   (int32_t*)nullptr);
                rc = check(hipGetLastError(), "D_k, C_k / ||D_k||");
            }
            if (rc) return rc;
            b->latest = 0;
            b->iter++;
            return TFS_OK;
        }
        if (!rc) rc = inverse_update(b, b->slot[1], Ck);
        if (!rc) {
            // D_k = new - old, then C_k and D_k over ||D_k||
            hipLaunchKernelGGL(diff_sumsq_kernel, dim3((unsigned)b->G), dim3(256), 0, b->stream,
                               (const double*)b->slot[1], (const double*)b->slot[0], Dk, n, b->partial);
            hipLaunchKernelGGL(scale_pair_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, Ck,
                               Dk, n, (const double*)b->partial, b->G, b->scal);
            rc = check(hipGetLastError(), "D_k, C_k / ||D_k||");
        }
        if (rc) return rc;
        std::swap(b->slot[0], b->slot[1]);   // error <- new_error
        b->latest = 0;
    } else {
        // no update: the engine's error stays in slot 0 (broyden_solver.cpp:69-129),
        // the new residual is kept in slot 1
        if (new_error) {
            int rc = copy_in(b, b->slot[1], new_error, n, "tfs_broyden_end_of_iter");
            if (rc) return rc;
        }
        b->latest = 1;
    }
    b->iter++;
    return TFS_OK;
}

int tfs_broyden_status(tfs_broyden* b, int32_t* iter, double* error_norm) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_status: null handle");
    if (iter) *iter = b->iter;
    if (error_norm) {
        int rc = check(k_norm(b->slot[0], b->n, b->partial, b->G, b->scal, b->stream), "||error||");
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(b->host_scal, b->scal, 8, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) return hip_fail(e, "tfs_broyden_status");
        *error_norm = *b->host_scal;
    }
    return TFS_OK;
}

}  // extern "C"

// Checks the state, stages host arrays (host memory), enqueues the residual
// kernels and (copy_flags) the copy of its two flags (too large, not finite)
// into host_scal[0..1], without waiting.
static int errors_enqueue(tfs_broyden* b, const tfs_broyden_state* st, int32_t slot, std::vector<void*>& staged,
                          bool copy_flags = true) {
    if (!b || !st || (slot != 0 && slot != 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_compute_errors: null argument or bad slot");
    if (st->num_swim < 1 || st->num_fil < 0 || st->num_seg < 0 || st->blobs_per_swim < 0 ||
        6LL * ((int64_t)st->num_swim * st->num_fil * st->num_seg + st->num_swim) +
                3LL * st->num_swim * st->blobs_per_swim != b->n)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_broyden_compute_errors: 6*(NSWIM*NFIL*NSEG + NSWIM) + 3*NSWIM*blobs_per_swim != n");
    if (st->blobs_per_swim > 0 && (!st->x_blobs || !st->v_blobs))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_compute_errors: blob rows need x_blobs and v_blobs");
    if (st->free_bodies && (st->blobs_per_swim <= 0 || !st->f_blobs ||
                            (st->num_fil > 0 && st->num_seg > 0 && !st->f_segs)))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_broyden_compute_errors: free bodies take the Broyden-for-everything form with f_segs and f_blobs");
    const bool nsegs = (int64_t)st->num_fil * st->num_seg > 0;
    if (!(st->dt > 0) || !(st->dl > 0) || !(st->tol > 0) || !st->body_x || !st->body_xm1 || !st->body_u ||
        !st->v_bodies || (st->bdf2 && (!st->body_xm2 || !st->body_um1)) ||
        (nsegs && (!st->seg_x || !st->seg_xm1 || !st->seg_u || !st->v_segs ||
                   (st->bdf2 && (!st->seg_xm2 || !st->seg_um1)))))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_compute_errors: missing state array or bad step");
    const int64_t nseg_tot = (int64_t)st->num_swim * st->num_fil * st->num_seg;
    const int64_t nb = st->num_swim;
    ResidualArgs a{};
    a.nswim = st->num_swim;
    a.nfil = st->num_fil;
    a.nseg = st->num_seg;
    a.bdf2 = st->bdf2 ? 1 : 0;
    a.dt = st->dt;
    a.pos_tol = st->dl * st->tol;
    a.lie_tol = st->tol;
    auto stage = [&](const double* src, int64_t count) -> const double* {
        if (!src || count == 0) return src;
        if (!host_mem(b)) return src;
        void* d = nullptr;
        if (hipMalloc(&d, count * 8) != hipSuccess) return nullptr;
        staged.push_back(d);
        if (hipMemcpyAsync(d, src, count * 8, hipMemcpyHostToDevice, b->stream) != hipSuccess) return nullptr;
        return (const double*)d;
    };
    a.x = stage(st->seg_x, 3 * nseg_tot);
    a.xm1 = stage(st->seg_xm1, 3 * nseg_tot);
    a.xm2 = a.bdf2 ? stage(st->seg_xm2, 3 * nseg_tot) : nullptr;
    a.u = stage(st->seg_u, 3 * nseg_tot);
    a.um1 = a.bdf2 ? stage(st->seg_um1, 3 * nseg_tot) : nullptr;
    a.v = stage(st->v_segs, 6 * nseg_tot);
    a.bx = stage(st->body_x, 3 * nb);
    a.bxm1 = stage(st->body_xm1, 3 * nb);
    a.bxm2 = a.bdf2 ? stage(st->body_xm2, 3 * nb) : nullptr;
    a.bu = stage(st->body_u, 3 * nb);
    a.bum1 = a.bdf2 ? stage(st->body_um1, 3 * nb) : nullptr;
    a.bv = stage(st->v_bodies, 6 * nb);
    const int64_t nblob_tot = (int64_t)st->num_swim * st->blobs_per_swim;
    a.nblob = st->blobs_per_swim;
    a.xb = stage(st->x_blobs, 3 * nblob_tot);
    a.vb = stage(st->v_blobs, 3 * nblob_tot);
    a.free_bfe = st->free_bodies ? 1 : 0;
    if (a.free_bfe) {
        a.fs = stage(st->f_segs, 6 * nseg_tot);
        a.fb = stage(st->f_blobs, 3 * nblob_tot);
    }
    int rc = TFS_OK;
    if (nblob_tot && (!a.xb || !a.vb)) rc = fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_compute_errors: staging");
    if (!rc && a.free_bfe && (!a.fb || (nseg_tot && !a.fs)))
        rc = fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_compute_errors: staging");
    if (!rc && ((nsegs && (!a.x || !a.xm1 || !a.u || !a.v || (a.bdf2 && (!a.xm2 || !a.um1)))) || !a.bx ||
                !a.bxm1 || !a.bu || !a.bv || (a.bdf2 && (!a.bxm2 || !a.bum1))))
        rc = fail(TFS_ERR_OUT_OF_MEMORY, "tfs_broyden_compute_errors: staging");
    const int blocks = (int)((nseg_tot + nb + nblob_tot + 255) / 256);
    if (!rc && 2LL * blocks > std::max<int64_t>((int64_t)b->G * (b->max_iter + 1), 2LL * b->RB))
        rc = fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_compute_errors: state larger than the engine");
    if (!rc && a.free_bfe) {
        hipLaunchKernelGGL(body_balance_kernel, dim3((unsigned)a.nswim), dim3(256), 0, b->stream, a, b->slot[slot]);
        rc = check(hipGetLastError(), "force and torque balance");
    }
    if (!rc) {
        hipLaunchKernelGGL(residual_kernel, dim3(blocks), dim3(256), 0, b->stream, a, b->slot[slot], b->partial);
        hipLaunchKernelGGL(residual_finish_kernel, dim3(1), dim3(256), 0, b->stream, b->partial, blocks,
                           b->scal + 2, b->book, b->book_max);
        rc = check(hipGetLastError(), "residual");
    }
    if (!rc && copy_flags)
        rc = check(hipMemcpyAsync(b->host_scal, b->scal + 2, 16, hipMemcpyDeviceToHost, b->stream),
                   "tfs_broyden_compute_errors");
    return rc;
}

// The flags errors_enqueue copied, once the stream has been waited for.
static int errors_result(tfs_broyden* b, int32_t* error_is_too_large) {
    if (error_is_too_large) *error_is_too_large = b->host_scal[0] > 1.0 ? 1 : 0;
    if (b->host_scal[1] != 0.0) return fail(TFS_ERR_NON_FINITE, "tfs_broyden_compute_errors: residual not finite");
    return TFS_OK;
}

extern "C" {

int tfs_broyden_compute_errors(tfs_broyden* b, const tfs_broyden_state* st, int32_t slot,
                               int32_t* error_is_too_large) {
    std::vector<void*> staged;
    int rc = errors_enqueue(b, st, slot, staged);
    hipError_t e = b ? hipStreamSynchronize(b->stream) : hipSuccess;
    for (void* p : staged) (void)hipFree(p);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "tfs_broyden_compute_errors");
    return errors_result(b, error_is_too_large);
}

}  // extern "C"

hipStream_t tfs::broyden_stream(const tfs_broyden* b) { return b->stream; }

// ---- Device-counted iterations: every launch has fixed arguments, so a
// caller can capture find_update and end_of_iter in hipGraphs ----

bool tfs::broyden_dev_supported(const tfs_broyden* b) { return !host_mem(b) && b->n <= kWaveDotsMax; }

int tfs::broyden_dev_begin(tfs_broyden* b) {
    hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(64), 0, b->stream, b->d_iter, (int32_t)b->iter);
    return check(hipGetLastError(), "device iteration count");
}

// inverse_update with the iteration count read on the device:
// out = -c J0^{-1} e - C(:, 0:k) (D(:, 0:k)^T e), k = min(*d_iter, max_iter - 1).
static int inverse_update_dev(tfs_broyden* b, const double* e, double* out) {
    int rc = apply_jinv(b, e, out, -b->desc.confidence);
    const int kmax = b->max_iter - 1;
    if (rc || kmax < 1) return rc;
    hipLaunchKernelGGL(lowrank_dots_kernel, dim3((unsigned)kmax), dim3(256), 0, b->stream, (const double*)b->D,
                       (int64_t)b->ld, kmax, e, b->n, b->h, (const int32_t*)b->d_iter);
    hipLaunchKernelGGL(lowrank_gemv_kernel, dim3((unsigned)((b->n + 255) / 256)), dim3(256), 0, b->stream,
                       (const double*)b->C, (int64_t)b->ld, kmax, (const double*)b->h, out, b->n,
                       (const int32_t*)b->d_iter);
    return check(hipGetLastError(), "low-rank update");
}

int tfs::broyden_dev_find_update(tfs_broyden* b, double* update) { return inverse_update_dev(b, b->slot[0], update); }

int tfs::broyden_dev_end_of_iter(tfs_broyden* b, int32_t* pending) {
    const int n = b->n;
    int rc = inverse_update_dev(b, b->slot[1], b->tC);
    if (rc) return rc;
    if (n <= kEndSmall) {
        hipLaunchKernelGGL(end_small_kernel, dim3(1), dim3(1024), 0, b->stream, (const double*)b->tC, b->tD, b->C,
                           b->D, (int64_t)b->ld, n, b->slot[0], (const double*)b->slot[1],
                           (const double*)(b->scal + 2), b->d_iter, 0, b->max_iter, pending);
        return check(hipGetLastError(), "end of iteration");
    }
    hipLaunchKernelGGL(diff_sumsq_kernel, dim3((unsigned)b->G), dim3(256), 0, b->stream, (const double*)b->slot[1],
                       (const double*)b->slot[0], b->tD, n, b->partial);
    hipLaunchKernelGGL(commit_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                       (const double*)b->tC, (const double*)b->tD, b->C, b->D, (int64_t)b->ld, n, b->slot[0],
                       (const double*)b->slot[1], (const double*)b->partial, b->G, (const double*)(b->scal + 2),
                       (const int32_t*)b->d_iter, b->max_iter, (const int32_t*)pending);
    hipLaunchKernelGGL(advance_count_kernel, dim3(1), dim3(64), 0, b->stream, b->d_iter, pending);
    return check(hipGetLastError(), "end of iteration");
}

void tfs::broyden_dev_finish(tfs_broyden* b, int32_t iter, int32_t latest) {
    b->iter = iter;
    b->latest = latest;
}

int tfs::broyden_errors_enqueue(tfs_broyden* b, const tfs_broyden_state* st, int32_t slot, bool copy_flags) {
    std::vector<void*> staged;  // device memory: nothing is staged
    int rc = errors_enqueue(b, st, slot, staged, copy_flags);
    for (void* p : staged) (void)hipFree(p);
    return rc;
}

int tfs::broyden_errors_result(tfs_broyden* b, int32_t* too_large) { return errors_result(b, too_large); }

const double* tfs::broyden_error_flags(const tfs_broyden* b) { return b->scal + 2; }

void tfs::broyden_set_bookkeeping(tfs_broyden* b, int32_t* flags, int max_iter) {
    b->book = flags;
    b->book_max = max_iter;
}

double tfs::broyden_last_measure(const tfs_broyden* b) { return b->host_scal[0]; }

int tfs::broyden_copy_error_flags(tfs_broyden* b) {
    return check(hipMemcpyAsync(b->host_scal, b->scal + 2, 16, hipMemcpyDeviceToHost, b->stream),
                 "residual flags");
}

int tfs::broyden_latest_norm(tfs_broyden* b, double* out) {
    int rc = check(k_norm(b->slot[b->latest], b->n, b->partial, b->G, b->scal, b->stream), "||residual||");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(b->host_scal, b->scal, 8, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "broyden_latest_norm");
    *out = *b->host_scal;
    return TFS_OK;
}

extern "C" {

int tfs_broyden_read_error(tfs_broyden* b, int32_t slot, double* out) {
    if (!b || !out || (slot != 0 && slot != 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_broyden_read_error: null argument or bad slot");
    hipError_t e = hipMemcpyAsync(out, b->slot[slot], (size_t)b->n * 8,
                                  host_mem(b) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess && (host_mem(b) || b->own_stream)) e = hipStreamSynchronize(b->stream);
    return check(e, "tfs_broyden_read_error");
}

}  // extern "C"
