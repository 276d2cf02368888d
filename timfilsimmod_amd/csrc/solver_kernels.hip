// solver_kernels.hip -- device kernels of the per-timestep solve (gfx950).
//
// * Krylov vector kernels for GMRES with FP64 Arnoldi: classical Gram-Schmidt
//   done twice (CGS2) as transposed/untransposed GEMVs over the basis V,
//   block reductions with wavefront shuffles, Givens rotations and the
//   triangular solve on one wave.
// * The saddle-point couplings of the prescribed-cilia system
//   (mobility_solver.cpp:1806-2090 system_matrix_mult_new, the K_phi / K_theta
//   terms) and its right preconditioner (mobility_solver.cpp:1040-1513
//   apply_preconditioner_new), one wave per filament.
// * Prescribed-beat kinematics (filament.cpp:680-983 Fourier fit of the
//   Fulford & Blake beat, filament.cpp:1083-1273 positions and phase/angle
//   velocity directions), the generalised-force right-hand side
//   (mobility_solver.cpp:240-280 cyclic cubic Hermite interpolation,
//   :282-440 assemble_rhs) and the explicit phase update (filament.cpp:1087-1097).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rpy_pair.h"
#include "solver_kernels.h"
#include "wave_ops.h"

namespace tfs {
namespace {

constexpr int kWave = 64;
constexpr int kMaxBasis = 352;   // columns of V an Arnoldi step can take (restart <= 350, plus one)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;  // valid in lane 0
}

// Sum over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return s;
}

// partial[c*G + b] = sum over the rows of block b of V[r, c] * w[r], c < ncols.
__global__ __launch_bounds__(256) void dot_cols_kernel(const double* __restrict__ V, int64_t ld, int ncols,
                                                      const double* __restrict__ w, int n,
                                                      double* __restrict__ partial) {
    __shared__ double sh[4];
    const int G = gridDim.x;
    const int r0 = blockIdx.x * 256 + threadIdx.x;
    for (int c = 0; c < ncols; ++c) {
        const double* v = V + (int64_t)c * ld;
        double s = 0.0;
        for (int r = r0; r < n; r += 256 * G) s += v[r] * w[r];
        s = block_sum256(s, sh);
        if (threadIdx.x == 0) partial[(int64_t)c * G + blockIdx.x] = s;
    }
}

// h[c] = sum_b partial[c*G + b]; H[c] += h[c] if H is given.  One block per column.
__global__ __launch_bounds__(256) void reduce_cols_kernel(const double* __restrict__ partial, int G,
                                                         double* __restrict__ h, double* __restrict__ Hcol) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) s += partial[(int64_t)c * G + b];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) {
        h[c] = s;
        if (Hcol) Hcol[c] += s;
    }
}

// w[r] -= sum_c V[r, c] h[c]  (alpha = -1) or out[r] = sum_c V[r,c] h[c] (alpha = 0).
__global__ __launch_bounds__(256) void gemv_kernel(const double* __restrict__ V, int64_t ld, int ncols,
                                                  const double* __restrict__ h, double* __restrict__ w, int n,
                                                  int subtract) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int c = 0; c < ncols; ++c) s += V[(int64_t)c * ld + r] * h[c];
    if (subtract) w[r] -= s;
    else w[r] = s;
}

// partial[b] = sum over block b of w[r]^2 (grid-stride).
__global__ __launch_bounds__(256) void sumsq_kernel(const double* __restrict__ w, int n,
                                                   double* __restrict__ partial) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += 256 * gridDim.x) s += w[r] * w[r];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] = sqrt(sum_b partial[b]); one block.
__global__ __launch_bounds__(256) void norm_finish_kernel(const double* __restrict__ partial, int G,
                                                         double* __restrict__ out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) s += partial[b];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) out[0] = __builtin_sqrt(s);
}

// One Arnoldi step's bookkeeping (restates the rotation logic of
// mobility_solver.cpp:2370-2419): H(j+1,j) = ||w||; apply the previous
// rotations to column j; form the new rotation; update g; publish |g(j+1)|
// and 1/||w|| for the basis normalisation.
__global__ void givens_kernel(double* __restrict__ H, int ldh, int j, double* __restrict__ cs,
                              double* __restrict__ sn, double* __restrict__ g,
                              const double* __restrict__ wnorm, double* __restrict__ scal) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* col = H + (int64_t)j * ldh;
    const double hn = wnorm[0];
    col[j + 1] = hn;
    scal[1] = hn > 0.0 ? 1.0 / hn : 0.0;   // scale of the next basis vector
    for (int i = 1; i <= j; ++i) {
        const double t = cs[i - 1] * col[i - 1] + sn[i - 1] * col[i];
        col[i] = cs[i - 1] * col[i] - sn[i - 1] * col[i - 1];
        col[i - 1] = t;
    }
    double c, s;
    const double a = col[j], b = col[j + 1];
    if (b == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (__builtin_fabs(b) > __builtin_fabs(a)) {
        const double t = a / b;
        s = (b > 0 ? 1.0 : -1.0) / __builtin_sqrt(1.0 + t * t);
        c = t * s;
    } else {
        const double t = b / a;
        c = (a > 0 ? 1.0 : -1.0) / __builtin_sqrt(1.0 + t * t);
        s = t * c;
    }
    cs[j] = c;
    sn[j] = s;
    col[j] = c * a + s * b;
    col[j + 1] = 0.0;
    g[j + 1] = -s * g[j];
    g[j] = c * g[j];
    scal[0] = __builtin_fabs(g[j + 1]);   // residual norm estimate
}

// v = w * scal[idx]
__global__ __launch_bounds__(256) void scale_kernel(double* __restrict__ v, const double* __restrict__ w,
                                                   const double* __restrict__ scal, int idx, int n) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) v[r] = w[r] * scal[idx];
}

// Back substitution H(0:k,0:k) y = g (mobility_solver.cpp:2421-2438).
// Back substitution H(0:k, 0:k) y = g for the upper-triangular H (column
// major, leading dimension ldh), by columns: once y_i is known, every earlier
// right-hand side loses H(j, i) y_i in parallel.  With STAGE the triangle is
// first copied into LDS (packed by columns, k(k+1)/2 entries) in one burst of
// loads, so the k dependent steps read no global memory.
template <bool STAGE>
__global__ __launch_bounds__(256) void trisolve_kernel(const double* __restrict__ H, int ldh,
                                                      const double* __restrict__ g, int k, double* __restrict__ y) {
    extern __shared__ double tri[];  // STAGE: k(k+1)/2 entries; then k right-hand sides
    __shared__ double yi;
    double* rhs = tri + (STAGE ? (int64_t)k * (k + 1) / 2 : 0);
    for (int j = threadIdx.x; j < k; j += 256) rhs[j] = g[j];
    if (STAGE) {
        // flat over the k x k square so that each thread's loads are independent
#pragma unroll 8
        for (int e = threadIdx.x; e < k * k; e += 256) {
            const int i = e / k, j = e - i * k;
            if (j <= i) tri[i * (i + 1) / 2 + j] = H[(int64_t)i * ldh + j];
        }
    }
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double* col = STAGE ? tri + (int64_t)i * (i + 1) / 2 : H + (int64_t)i * ldh;
        if (threadIdx.x == 0) {
            yi = rhs[i] / col[i];
            y[i] = yi;
        }
        __syncthreads();
        const double v = yi;
        for (int j = threadIdx.x; j < i; j += 256) rhs[j] -= col[j] * v;
        __syncthreads();
    }
}

// Warm start from the solution history, and the history shifted by one step:
// x holds x_{n-1}, h1 x_{n-2}, h2 x_{n-3}; afterwards x is the guess for x_n
// (order 0: x_{n-1}; 1: 2 x_{n-1} - x_{n-2}; 2: 3 x_{n-1} - 3 x_{n-2} + x_{n-3}),
// h1 = x_{n-1}, h2 = x_{n-2}.
__global__ __launch_bounds__(256) void extrapolate_kernel(double* __restrict__ x, double* __restrict__ h1,
                                                          double* __restrict__ h2, int n, int order) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = h1[i], c = h2[i];
    x[i] = order == 2 ? 3.0 * (a - b) + c : (order == 1 ? 2.0 * a - b : a);
    h2[i] = b;
    h1[i] = a;
}

// y = a*x + b*y
__global__ __launch_bounds__(256) void axpby_kernel(double a, const double* __restrict__ x, double b,
                                                   double* __restrict__ y, int n) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) y[r] = b == 0.0 ? a * x[r] : a * x[r] + b * y[r];
}

// out = x - y
__global__ __launch_bounds__(256) void sub_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                 double* __restrict__ out, int n) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) out[r] = x[r] - y[r];
}

// ---------------------------------------------------------------------------
// Arnoldi step (CGS2) in four launches.  Workgroup b owns the row slices
// b, b + G, b + 2G, ... of kSlice rows.  In a dot phase its four waves take
// the basis columns round robin and every lane sums its rows, so the column
// reductions need no workgroup barrier.  A launch that needs the previous
// launch's coefficients re-reduces that launch's per-workgroup partial sums
// itself -- every workgroup in the same order, so all of them see identical
// values -- instead of synchronising across workgroups.
//   dots:          p1 = partial V^T w
//   update_dots:   h1 = sum p1; w -= V h1; p2 = partial V^T w
//   update_sumsq:  h2 = sum p2; w -= V h2; p3 = partial ||w||^2
//   close:         ||w||; Givens; 1/||w||; est[j]
// Workgroup 0 writes H(:, j) = h1 + h2.  All sums run in a fixed order: the
// step is deterministic.
// ---------------------------------------------------------------------------
constexpr int kSlice = 512;

struct ArnoldiArgs {
    const double* V;     // basis, column-major
    int64_t ld;
    int ncols;           // j + 1
    double* w;
    int n;
    const int32_t* skip; // stop flag: nonzero -> the launch does nothing
};

// partial[c*G + b] = sum over the rows of workgroup b of V[r, c] * w[r].
// Each wave takes four columns at a time (kSlice/64 rows x 4 columns of loads
// in flight per lane), with the lane's rows of w held in registers.
__device__ __forceinline__ void slice_dots(const ArnoldiArgs& a, double* __restrict__ partial, int b, int G) {
    constexpr int Q = kSlice / kWave;
    const double* __restrict__ V = a.V;
    const double* __restrict__ w = a.w;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int64_t last = a.n - 1;
    for (int c0 = 4 * wid; c0 < a.ncols; c0 += 16) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        const int nc = a.ncols - c0 < 4 ? a.ncols - c0 : 4;
        for (int64_t r0 = (int64_t)b * kSlice; r0 < a.n; r0 += (int64_t)G * kSlice) {
            // Branch-free tail: rows past n read row n-1 and weigh it by zero.
            // All loads of the slice are issued before the first product.
            int64_t idx[Q];
            double wv[Q], vv[4][Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int64_t r = r0 + lane + q * kWave;
                idx[q] = r < a.n ? r : last;
                wv[q] = w[idx[q]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double* v = V + (int64_t)(c0 + (u < nc ? u : 0)) * a.ld;
#pragma unroll
                for (int q = 0; q < Q; ++q) vv[u][q] = v[idx[q]];
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const double wq = r0 + lane + q * kWave < a.n ? wv[q] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u) s[u] += vv[u][q] * wq;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double t = wave_sum(s[u]);
            if (lane == 0 && u < nc) partial[(int64_t)(c0 + u) * G + b] = t;
        }
    }
}

// Sum of p[0 .. G) by one wave (lanes stride over the entries, then a
// shuffle tree): valid in lane 0.
__device__ __forceinline__ double wave_sum_strided(const double* p, int G) {
    const int lane = threadIdx.x & (kWave - 1);
    double s = 0.0;
    for (int b = lane; b < G; b += kWave) s += p[b];
    return wave_sum(s);
}

// hs[c] = sum_b partial[c*G + b] into LDS; workgroup 0 also adds it to Hcol.
// Waves take four columns at a time; lanes stride over b with every load of a
// column group issued before the sums (entries past G are weighed by zero).
// Ends with a barrier.
__device__ __forceinline__ void gather_coeffs(const double* __restrict__ partial, int ncols, int G, double* hs,
                                              double* __restrict__ Hcol, bool first) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    for (int c0 = 4 * wid; c0 < ncols; c0 += 16) {
        const int nc = ncols - c0 < 4 ? ncols - c0 : 4;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b0 = 0; b0 < G; b0 += 4 * kWave) {
            double pv[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int bb = b0 + lane + q * kWave;
                    pv[u][q] = partial[(int64_t)(c0 + (u < nc ? u : 0)) * G + (bb < G ? bb : G - 1)];
                }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in = b0 + lane + q * kWave < G;
#pragma unroll
                for (int u = 0; u < 4; ++u) s[u] += in ? pv[u][q] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double t = wave_sum(s[u]);
            if (lane == 0 && u < nc) {
                hs[c0 + u] = t;
                if (blockIdx.x == 0) Hcol[c0 + u] = first ? t : Hcol[c0 + u] + t;
            }
        }
    }
    __syncthreads();
}

// w[r] -= sum_c V[r, c] h[c] over the rows of workgroup b (columns eight at
// a time, their loads issued together).
__device__ __forceinline__ void slice_update(const ArnoldiArgs& a, const double* h, int b, int G) {
    constexpr int Q = kSlice / 256;
    const double* __restrict__ V = a.V;
    double* __restrict__ w = a.w;
    for (int64_t r0 = (int64_t)b * kSlice + threadIdx.x; r0 < a.n; r0 += (int64_t)G * kSlice) {
        double s[Q];
        int64_t idx[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            s[q] = 0.0;
            idx[q] = r0 + q * 256 < a.n ? r0 + q * 256 : a.n - 1;
        }
        int c = 0;
        for (; c + 8 <= a.ncols; c += 8) {
            double vv[8][Q];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) vv[u][q] = V[(int64_t)(c + u) * a.ld + idx[q]];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] += vv[u][q] * h[c + u];
        }
        for (; c + 4 <= a.ncols; c += 4) {
            double vv[4][Q];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) vv[u][q] = V[(int64_t)(c + u) * a.ld + idx[q]];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) s[q] += vv[u][q] * h[c + u];
        }
        for (; c < a.ncols; ++c)
#pragma unroll
            for (int q = 0; q < Q; ++q) s[q] += V[(int64_t)c * a.ld + idx[q]] * h[c];
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (r0 + q * 256 < a.n) w[r0 + q * 256] -= s[q];
    }
}

__global__ __launch_bounds__(256) void arnoldi_dots_kernel(ArnoldiArgs a, double* __restrict__ p_out) {
    if (*a.skip) return;
    slice_dots(a, p_out, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void arnoldi_update_dots_kernel(ArnoldiArgs a, const double* __restrict__ p_in,
                                                                  double* __restrict__ p_out,
                                                                  double* __restrict__ Hcol) {
    __shared__ double hs[kMaxBasis];
    if (*a.skip) return;
    gather_coeffs(p_in, a.ncols, gridDim.x, hs, Hcol, true);
    slice_update(a, hs, blockIdx.x, gridDim.x);
    __syncthreads();  // the slice's rows of w are final before the dots read them
    slice_dots(a, p_out, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void arnoldi_update_sumsq_kernel(ArnoldiArgs a, const double* __restrict__ p_in,
                                                                   double* __restrict__ p_out,
                                                                   double* __restrict__ Hcol, bool first) {
    __shared__ double hs[kMaxBasis];
    __shared__ double sh[4];
    if (*a.skip) return;
    gather_coeffs(p_in, a.ncols, gridDim.x, hs, Hcol, first);
    slice_update(a, hs, blockIdx.x, gridDim.x);
    __syncthreads();
    double s = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kSlice + threadIdx.x; r < a.n; r += (int64_t)gridDim.x * kSlice) {
#pragma unroll
        for (int q = 0; q < kSlice / 256; ++q) {
            const int64_t rr = r + q * 256;
            if (rr < a.n) s += a.w[rr] * a.w[rr];
        }
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) p_out[blockIdx.x] = s;
}

// Givens bookkeeping of step j given ||w|| = hn (givens_kernel's rotations):
// the column of H and the previous rotations are staged in LDS so that
// thread 0's sequential rotations read no memory; est[j] = |g(j+1)|,
// scal[1] = 1/||w||.  The first step whose estimate meets abs_tol sets
// *stop = j + 1 (a non-finite estimate sets -(j + 1)); later launches of the
// solve then skip.  The value also goes to *stop_mirror and est[j] to
// *est_mirror if given (host memory the host reads).  Called by every thread
// of one workgroup.
__device__ void arnoldi_close(double hn, double* __restrict__ H, int ldh, int j, double* __restrict__ cs,
                              double* __restrict__ sn, double* __restrict__ g, double* __restrict__ scal,
                              double* __restrict__ est, int32_t* __restrict__ stop, double abs_tol,
                              int32_t* __restrict__ stop_mirror = nullptr, double* __restrict__ est_mirror = nullptr) {
    __shared__ double col[kMaxBasis + 1], lc[kMaxBasis], ls[kMaxBasis];
    double* Hc = H + (int64_t)j * ldh;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) {
        col[i] = Hc[i];
        lc[i] = cs[i];
        ls[i] = sn[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        col[j + 1] = hn;
        scal[1] = hn > 0.0 ? 1.0 / hn : 0.0;
        for (int i = 1; i <= j; ++i) {
            const double u = lc[i - 1] * col[i - 1] + ls[i - 1] * col[i];
            col[i] = lc[i - 1] * col[i] - ls[i - 1] * col[i - 1];
            col[i - 1] = u;
        }
        double c, sv;
        const double x = col[j], y = col[j + 1];
        if (y == 0.0) {
            c = 1.0;
            sv = 0.0;
        } else if (__builtin_fabs(y) > __builtin_fabs(x)) {
            const double q = x / y;
            sv = (y > 0 ? 1.0 : -1.0) / __builtin_sqrt(1.0 + q * q);
            c = q * sv;
        } else {
            const double q = y / x;
            c = (x > 0 ? 1.0 : -1.0) / __builtin_sqrt(1.0 + q * q);
            sv = q * c;
        }
        cs[j] = c;
        sn[j] = sv;
        col[j] = c * x + sv * y;
        col[j + 1] = 0.0;
        const double gj = g[j];
        g[j + 1] = -sv * gj;
        g[j] = c * gj;
        scal[0] = __builtin_fabs(g[j + 1]);
        est[j] = scal[0];
        if (est_mirror)
            __hip_atomic_store(reinterpret_cast<unsigned long long*>(est_mirror),
                               (unsigned long long)__double_as_longlong(scal[0]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        const int32_t sv_stop = !__builtin_isfinite(scal[0]) ? -(j + 1) : (scal[0] <= abs_tol ? j + 1 : 0);
        if (sv_stop) {
            *stop = sv_stop;
            if (stop_mirror) __hip_atomic_store(stop_mirror, sv_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= j + 1; i += blockDim.x) Hc[i] = col[i];
}

// The next basis vector and its preconditioned image in one pass:
// v = w * sc, out = P^{-1} v, with the filament-wave layout of precond_kernel.
__device__ void scale_precond_body(const double* __restrict__ w, double sc, double* __restrict__ v,
                                   double* __restrict__ out, const double* __restrict__ kphi,
                                   const double* __restrict__ kth, const SaddleDims& d, double gs, double gb) {
    const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int m = gw; m < d.nfil; m += nwaves) {
        const int len = 3 * d.nseg;
        const int64_t seg0 = (int64_t)m * len;
        double dp = 0.0, dt = 0.0, c11 = 0.0, c22 = 0.0, c12 = 0.0;
        for (int k = lane; k < len; k += kWave) {
            const double kp = kphi[seg0 + k], kt = kth[seg0 + k];
            const double s = gs * (w[seg0 + k] * sc);
            dp += kp * s;
            dt += kt * s;
            c11 += kp * kp;
            c22 += kt * kt;
            c12 += kp * kt;
        }
        dp = wave_sum(dp);
        dt = wave_sum(dt);
        c11 = wave_sum(c11);
        c22 = wave_sum(c22);
        c12 = wave_sum(c12);
        double p = 0.0, q = 0.0;
        if (lane == 0) {
            const double inp = d.dyn_phase ? w[d.off_phase + m] * sc : 0.0;
            const double ina = d.dyn_angle ? w[d.off_angle + m] * sc : 0.0;
            if (d.dyn_phase && d.dyn_angle) {
                const double r1 = -inp - dp;
                const double r2 = -ina - dt;
                const double det = c11 * c22 - c12 * c12;
                p = (c22 * r1 - c12 * r2) / det / gs;
                q = (-c12 * r1 + c11 * r2) / det / gs;
            } else if (d.dyn_phase) {
                p = (-inp - dp) / (c11 * gs);
            } else if (d.dyn_angle) {
                q = (-ina - dt) / (c22 * gs);
            }
            if (d.dyn_phase) {
                v[d.off_phase + m] = inp;
                out[d.off_phase + m] = p;
            }
            if (d.dyn_angle) {
                v[d.off_angle + m] = ina;
                out[d.off_angle + m] = q;
            }
        }
        p = __shfl(p, 0, kWave);
        q = __shfl(q, 0, kWave);
        for (int k = lane; k < len; k += kWave) {
            const double in = w[seg0 + k] * sc;
            v[seg0 + k] = in;
            out[seg0 + k] = gs * (in + p * kphi[seg0 + k] + q * kth[seg0 + k]);
        }
    }
    const int nb3 = 3 * d.nblob;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb3; i += gridDim.x * blockDim.x) {
        const double in = w[d.off_blob + i] * sc;
        v[d.off_blob + i] = in;
        out[d.off_blob + i] = gb * in;
    }
}

// y_lane = sum_c B[c][lane] x_c for a symmetric n x n block (n <= 64), row
// `lane` of the product on each lane (0 past n); x_c is lane c's x.  The
// column loads go out sixteen at a time, so a wave keeps many in flight.
__device__ __forceinline__ double small_symv(const double* __restrict__ B, int n, int lane, double x) {
    const int r = lane < n ? lane : n - 1;
    double y = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < kWave; c0 += 16) {
        if (c0 < n) {
            double b[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) b[u] = c0 + u < n ? B[(int64_t)(c0 + u) * n + r] : 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u)
                if (c0 + u < n) y += b[u] * lane_bcast(x, c0 + u);
        }
    }
    return lane < n ? y : 0.0;
}

// One filament (segment rows and its phase / angle rows) or one blob patch of
// the block preconditioner, by one wave: v = w * sc (if v), out = P^{-1} v.
// Rows are spread over the lanes (row r on lane r % 64, register r / 64); the
// dense block products take their operands from the owning lanes by shuffle.
__device__ void block_precond_wave(const PrecondArgs& P, const SaddleDims& d, int item, int lane,
                                   const double* __restrict__ w, double sc, double* __restrict__ v,
                                   double* __restrict__ out) {
    constexpr int RMAX = 3 * kMaxBlockFilSeg / kWave;   // registers per lane
    if (item < d.nfil && 3 * d.nseg <= kWave) {
        // one row per lane
        const int m = item;
        const int len = 3 * d.nseg;
        const int64_t seg0 = (int64_t)m * len;
        const bool act = lane < len;
        const double x = act ? w[seg0 + lane] * sc : 0.0;
        if (v && act) v[seg0 + lane] = x;
        const double du = small_symv(P.fil_inv + (int64_t)m * len * len, len, lane, x);
        double tp = wave_sum(act ? P.kphi[seg0 + lane] * du : 0.0);
        double tt = wave_sum(act ? P.kth[seg0 + lane] * du : 0.0);
        double yp = 0.0, ya = 0.0;
        if (lane == 0) {
            const double* Si = P.fil_sinv + 4 * (int64_t)m;
            const double vp = d.dyn_phase ? w[d.off_phase + m] * sc : 0.0;
            const double va = d.dyn_angle ? w[d.off_angle + m] * sc : 0.0;
            if (d.dyn_phase && d.dyn_angle) {
                yp = -(Si[0] * (vp + tp) + Si[1] * (va + tt));
                ya = -(Si[2] * (vp + tp) + Si[3] * (va + tt));
            } else if (d.dyn_phase) {
                yp = -Si[0] * (vp + tp);
            } else if (d.dyn_angle) {
                ya = -Si[0] * (va + tt);
            }
            if (d.dyn_phase) {
                if (v) v[d.off_phase + m] = vp;
                out[d.off_phase + m] = yp;
            }
            if (d.dyn_angle) {
                if (v) v[d.off_angle + m] = va;
                out[d.off_angle + m] = ya;
            }
        }
        yp = lane_bcast(yp, 0);
        ya = lane_bcast(ya, 0);
        if (act) {
            const double* dkp = P.fil_dk + (int64_t)m * 2 * len;
            out[seg0 + lane] = du + dkp[lane] * yp + dkp[len + lane] * ya;
        }
        return;
    }
    if (item < d.nfil) {
        const int m = item;
        const int len = 3 * d.nseg;
        const int64_t seg0 = (int64_t)m * len;
        const double* Bi = P.fil_inv + (int64_t)m * len * len;
        double x[RMAX], du[RMAX];
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int r = q * kWave + lane;
            x[q] = r < len ? w[seg0 + r] * sc : 0.0;
            if (v && r < len) v[seg0 + r] = x[q];
            du[q] = 0.0;
        }
        // du = D_m^{-1} x (symmetric: entry (r, c) read as (c, r), coalesced over r)
#pragma unroll
        for (int qc = 0; qc < RMAX; ++qc) {
            if (qc * kWave >= len) break;
            for (int l = 0; l < kWave; ++l) {
                const int c = qc * kWave + l;
                if (c >= len) break;
                const double xc = __shfl(x[qc], l, kWave);
#pragma unroll
                for (int q = 0; q < RMAX; ++q) {
                    const int r = q * kWave + lane;
                    if (r < len) du[q] += Bi[(int64_t)c * len + r] * xc;
                }
            }
        }
        double tp = 0.0, tt = 0.0;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int r = q * kWave + lane;
            if (r < len) {
                tp += P.kphi[seg0 + r] * du[q];
                tt += P.kth[seg0 + r] * du[q];
            }
        }
        tp = wave_sum(tp);
        tt = wave_sum(tt);
        double yp = 0.0, ya = 0.0;
        if (lane == 0) {
            const double* Si = P.fil_sinv + 4 * (int64_t)m;
            const double vp = d.dyn_phase ? w[d.off_phase + m] * sc : 0.0;
            const double va = d.dyn_angle ? w[d.off_angle + m] * sc : 0.0;
            // y = -S^{-1} (v + K^T D^{-1} x)
            if (d.dyn_phase && d.dyn_angle) {
                yp = -(Si[0] * (vp + tp) + Si[1] * (va + tt));
                ya = -(Si[2] * (vp + tp) + Si[3] * (va + tt));
            } else if (d.dyn_phase) {
                yp = -Si[0] * (vp + tp);
            } else if (d.dyn_angle) {
                ya = -Si[0] * (va + tt);
            }
            if (d.dyn_phase) {
                if (v) v[d.off_phase + m] = vp;
                out[d.off_phase + m] = yp;
            }
            if (d.dyn_angle) {
                if (v) v[d.off_angle + m] = va;
                out[d.off_angle + m] = ya;
            }
        }
        yp = __shfl(yp, 0, kWave);
        ya = __shfl(ya, 0, kWave);
        // x = D^{-1} (x + K y) = du + (D^{-1} K_phi) yp + (D^{-1} K_theta) ya
        const double* dkp = P.fil_dk + (int64_t)m * 2 * len;
        const double* dka = dkp + len;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int r = q * kWave + lane;
            if (r < len) out[seg0 + r] = du[q] + dkp[r] * yp + dka[r] * ya;
        }
        return;
    }
    // blob patch: at most kPatchBlobs blobs, 3 * kPatchBlobs <= 64 rows
    const int p = item - d.nfil;
    const int b0 = P.patch_start[p];
    const int rows = 3 * (P.patch_start[p + 1] - b0);
    const double* Bi = P.patch_inv + P.patch_off[p];
    int64_t idx = 0;
    double x = 0.0;
    if (lane < rows) {
        idx = d.off_blob + 3 * (int64_t)P.patch_blob[b0 + lane / 3] + lane % 3;
        x = w[idx] * sc;
        if (v) v[idx] = x;
    }
    const double o = small_symv(Bi, rows, lane, x);
    if (lane < rows) out[idx] = o;
}

// Every wave of the grid takes filaments, then patches, round robin.
__device__ void block_precond_body(const PrecondArgs& P, const SaddleDims& d, const double* __restrict__ w,
                                   double sc, double* __restrict__ v, double* __restrict__ out) {
    const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = gridDim.x * blockDim.x / kWave;
    const int items = d.nfil + P.npatch;
    for (int it = gw; it < items; it += nwaves) block_precond_wave(P, d, it, lane, w, sc, v, out);
}

// scale = *dev_scale[1], or host_scale when dev_scale is null.
__global__ __launch_bounds__(256) void scale_precond_kernel(const double* __restrict__ w,
                                                           const double* __restrict__ dev_scale,
                                                           double host_scale, double* __restrict__ v,
                                                           double* __restrict__ out, PrecondArgs P, SaddleDims d,
                                                           const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const double sc = dev_scale ? dev_scale[1] : host_scale;
    if (P.kind == 1) block_precond_body(P, d, w, sc, v, out);
    else scale_precond_body(w, sc, v, out, P.kphi, P.kth, d, P.gs, P.gb);
}


// Closes Arnoldi step j and opens step j + 1 in one launch.  Every workgroup
// sums the G partial sums of ||w||^2 in the same order (so all of them hold
// the same 1/||w||); workgroup 0 does the Givens bookkeeping of step j; all
// of them write v_{j+1} = w / ||w|| and z = P^{-1} v_{j+1} (skipped when v is
// null: the close of the last step of a cycle).  Workgroups that start after
// step j met the tolerance may still write v_{j+1} and z: neither is read
// after convergence.
__global__ __launch_bounds__(256) void arnoldi_close_next_kernel(
    const double* __restrict__ p_in, int G, double* __restrict__ H, int ldh, int j, double* __restrict__ cs,
    double* __restrict__ sn, double* __restrict__ g, double* __restrict__ scal, double* __restrict__ est,
    int32_t* __restrict__ stop, int32_t* __restrict__ stop_mirror, double* __restrict__ est_mirror, double abs_tol,
    const double* __restrict__ w, double* __restrict__ v, double* __restrict__ out, PrecondArgs P, SaddleDims d) {
    __shared__ double nrm;
    if (*stop) return;
    if (threadIdx.x < kWave) {
        const double t = wave_sum_strided(p_in, G);
        if (threadIdx.x == 0) nrm = __builtin_sqrt(t);
    }
    __syncthreads();
    const double hn = nrm;
    if (blockIdx.x == 0) arnoldi_close(hn, H, ldh, j, cs, sn, g, scal, est, stop, abs_tol, stop_mirror, est_mirror);
    if (v) {
        const double sc = hn > 0.0 ? 1.0 / hn : 0.0;
        if (P.kind == 1) block_precond_body(P, d, w, sc, v, out);
        else scale_precond_body(w, sc, v, out, P.kphi, P.kth, d, P.gs, P.gb);
    }
}

// Right preconditioner P^{-1}: the exact inverse of [D -K; -K^T 0] with
// D = I/(6 pi mu a) (mobility_solver.cpp:1040-1116).  Blob rows are scaled by
// 6 pi mu a_b.  Filament waves first, then a grid-stride over blob entries.
__global__ __launch_bounds__(256) void precond_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                     PrecondArgs P, SaddleDims d) {
    if (P.kind == 1) {
        block_precond_body(P, d, in, 1.0, nullptr, out);
        return;
    }
    const double* __restrict__ kphi = P.kphi;
    const double* __restrict__ kth = P.kth;
    const double gs = P.gs, gb = P.gb;
    const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int m = gw; m < d.nfil; m += nwaves) {
        const int len = 3 * d.nseg;
        const int64_t seg0 = (int64_t)m * len;
        double dp = 0.0, dt = 0.0, c11 = 0.0, c22 = 0.0, c12 = 0.0;
        for (int k = lane; k < len; k += kWave) {
            const double kp = kphi[seg0 + k], kt = kth[seg0 + k];
            const double s = gs * in[seg0 + k];
            dp += kp * s;
            dt += kt * s;
            c11 += kp * kp;
            c22 += kt * kt;
            c12 += kp * kt;
        }
        dp = wave_sum(dp);
        dt = wave_sum(dt);
        c11 = wave_sum(c11);
        c22 = wave_sum(c22);
        c12 = wave_sum(c12);
        double p = 0.0, q = 0.0;
        if (lane == 0) {
            if (d.dyn_phase && d.dyn_angle) {
                const double r1 = -in[d.off_phase + m] - dp;
                const double r2 = -in[d.off_angle + m] - dt;
                const double det = c11 * c22 - c12 * c12;
                p = (c22 * r1 - c12 * r2) / det / gs;
                q = (-c12 * r1 + c11 * r2) / det / gs;
            } else if (d.dyn_phase) {
                p = (-in[d.off_phase + m] - dp) / (c11 * gs);
            } else if (d.dyn_angle) {
                q = (-in[d.off_angle + m] - dt) / (c22 * gs);
            }
            if (d.dyn_phase) out[d.off_phase + m] = p;
            if (d.dyn_angle) out[d.off_angle + m] = q;
        }
        p = __shfl(p, 0, kWave);
        q = __shfl(q, 0, kWave);
        for (int k = lane; k < len; k += kWave)
            out[seg0 + k] = gs * (in[seg0 + k] + p * kphi[seg0 + k] + q * kth[seg0 + k]);
    }
    const int nb3 = 3 * d.nblob;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb3; i += gridDim.x * blockDim.x)
        out[d.off_blob + i] = gb * in[d.off_blob + i];
}

// Block preconditioner set-up: the 3 x 3 RPY blocks (the reference's pair
// tensor, every branch, in FP64) of the pairs inside one filament or one blob
// patch, row-major; one workgroup per filament / patch.
__device__ void rpy_block_entries(const double* xi, const double* xj, double a, double mu, double* blk3x3) {
    const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
    for (int c = 0; c < 3; ++c) {
        double f[3] = {0.0, 0.0, 0.0}, u[3] = {0.0, 0.0, 0.0};
        f[c] = 1.0;
        pair_near<3, 3>(u, dx, dy, dz, f, a, a, mu);
        for (int r = 0; r < 3; ++r) blk3x3[3 * r + c] = u[r];
    }
}

__global__ __launch_bounds__(256) void fil_blocks_kernel(const double* __restrict__ x_segs, int nseg, double a,
                                                        double mu, double* __restrict__ blk) {
    const int m = blockIdx.x;
    const int len = 3 * nseg;
    const double* x = x_segs + (int64_t)m * len;
    double* B = blk + (int64_t)m * len * len;
    for (int e = threadIdx.x; e < nseg * nseg; e += blockDim.x) {
        const int i = e / nseg, j = e % nseg;
        double t[9];
        rpy_block_entries(x + 3 * i, x + 3 * j, a, mu, t);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[(int64_t)(3 * i + r) * len + 3 * j + c] = t[3 * r + c];
    }
}

__global__ __launch_bounds__(256) void patch_blocks_kernel(const double* __restrict__ x_blobs,
                                                          const int32_t* __restrict__ patch_blob,
                                                          const int32_t* __restrict__ patch_start,
                                                          const int64_t* __restrict__ patch_off, double a,
                                                          double mu, double* __restrict__ blk) {
    const int p = blockIdx.x;
    const int b0 = patch_start[p], cnt = patch_start[p + 1] - b0;
    const int rows = 3 * cnt;
    double* B = blk + patch_off[p];
    for (int e = threadIdx.x; e < cnt * cnt; e += blockDim.x) {
        const int i = e / cnt, j = e % cnt;
        double t[9];
        rpy_block_entries(x_blobs + 3 * (int64_t)patch_blob[b0 + i], x_blobs + 3 * (int64_t)patch_blob[b0 + j], a,
                          mu, t);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) B[(int64_t)(3 * i + r) * rows + 3 * j + c] = t[3 * r + c];
    }
}

// The inverse of the Schur complement [s11 s12; s21 s22] of the dynamic
// couplings (2 x 2, or 1 x 1 in S[0]).
__device__ void write_sinv(double* S, const SaddleDims& d, double s11, double s12, double s21, double s22) {
    if (d.dyn_phase && d.dyn_angle) {
        const double det = s11 * s22 - s12 * s21;
        S[0] = s22 / det;
        S[1] = -s12 / det;
        S[2] = -s21 / det;
        S[3] = s11 / det;
    } else {
        S[0] = d.dyn_phase ? 1.0 / s11 : (d.dyn_angle ? 1.0 / s22 : 0.0);
        S[1] = S[2] = S[3] = 0.0;
    }
}

// fil_dk = D_m^{-1} [K_phi K_theta] and fil_sinv = (K^T D_m^{-1} K)^{-1} (the
// dynamic couplings only), one wave per filament.
__global__ __launch_bounds__(256) void fil_schur_kernel(PrecondArgs P, SaddleDims d, double* __restrict__ dk,
                                                       double* __restrict__ sinv) {
    constexpr int RMAX = 3 * kMaxBlockFilSeg / kWave;
    const int m = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (m >= d.nfil) return;
    const int len = 3 * d.nseg;
    const int64_t seg0 = (int64_t)m * len;
    const double* Bi = P.fil_inv + (int64_t)m * len * len;
    if (len <= kWave) {
        const bool act = lane < len;
        const double kp1 = act ? P.kphi[seg0 + lane] : 0.0, kt1 = act ? P.kth[seg0 + lane] : 0.0;
        const double up1 = small_symv(Bi, len, lane, kp1), ut1 = small_symv(Bi, len, lane, kt1);
        if (act) {
            dk[(int64_t)m * 2 * len + lane] = up1;
            dk[(int64_t)m * 2 * len + len + lane] = ut1;
        }
        const double s11 = wave_sum(kp1 * up1), s12 = wave_sum(kp1 * ut1);
        const double s21 = wave_sum(kt1 * up1), s22 = wave_sum(kt1 * ut1);
        if (lane == 0) write_sinv(sinv + 4 * (int64_t)m, d, s11, s12, s21, s22);
        return;
    }
    double kp[RMAX], kt[RMAX], up[RMAX], ut[RMAX];
#pragma unroll
    for (int q = 0; q < RMAX; ++q) {
        const int r = q * kWave + lane;
        kp[q] = r < len ? P.kphi[seg0 + r] : 0.0;
        kt[q] = r < len ? P.kth[seg0 + r] : 0.0;
        up[q] = ut[q] = 0.0;
    }
#pragma unroll
    for (int qc = 0; qc < RMAX; ++qc) {
        if (qc * kWave >= len) break;
        for (int l = 0; l < kWave; ++l) {
            const int c = qc * kWave + l;
            if (c >= len) break;
            const double pc = __shfl(kp[qc], l, kWave), tc = __shfl(kt[qc], l, kWave);
#pragma unroll
            for (int q = 0; q < RMAX; ++q) {
                const int r = q * kWave + lane;
                if (r < len) {
                    const double bv = Bi[(int64_t)c * len + r];
                    up[q] += bv * pc;
                    ut[q] += bv * tc;
                }
            }
        }
    }
    double s11 = 0.0, s12 = 0.0, s21 = 0.0, s22 = 0.0;
#pragma unroll
    for (int q = 0; q < RMAX; ++q) {
        const int r = q * kWave + lane;
        if (r < len) {
            dk[(int64_t)m * 2 * len + r] = up[q];
            dk[(int64_t)m * 2 * len + len + r] = ut[q];
            s11 += kp[q] * up[q];
            s12 += kp[q] * ut[q];
            s21 += kt[q] * up[q];
            s22 += kt[q] * ut[q];
        }
    }
    s11 = wave_sum(s11);
    s12 = wave_sum(s12);
    s21 = wave_sum(s21);
    s22 = wave_sum(s22);
    if (lane == 0) write_sinv(sinv + 4 * (int64_t)m, d, s11, s12, s21, s22);
}

// ---------------------------------------------------------------------------
// Prescribed beat.
// ---------------------------------------------------------------------------
__constant__ double c_beat_A[2][4][3];  // [x|y][mode 0..3][power s, s^2, s^3]
__constant__ double c_beat_B[2][3][3];  // [x|y][mode 1..3][power]

// Local beat-plane shape (x, y) and its phase derivative at arclength
// parameter s (filament.cpp:770-833).
__device__ void beat_shape(double s, double phase, double* px, double* py, double* dx, double* dy) {
    const double sv[3] = {s, s * s, s * s * s};
    double cs[4], sn[4];
    cs[0] = 1.0;
    sn[0] = 0.0;
    for (int n = 1; n < 4; ++n) {
        cs[n] = cos(n * phase);
        sn[n] = sin(n * phase);
    }
    double r[2] = {0, 0}, v[2] = {0, 0};
    for (int ax = 0; ax < 2; ++ax) {
        for (int p = 0; p < 3; ++p) {
            double a = 0.0, da = 0.0;
            for (int n = 0; n < 4; ++n) {
                a += cs[n] * c_beat_A[ax][n][p];
                da += -n * sn[n] * c_beat_A[ax][n][p];
            }
            for (int n = 1; n < 4; ++n) {
                a += sn[n] * c_beat_B[ax][n - 1][p];
                da += n * cs[n] * c_beat_B[ax][n - 1][p];
            }
            r[ax] += a * sv[p];
            v[ax] += da * sv[p];
        }
    }
    *px = r[0];
    *py = r[1];
    *dx = v[0];
    *dy = v[1];
}

// One thread per segment: positions (into x_segs, 3 per segment), K_phi and
// K_theta columns (3 per segment).
__global__ __launch_bounds__(256) void beat_kernel(BeatArgs a, double* __restrict__ x_segs,
                                                  double* __restrict__ kphi, double* __restrict__ kth) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nfil * a.nseg) return;
    const int m = t / a.nseg, n = t - m * a.nseg;
    const double* fr = a.frames + 9 * m;  // e1 (normal), e2 (beat direction), e3
    const double* x0 = a.bases + 3 * m;
    double pos[3] = {x0[0], x0[1], x0[2]}, kp[3] = {0, 0, 0}, kt[3] = {0, 0, 0};
    if (n > 0) {
        const double s = a.s_values[n];
        double px, py, dx, dy;
        beat_shape(s, a.phase[m], &px, &py, &dx, &dy);
        const double th = a.angle ? a.angle[m] : 0.0;
        const double ct = cos(th), st = sin(th);
        // shape rotated about the local z axis by theta, scaled by the length
        const double rx = a.length * (ct * px - st * py), ry = a.length * (st * px + ct * py);
        const double vx = a.length * (ct * dx - st * dy), vy = a.length * (st * dx + ct * dy);
        for (int c = 0; c < 3; ++c) {
            pos[c] += rx * fr[c] + ry * fr[3 + c];
            kp[c] = vx * fr[c] + vy * fr[3 + c];
            // e_z x ref = (-ref_y, ref_x, 0) in the local frame
            kt[c] = -ry * fr[c] + rx * fr[3 + c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        x_segs[3 * t + c] = pos[c];
        kphi[3 * t + c] = kp[c];
        kth[3 * t + c] = kt[c];
    }
}

// Cyclic cubic Hermite interpolation of a table sampled uniformly over one
// period (mobility_solver.cpp:240-280).
__device__ double cyclic_hermite(double phase, const double* tab, int nt) {
    const double PI = 3.14159265358979323846;
    double u = 0.5 * phase / PI;
    u -= __builtin_floor(u);
    u *= nt;
    if (!(u >= 0.0)) u = 0.0;  // a non-finite phase must not index the table
    int lo = (int)u;
    if (lo >= nt) return tab[0];
    const int hi = (lo == nt - 1) ? 0 : lo + 1;
    const double ql = tab[lo];
    const double dql = (lo == 0) ? (tab[1] - tab[0]) : 0.5 * (tab[hi] - tab[lo - 1]);
    const double qh = tab[hi];
    const double dqh = (hi == 0) ? (tab[0] - tab[nt - 1]) : 0.5 * (tab[(hi == nt - 1) ? 0 : hi + 1] - tab[lo]);
    const double t = u - lo, t2 = t * t, t3 = t2 * t;
    return ql * (2 * t3 - 3 * t2 + 1) + dql * (t3 - 2 * t2 + t) + qh * (3 * t2 - 2 * t3) + dqh * (t3 - t2);
}

// Right-hand side (mobility_solver.cpp:282-440 with PRESCRIBED_BODY_VELOCITIES):
// minus the generalised driving forces in the phase/angle rows, and the rigid
// body motion K V_body in the segment and blob rows.
__global__ __launch_bounds__(256) void rhs_kernel(RhsArgs a, const double* __restrict__ x_segs,
                                                 const double* __restrict__ x_blobs, double* __restrict__ rhs) {
    const double PI = 3.14159265358979323846;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ns = a.d.nfil * a.d.nseg;
    if (t < ns + a.d.nblob) {
        const double* x = t < ns ? x_segs + 3 * t : x_blobs + 3 * (t - ns);
        const double r0 = x[0] - a.body_x[0], r1 = x[1] - a.body_x[1], r2 = x[2] - a.body_x[2];
        const double* V = a.body_v;  // (V, Omega)
        double* o = t < ns ? rhs + 3 * t : rhs + a.d.off_blob + 3 * (t - ns);
        o[0] = V[0] + V[4] * r2 - V[5] * r1;
        o[1] = V[1] + V[5] * r0 - V[3] * r2;
        o[2] = V[2] + V[3] * r1 - V[4] * r0;
    }
    if (a.d.free_body && t < 6 * a.d.nswim) rhs[a.d.off_body + t] = 0.0;  // force- and torque-free
    if (t < a.d.nfil) {
        const double w = a.omega0[t] * 0.5 / PI;
        if (a.d.dyn_phase) {
            const double q = cyclic_hermite(a.phase[t], a.gen_phase, a.n_gen_phase) * w * a.gen_force_factor;
            rhs[a.d.off_phase + t] = -q;
        }
        if (a.d.dyn_angle) {
            double q = cyclic_hermite(a.phase[t], a.gen_angle, a.n_gen_angle) * w;
            const double L = a.length;
            q -= a.torsion_factor * (a.omega0[t] * L * L * L) * a.angle[t];
            rhs[a.d.off_angle + t] = -q;
        }
    }
}

// phase += phase_dot*dt, angle += angle_dot*dt (filament.cpp:1087-1097).
__global__ void phase_update_kernel(const double* __restrict__ x, SaddleDims d, double dt, double* phase,
                                    double* angle) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= d.nfil) return;
    if (d.dyn_phase) phase[m] += x[d.off_phase + m] * dt;
    if (d.dyn_angle) angle[m] += x[d.off_angle + m] * dt;
}

// ---- free bodies ------------------------------------------------------------
__device__ __forceinline__ int swimmer_of(const SaddleDims& d, int64_t e) {
    if (e < d.off_blob) return (int)((e / 3) / d.segs_per_swim);
    if (e < d.off_body) return (int)(((e - d.off_blob) / 3) / d.blobs_per_swim);
    if (e < d.off_phase) return (int)((e - d.off_body) / 6);
    if (e < d.off_angle) return (int)((e - d.off_phase) / d.fil_per_swim);
    return (int)((e - d.off_angle) / d.fil_per_swim);
}

__global__ __launch_bounds__(256) void body_moments_kernel(const double* __restrict__ vec, SaddleDims d,
                                                          const double* __restrict__ xs,
                                                          const double* __restrict__ xb,
                                                          const double* __restrict__ xbod, double sign,
                                                          double* __restrict__ out,
                                                          const int32_t* __restrict__ skip) {
    __shared__ double sh[256];
    if (skip && *skip) return;
    const int sw = blockIdx.x;
    const double X0 = xbod[3 * sw], X1 = xbod[3 * sw + 1], X2 = xbod[3 * sw + 2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int nparts = d.segs_per_swim + d.blobs_per_swim;
    for (int q = threadIdx.x; q < nparts; q += 256) {
        const bool seg = q < d.segs_per_swim;
        const int64_t i = seg ? (int64_t)sw * d.segs_per_swim + q : (int64_t)sw * d.blobs_per_swim + (q - d.segs_per_swim);
        const double* x = seg ? xs + 3 * i : xb + 3 * i;
        const double* f = seg ? vec + 3 * i : vec + d.off_blob + 3 * i;
        const double r0 = x[0] - X0, r1 = x[1] - X1, r2 = x[2] - X2;
        const double f0 = f[0], f1 = f[1], f2 = f[2];
        acc[0] += f0;
        acc[1] += f1;
        acc[2] += f2;
        acc[3] += r1 * f2 - r2 * f1;
        acc[4] += r2 * f0 - r0 * f2;
        acc[5] += r0 * f1 - r1 * f0;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const double tot = block_sum256(acc[c], sh);
        if (threadIdx.x == 0) out[6 * sw + c] = sign * tot;
    }
}

__global__ __launch_bounds__(256) void body_column_kernel(double* __restrict__ col, int c, SaddleDims d,
                                                         const double* __restrict__ xs,
                                                         const double* __restrict__ xb,
                                                         const double* __restrict__ xbod) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= d.n) return;
    double val = 0.0;
    if (e < d.off_body) {
        const bool seg = e < d.off_blob;
        const int64_t i = seg ? e / 3 : (e - d.off_blob) / 3;
        const int k = (int)((seg ? e : e - d.off_blob) % 3);
        const int sw = swimmer_of(d, e);
        if (c < 3) {
            val = k == c ? 1.0 : 0.0;
        } else {
            const double* x = seg ? xs + 3 * i : xb + 3 * i;
            const double r[3] = {x[0] - xbod[3 * sw], x[1] - xbod[3 * sw + 1], x[2] - xbod[3 * sw + 2]};
            // (e_a x r)_k with a = c - 3: -r[a+2] at k = a+1, r[a+1] at k = a+2
            const int a = c - 3;
            const int k1 = (a + 1) % 3, k2 = (a + 2) % 3;
            val = k == k1 ? -r[k2] : (k == k2 ? r[k1] : 0.0);
        }
    }
    col[e] = val;
}

__global__ __launch_bounds__(256) void body_correct_kernel(const double* __restrict__ in, double host_sc,
                                                          const double* __restrict__ p_in, int G,
                                                          double* __restrict__ v, double* __restrict__ z,
                                                          const double* __restrict__ Z, int64_t ldz,
                                                          const double* __restrict__ t,
                                                          const double* __restrict__ sinv, SaddleDims d,
                                                          const int32_t* __restrict__ skip) {
    __shared__ double scs;
    if (skip && *skip) return;
    if (p_in) {
        if (threadIdx.x < kWave) {
            const double tt = wave_sum_strided(p_in, G);
            if (threadIdx.x == 0) {
                const double hn = __builtin_sqrt(tt);
                scs = hn > 0.0 ? 1.0 / hn : 0.0;
            }
        }
        __syncthreads();
    }
    const double sc = p_in ? scs : host_sc;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < d.n; e += (int64_t)gridDim.x * 256) {
        const int sw = swimmer_of(d, e);
        double rb[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) rb[c] = in[d.off_body + 6 * sw + c] * sc + t[6 * sw + c];
        double y[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc += sinv[36 * sw + 6 * c + k] * rb[k];
            y[c] = -acc;
        }
        if (e >= d.off_body && e < d.off_phase) {
            const int c = (int)((e - d.off_body) % 6);
            z[e] = y[c];
            if (v) v[e] = in[e] * sc;
        } else {
            double acc = z[e];
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += Z[(int64_t)c * ldz + e] * y[c];
            z[e] = acc;
        }
    }
}

// Dense blob mobility (row-major n_pad x n_pad, n = 3 nb): 3 x 3 RPY blocks of
// every blob pair (the pair tensor's every branch, self included); the
// padding past n is the identity.  One thread per (i, j) block.
__global__ __launch_bounds__(256) void blob_mobility_dense_kernel(const double* __restrict__ xb, int nb, double a,
                                                                 double mu, double* __restrict__ A, int n_pad) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nbb = (int64_t)nb * nb;
    if (idx < nbb) {
        const int i = (int)(idx / nb), j = (int)(idx % nb);
        double blk[9];
        rpy_block_entries(xb + 3 * i, xb + 3 * j, a, mu, blk);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) A[(int64_t)(3 * i + r) * n_pad + 3 * j + c] = blk[3 * r + c];
        return;
    }
    // padding: identity rows and columns
    const int64_t p = idx - nbb;
    const int n = 3 * nb, extra = n_pad - n;
    if (p >= (int64_t)extra * n_pad) return;
    const int r = n + (int)(p / n_pad), c = (int)(p % n_pad);
    A[(int64_t)r * n_pad + c] = r == c ? 1.0 : 0.0;
    A[(int64_t)c * n_pad + r] = r == c ? 1.0 : 0.0;
}

// Pose of a swimming body (X, q): filament bases X + R b_ref, frames R e_ref,
// blob positions X + R x_ref (rigid_body::blob_positions, rigid_body.cpp:482-500);
// thread 0 also writes body_x = X.
__global__ __launch_bounds__(256) void body_pose_kernel(const double* __restrict__ pose,
                                                       const double* __restrict__ base_ref,
                                                       const double* __restrict__ frame_ref,
                                                       const double* __restrict__ blob_ref, int nfil, int nblob,
                                                       double* __restrict__ bases, double* __restrict__ frames,
                                                       double* __restrict__ x_blobs, double* __restrict__ body_x) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const double X[3] = {pose[0], pose[1], pose[2]};
    const double w = pose[3], x = pose[4], y = pose[5], z = pose[6];
    const double R[9] = {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z),       2.0 * (x * z + w * y),
                         2.0 * (x * y + w * z),       1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
                         2.0 * (x * z - w * y),       2.0 * (y * z + w * x),       1.0 - 2.0 * (x * x + y * y)};
    auto rot = [&](const double* v, double* o, bool shift) {
        for (int r = 0; r < 3; ++r) o[r] = (shift ? X[r] : 0.0) + R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
    };
    if (t == 0)
        for (int r = 0; r < 3; ++r) body_x[r] = X[r];
    if (t < nfil) {
        rot(base_ref + 3 * t, bases + 3 * t, true);
        for (int e = 0; e < 3; ++e) rot(frame_ref + 9 * t + 3 * e, frames + 9 * t + 3 * e, false);
    }
    if (t < nblob) rot(blob_ref + 3 * t, x_blobs + 3 * t, true);
}

// Explicit body update after a solve (cilia_sim_main.cu:677-693):
// X += dt V, q = exp(dt Omega) q.
__global__ void body_advance_kernel(const double* __restrict__ sol, int64_t off_body, double dt,
                                    double* __restrict__ pose) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double* V = sol + off_body;
    for (int r = 0; r < 3; ++r) pose[r] += dt * V[r];
    const double u[3] = {dt * V[3], dt * V[4], dt * V[5]};
    const double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double e0 = 1.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
    if (th >= 1e-6) {  // lie_exp (quaternion.cpp:522-545)
        const double sn = sin(0.5 * th) / th;
        e0 = cos(0.5 * th);
        e1 = sn * u[0];
        e2 = sn * u[1];
        e3 = sn * u[2];
    }
    const double q0 = pose[3], q1 = pose[4], q2 = pose[5], q3 = pose[6];
    pose[3] = e0 * q0 - (e1 * q1 + e2 * q2 + e3 * q3);
    pose[4] = e0 * q1 + q0 * e1 + e2 * q3 - e3 * q2;
    pose[5] = e0 * q2 + q0 * e2 + e3 * q1 - e1 * q3;
    pose[6] = e0 * q3 + q0 * e3 + e1 * q2 - e2 * q1;
}


// ---------------------------------------------------------------------------
// Independent bodies in lock-step: one GMRES per body.  The Krylov basis is
// kept in the bodies' own layout (body b's unknowns contiguous at koff[b]), so
// the per-vector device functions above run per body on a 2-D grid: x over
// the body's row slices (gridDim.x workgroups), y over the bodies.  A body
// whose stop flag is set does nothing.
// ---------------------------------------------------------------------------
__device__ __forceinline__ ArnoldiArgs ens_slice(const EnsArgs& e, int b, int ncols) {
    return ArnoldiArgs{e.V + e.koff[b], e.ld, ncols, e.w + e.koff[b], (int)(e.koff[b + 1] - e.koff[b]), nullptr};
}

__global__ __launch_bounds__(256) void ens_dots_kernel(EnsArgs e, int ncols, int to_pb) {
    const int b = blockIdx.y;
    if (e.stop[b]) return;
    slice_dots(ens_slice(e, b, ncols), (to_pb ? e.pb : e.pa) + (int64_t)b * e.pstride, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void ens_update_dots_kernel(EnsArgs e, int ncols, int j) {
    __shared__ double hs[kMaxBasis];
    const int b = blockIdx.y;
    if (e.stop[b]) return;
    const ArnoldiArgs a = ens_slice(e, b, ncols);
    gather_coeffs(e.pa + (int64_t)b * e.pstride, ncols, gridDim.x, hs, e.H + (int64_t)b * e.hstride + (int64_t)j * e.ldh,
                  true);
    slice_update(a, hs, blockIdx.x, gridDim.x);
    __syncthreads();
    slice_dots(a, e.pb + (int64_t)b * e.pstride, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void ens_update_sumsq_kernel(EnsArgs e, int ncols, int j, int first) {
    __shared__ double hs[kMaxBasis];
    __shared__ double sh[4];
    const int b = blockIdx.y;
    if (e.stop[b]) return;
    const ArnoldiArgs a = ens_slice(e, b, ncols);
    gather_coeffs(e.pb + (int64_t)b * e.pstride, ncols, gridDim.x, hs, e.H + (int64_t)b * e.hstride + (int64_t)j * e.ldh,
                  first != 0);
    slice_update(a, hs, blockIdx.x, gridDim.x);
    __syncthreads();
    double s = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * kSlice + threadIdx.x; r < a.n; r += (int64_t)gridDim.x * kSlice) {
#pragma unroll
        for (int q = 0; q < kSlice / 256; ++q) {
            const int64_t rr = r + q * 256;
            if (rr < a.n) s += a.w[rr] * a.w[rr];
        }
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) e.pa[(int64_t)b * e.pstride + blockIdx.x] = s;
}

// Closes step j of every running body (||w|| from its partial sums, Givens,
// estimate, stop flag; the body's workgroup 0) and, with write_next, writes
// v_{j+1} = w / ||w|| into column j + 1 of its basis.  A body that stops here
// is counted in *nstop and stamped (wall clock) in stamps[b].
__global__ __launch_bounds__(256) void ens_close_kernel(EnsArgs e, int j, int write_next) {
    __shared__ double nrm;
    const int b = blockIdx.y;
    if (e.stop[b]) return;
    if (threadIdx.x < kWave) {
        const double t = wave_sum_strided(e.pa + (int64_t)b * e.pstride, gridDim.x);
        if (threadIdx.x == 0) nrm = __builtin_sqrt(t);
    }
    __syncthreads();
    const double hn = nrm;
    if (blockIdx.x == 0) {
        const int64_t mo = (int64_t)b * e.mstride;
        arnoldi_close(hn, e.H + (int64_t)b * e.hstride, e.ldh, j, e.cs + mo, e.sn + mo, e.g + mo, e.scal + 2 * b,
                      e.est + mo, e.stop + b, e.abs_tol[b]);
        if (threadIdx.x == 0 && e.stop[b] != 0) {
            atomicAdd(e.nstop, 1);
            if (e.stamps) e.stamps[b] = (int64_t)wall_clock64();
        }
    }
    if (write_next) {
        const double sc = hn > 0.0 ? 1.0 / hn : 0.0;
        const int64_t o = e.koff[b], n = e.koff[b + 1] - o;
        double* v = e.V + (int64_t)(j + 1) * e.ld + o;
        const double* w = e.w + o;
        for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) v[r] = w[r] * sc;
    }
}

// Column 0 of every running body's basis: v_0 = r / ||r||, and g = (||r||, 0, ...).
__global__ __launch_bounds__(256) void ens_open_kernel(EnsArgs e, const double* __restrict__ r,
                                                       const double* __restrict__ rnorm) {
    const int b = blockIdx.y;
    if (e.stop[b]) return;
    const double rn = rnorm[b];
    const int64_t o = e.koff[b], n = e.koff[b + 1] - o;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        e.V[o + i] = r[o + i] / rn;
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < e.mstride; i += 256) e.g[(int64_t)b * e.mstride + i] = i == 0 ? rn : 0.0;
}

// dst[map[i]] = src[i] (to_operator) or dst[i] = src[map[i]].
__global__ __launch_bounds__(256) void ens_permute_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                          const int32_t* __restrict__ map, int64_t n, int to_operator) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        if (to_operator) dst[map[i]] = src[i];
        else dst[i] = src[map[i]];
    }
}

// out[b] = ||v_b|| over body b's rows (one workgroup per body, fixed order).
__global__ __launch_bounds__(256) void ens_norms_kernel(const double* __restrict__ v, const int64_t* __restrict__ koff,
                                                        double* __restrict__ out) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    double s = 0.0;
    for (int64_t i = koff[b] + threadIdx.x; i < koff[b + 1]; i += 256) s += v[i] * v[i];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) out[b] = __builtin_sqrt(s);
}

// Iterations of body b in the cycle just run: the stopping step, all m if it
// ran without stopping, none if it did not take part (or failed).
__device__ __forceinline__ int ens_steps(int32_t stop, int m) {
    return stop == 0 ? m : (stop > 0 && stop <= m ? stop : 0);
}

// Per body: H(0:k, 0:k) y = g by columns, the triangle staged in LDS (one
// workgroup per body); kdone[b] = k.
__global__ __launch_bounds__(256) void ens_trisolve_kernel(EnsArgs e, int m, double* __restrict__ y_all,
                                                           int32_t* __restrict__ kdone) {
    extern __shared__ double tri[];
    __shared__ double yi;
    const int b = blockIdx.x;
    const int k = ens_steps(e.stop[b], m);
    if (threadIdx.x == 0) kdone[b] = k;
    const double* H = e.H + (int64_t)b * e.hstride;
    const double* g = e.g + (int64_t)b * e.mstride;
    double* y = y_all + (int64_t)b * e.mstride;
    double* rhs = tri + (int64_t)m * (m + 1) / 2;
    for (int j = threadIdx.x; j < k; j += 256) rhs[j] = g[j];
    for (int q = threadIdx.x; q < k * k; q += 256) {
        const int i = q / k, j = q - i * k;
        if (j <= i) tri[i * (i + 1) / 2 + j] = H[(int64_t)i * e.ldh + j];
    }
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double* col = tri + (int64_t)i * (i + 1) / 2;
        if (threadIdx.x == 0) {
            yi = rhs[i] / col[i];
            y[i] = yi;
        }
        __syncthreads();
        const double v = yi;
        for (int j = threadIdx.x; j < i; j += 256) rhs[j] -= col[j] * v;
        __syncthreads();
    }
}

// out_b = V_b(:, 0:k_b) y_b for every body (zero rows where k_b = 0).
__global__ __launch_bounds__(256) void ens_gemv_kernel(EnsArgs e, const double* __restrict__ y_all,
                                                       const int32_t* __restrict__ kdone, double* __restrict__ out) {
    const int b = blockIdx.y;
    const int k = kdone[b];
    const int64_t o = e.koff[b], n = e.koff[b + 1] - o;
    const double* y = y_all + (int64_t)b * e.mstride;
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) s += e.V[(int64_t)c * e.ld + o + r] * y[c];
        out[o + r] = s;
    }
}

// Per body with flag[b] set: x = 0 and r = rhs over its rows (map: Krylov
// index -> operator index): the zero-rhs and rubbish-guess rules per body.
__global__ __launch_bounds__(256) void ens_reset_kernel(const int32_t* __restrict__ flag,
                                                        const int64_t* __restrict__ koff, int nbody,
                                                        const int32_t* __restrict__ map, double* __restrict__ x,
                                                        double* __restrict__ r, const double* __restrict__ rhs) {
    const int b = blockIdx.y;
    if (b >= nbody || !flag[b]) return;
    for (int64_t i = koff[b] + (int64_t)blockIdx.x * 256 + threadIdx.x; i < koff[b + 1]; i += (int64_t)gridDim.x * 256) {
        const int32_t o = map[i];
        x[o] = 0.0;
        if (r) r[o] = rhs[o];
    }
}

// stamps[b] = now for every body (the start of a timestep).
__global__ void ens_stamp_kernel(int64_t* __restrict__ stamps, int nbody) {
    const int64_t t = (int64_t)wall_clock64();
    for (int b = threadIdx.x; b < nbody; b += blockDim.x) stamps[b] = t;
}

}  // namespace

static inline int blocks_for(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

hipError_t k_dot_cols(const double* V, int64_t ld, int ncols, const double* w, int n, double* partial, int G,
                      hipStream_t s) {
    hipLaunchKernelGGL(dot_cols_kernel, dim3(G), dim3(256), 0, s, V, ld, ncols, w, n, partial);
    return hipGetLastError();
}
hipError_t k_reduce_cols(const double* partial, int G, int ncols, double* h, double* Hcol, hipStream_t s) {
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(ncols), dim3(256), 0, s, partial, G, h, Hcol);
    return hipGetLastError();
}
hipError_t k_gemv(const double* V, int64_t ld, int ncols, const double* h, double* w, int n, bool subtract,
                  hipStream_t s) {
    hipLaunchKernelGGL(gemv_kernel, dim3(blocks_for(n)), dim3(256), 0, s, V, ld, ncols, h, w, n, (int)subtract);
    return hipGetLastError();
}
hipError_t k_norm(const double* w, int n, double* partial, int G, double* out, hipStream_t s) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(G), dim3(256), 0, s, w, n, partial);
    hipLaunchKernelGGL(norm_finish_kernel, dim3(1), dim3(256), 0, s, partial, G, out);
    return hipGetLastError();
}
hipError_t k_givens(double* H, int ldh, int j, double* cs, double* sn, double* g, const double* wnorm,
                    double* scal, hipStream_t s) {
    hipLaunchKernelGGL(givens_kernel, dim3(1), dim3(64), 0, s, H, ldh, j, cs, sn, g, wnorm, scal);
    return hipGetLastError();
}
hipError_t k_scale(double* v, const double* w, const double* scal, int idx, int n, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(256), 0, s, v, w, scal, idx, n);
    return hipGetLastError();
}
hipError_t k_trisolve(const double* H, int ldh, const double* g, int k, double* y, hipStream_t s) {
    const size_t staged = sizeof(double) * ((size_t)k * (k + 1) / 2 + k);
    if (staged <= 60 * 1024)
        hipLaunchKernelGGL(trisolve_kernel<true>, dim3(1), dim3(256), staged, s, H, ldh, g, k, y);
    else
        hipLaunchKernelGGL(trisolve_kernel<false>, dim3(1), dim3(256), sizeof(double) * k, s, H, ldh, g, k, y);
    return hipGetLastError();
}
hipError_t k_extrapolate(double* x, double* h1, double* h2, int n, int order, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(extrapolate_kernel, dim3(blocks_for(n)), dim3(256), 0, s, x, h1, h2, n, order);
    return hipGetLastError();
}
hipError_t k_axpby(double a, const double* x, double b, double* y, int n, hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, dim3(blocks_for(n)), dim3(256), 0, s, a, x, b, y, n);
    return hipGetLastError();
}
hipError_t k_arnoldi_step(const double* V, int64_t ld, int j, double* w, int n, double* pa, double* pb, int G,
                          double* H, int ldh, double* cs, double* sn, double* g, double* scal, double* est,
                          int32_t* stop, double abs_tol, int passes, hipStream_t s) {
    if (j + 1 > kMaxBasis || (passes != 1 && passes != 2)) return hipErrorInvalidValue;
    const ArnoldiArgs a{V, ld, j + 1, w, n, stop};
    double* Hcol = H + (int64_t)j * ldh;
    if (passes == 2) {
        hipLaunchKernelGGL(arnoldi_dots_kernel, dim3(G), dim3(256), 0, s, a, pa);
        hipLaunchKernelGGL(arnoldi_update_dots_kernel, dim3(G), dim3(256), 0, s, a, (const double*)pa, pb, Hcol);
        hipLaunchKernelGGL(arnoldi_update_sumsq_kernel, dim3(G), dim3(256), 0, s, a, (const double*)pb, pa, Hcol,
                           false);
    } else {
        // one pass: the projections go to pb, so that the norm partials
        // written to pa overlap no coefficient partial still being read
        hipLaunchKernelGGL(arnoldi_dots_kernel, dim3(G), dim3(256), 0, s, a, pb);
        hipLaunchKernelGGL(arnoldi_update_sumsq_kernel, dim3(G), dim3(256), 0, s, a, (const double*)pb, pa, Hcol,
                           true);
    }
    return hipGetLastError();
}
// Waves for a preconditioner launch: one per filament (or per patch in the
// block form), or enough threads for the blob entries.
static int64_t precond_work(const PrecondArgs& P, const SaddleDims& d) {
    const int64_t waves = P.kind == 1 ? (int64_t)d.nfil + P.npatch : d.nfil;
    const int64_t work = waves * kWave > 3LL * d.nblob ? waves * kWave : 3LL * d.nblob;
    return work > 0 ? work : 1;
}
hipError_t k_arnoldi_close_next(const double* pa, int G, double* H, int ldh, int j, double* cs, double* sn,
                                double* g, double* scal, double* est, int32_t* stop, int32_t* stop_mirror,
                                double* est_mirror, double abs_tol, const double* w, double* v, double* out, const PrecondArgs& P,
                                const SaddleDims& d, hipStream_t s) {
    if (j + 1 > kMaxBasis) return hipErrorInvalidValue;
    const int grid = v ? blocks_for(precond_work(P, d)) : 1;
    hipLaunchKernelGGL(arnoldi_close_next_kernel, dim3(grid), dim3(256), 0, s, pa, G, H, ldh, j, cs, sn, g, scal,
                       est, stop, stop_mirror, est_mirror, abs_tol, w, v, out, P, d);
    return hipGetLastError();
}
int arnoldi_blocks(int64_t n) {
    const int64_t g = (n + kSlice - 1) / kSlice;
    return (int)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
}
hipError_t k_scale_precond(const double* w, const double* dev_scale, double host_scale, double* v, double* out,
                           const PrecondArgs& P, const SaddleDims& d, const int32_t* skip, hipStream_t s) {
    hipLaunchKernelGGL(scale_precond_kernel, dim3(blocks_for(precond_work(P, d))), dim3(256), 0, s, w, dev_scale,
                       host_scale, v, out, P, d, skip);
    return hipGetLastError();
}
hipError_t k_sub(const double* x, const double* y, double* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(n)), dim3(256), 0, s, x, y, out, n);
    return hipGetLastError();
}
hipError_t k_precond(const double* in, double* out, const PrecondArgs& P, const SaddleDims& d, hipStream_t s) {
    hipLaunchKernelGGL(precond_kernel, dim3(blocks_for(precond_work(P, d))), dim3(256), 0, s, in, out, P, d);
    return hipGetLastError();
}
hipError_t k_fil_blocks(const double* x_segs, int nfil, int nseg, double a_seg, double mu, double* fil_blk,
                        hipStream_t s) {
    if (nfil <= 0) return hipSuccess;
    hipLaunchKernelGGL(fil_blocks_kernel, dim3((unsigned)nfil), dim3(256), 0, s, x_segs, nseg, a_seg, mu, fil_blk);
    return hipGetLastError();
}
hipError_t k_patch_blocks(const double* x_blobs, const int32_t* patch_blob, const int32_t* patch_start,
                          const int64_t* patch_off, int npatch, double a_blob, double mu, double* patch_blk,
                          hipStream_t s) {
    if (npatch <= 0) return hipSuccess;
    hipLaunchKernelGGL(patch_blocks_kernel, dim3((unsigned)npatch), dim3(256), 0, s, x_blobs, patch_blob,
                       patch_start, patch_off, a_blob, mu, patch_blk);
    return hipGetLastError();
}
hipError_t k_fil_schur(const PrecondArgs& P, const SaddleDims& d, double* fil_dk, double* fil_sinv, hipStream_t s) {
    if (d.nfil <= 0) return hipSuccess;
    hipLaunchKernelGGL(fil_schur_kernel, dim3(blocks_for((int64_t)d.nfil * kWave)), dim3(256), 0, s, P, d, fil_dk,
                       fil_sinv);
    return hipGetLastError();
}
hipError_t k_beat_coeffs(const double* A, const double* B) {
    TFS_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_beat_A), A, sizeof(double) * 24));
    TFS_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_beat_B), B, sizeof(double) * 18));
    return hipSuccess;
}
hipError_t k_beat(const BeatArgs& a, double* x_segs, double* kphi, double* kth, hipStream_t s) {
    const int n = a.nfil * a.nseg;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(beat_kernel, dim3(blocks_for(n)), dim3(256), 0, s, a, x_segs, kphi, kth);
    return hipGetLastError();
}
hipError_t k_rhs(const RhsArgs& a, const double* x_segs, const double* x_blobs, double* rhs, hipStream_t s) {
    const int n = a.d.nfil * a.d.nseg + a.d.nblob;
    const int m = n > a.d.nfil ? n : a.d.nfil;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(rhs_kernel, dim3(blocks_for(m)), dim3(256), 0, s, a, x_segs, x_blobs, rhs);
    return hipGetLastError();
}
hipError_t k_phase_update(const double* x, const SaddleDims& d, double dt, double* phase, double* angle,
                          hipStream_t s) {
    if (d.nfil == 0) return hipSuccess;
    hipLaunchKernelGGL(phase_update_kernel, dim3(blocks_for(d.nfil)), dim3(256), 0, s, x, d, dt, phase, angle);
    return hipGetLastError();
}

hipError_t k_body_moments(const double* vec, const SaddleDims& d, const double* x_segs, const double* x_blobs,
                          const double* x_bodies, double sign, double* out, const int32_t* skip, hipStream_t s) {
    hipLaunchKernelGGL(body_moments_kernel, dim3((unsigned)d.nswim), dim3(256), 0, s, vec, d, x_segs, x_blobs, x_bodies,
                       sign, out, skip);
    return hipGetLastError();
}

hipError_t k_body_column(double* col, int c, const SaddleDims& d, const double* x_segs, const double* x_blobs,
                         const double* x_bodies, hipStream_t s) {
    hipLaunchKernelGGL(body_column_kernel, dim3(blocks_for(d.n)), dim3(256), 0, s, col, c, d, x_segs, x_blobs,
                       x_bodies);
    return hipGetLastError();
}

hipError_t k_body_correct(const double* in, double host_sc, const double* p_in, int G, double* v, double* z,
                          const double* Z, int64_t ldz, const double* t, const double* sinv, const SaddleDims& d,
                          const int32_t* skip, hipStream_t s) {
    const int grid = std::min(blocks_for(d.n), 1024);
    hipLaunchKernelGGL(body_correct_kernel, dim3(grid), dim3(256), 0, s, in, host_sc, p_in, G, v, z, Z, ldz, t, sinv,
                       d, skip);
    return hipGetLastError();
}

hipError_t k_blob_mobility_dense(const double* x_blobs, int nb, double a, double mu, double* A, int n_pad,
                                 hipStream_t s) {
    const int64_t work = (int64_t)nb * nb + (int64_t)(n_pad - 3 * nb) * n_pad;
    hipLaunchKernelGGL(blob_mobility_dense_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, x_blobs, nb,
                       a, mu, A, n_pad);
    return hipGetLastError();
}

hipError_t k_body_pose(const double* pose, const double* base_ref, const double* frame_ref, const double* blob_ref,
                       int nfil, int nblob, double* bases, double* frames, double* x_blobs, double* body_x,
                       hipStream_t s) {
    const int n = std::max(std::max(nfil, nblob), 1);
    hipLaunchKernelGGL(body_pose_kernel, dim3(blocks_for(n)), dim3(256), 0, s, pose, base_ref, frame_ref, blob_ref,
                       nfil, nblob, bases, frames, x_blobs, body_x);
    return hipGetLastError();
}

hipError_t k_body_advance(const double* sol, int64_t off_body, double dt, double* pose, hipStream_t s) {
    hipLaunchKernelGGL(body_advance_kernel, dim3(1), dim3(64), 0, s, sol, off_body, dt, pose);
    return hipGetLastError();
}

// ---- independent bodies in lock-step ----------------------------------------
hipError_t k_ens_arnoldi_step(const EnsArgs& e, int j, int G, int passes, hipStream_t s) {
    if (j + 1 > kMaxBasis || (passes != 1 && passes != 2)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)G, (unsigned)e.nbody);
    if (passes == 2) {
        hipLaunchKernelGGL(ens_dots_kernel, grid, dim3(256), 0, s, e, j + 1, 0);
        hipLaunchKernelGGL(ens_update_dots_kernel, grid, dim3(256), 0, s, e, j + 1, j);
        hipLaunchKernelGGL(ens_update_sumsq_kernel, grid, dim3(256), 0, s, e, j + 1, j, 0);
    } else {
        // one pass: the projections go to pb (the norm partials go to pa)
        hipLaunchKernelGGL(ens_dots_kernel, grid, dim3(256), 0, s, e, j + 1, 1);
        hipLaunchKernelGGL(ens_update_sumsq_kernel, grid, dim3(256), 0, s, e, j + 1, j, 1);
    }
    return hipGetLastError();
}
hipError_t k_ens_close(const EnsArgs& e, int j, bool write_next, int G, hipStream_t s) {
    hipLaunchKernelGGL(ens_close_kernel, dim3((unsigned)G, (unsigned)e.nbody), dim3(256), 0, s, e, j,
                       write_next ? 1 : 0);
    return hipGetLastError();
}
hipError_t k_ens_open(const EnsArgs& e, const double* r, const double* rnorm, int G, hipStream_t s) {
    hipLaunchKernelGGL(ens_open_kernel, dim3((unsigned)G, (unsigned)e.nbody), dim3(256), 0, s, e, r, rnorm);
    return hipGetLastError();
}
hipError_t k_ens_permute(const double* src, double* dst, const int32_t* map, int64_t n, bool to_operator,
                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ens_permute_kernel, dim3((unsigned)std::min<int64_t>(blocks_for(n), 4096)), dim3(256), 0, s,
                       src, dst, map, n, to_operator ? 1 : 0);
    return hipGetLastError();
}
hipError_t k_ens_norms(const double* v, const int64_t* koff, int nbody, double* out, hipStream_t s) {
    hipLaunchKernelGGL(ens_norms_kernel, dim3((unsigned)nbody), dim3(256), 0, s, v, koff, out);
    return hipGetLastError();
}
hipError_t k_ens_update(const EnsArgs& e, int m, double* y, int32_t* kdone, double* out, int G, hipStream_t s) {
    const size_t shmem = sizeof(double) * ((size_t)m * (m + 1) / 2 + m);
    if (shmem > 60 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ens_trisolve_kernel, dim3((unsigned)e.nbody), dim3(256), shmem, s, e, m, y, kdone);
    hipLaunchKernelGGL(ens_gemv_kernel, dim3((unsigned)G, (unsigned)e.nbody), dim3(256), 0, s, e, (const double*)y,
                       (const int32_t*)kdone, out);
    return hipGetLastError();
}
hipError_t k_ens_reset(const int32_t* flag, const int64_t* koff, int nbody, const int32_t* map, double* x, double* r,
                       const double* rhs, int G, hipStream_t s) {
    hipLaunchKernelGGL(ens_reset_kernel, dim3((unsigned)G, (unsigned)nbody), dim3(256), 0, s, flag, koff, nbody, map, x,
                       r, rhs);
    return hipGetLastError();
}
hipError_t k_ens_stamp(int64_t* stamps, int nbody, hipStream_t s) {
    hipLaunchKernelGGL(ens_stamp_kernel, dim3(1), dim3(256), 0, s, stamps, nbody);
    return hipGetLastError();
}

}  // namespace tfs
