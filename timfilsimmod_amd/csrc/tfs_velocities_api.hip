// tfs_velocities_api.hip -- mobility_solver::compute_velocities for elastic
// (non-prescribed) filaments on a body with prescribed velocity
// (mobility_solver.cpp:2598-2655), behind include/tfs_solver.h.
//
// The reference runs Mss_mult and Mbs_mult, copies the blob velocities to the
// host, assembles -v_blobs + K V on the host, runs GMRES on M_bb with host
// Arnoldi, copies the blob forces back and runs Msb_mult.  Here every stage
// stays on the device: two matvec evaluations of one RPY handle, a kernel for
// the right-hand side, the device GMRES of a blob-only solver handle, and the
// accumulating segment-blob evaluation, all on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/tfs_solver.h"
#include "dense_kernels.h"
#include "solver_kernels.h"
#include "tfs_api_internal.h"

using namespace tfs;

namespace {

// rhs[3m + d] = -v_blobs[3m + d] + (V + Omega x (x_blob - x_body))_d, the body
// of blob m being m / blobs_per_swim (mobility_solver.cpp:442-446, 533-545).
__global__ __launch_bounds__(256) void blob_rhs_kernel(const double* __restrict__ v_blobs,
                                                       const double* __restrict__ x_blobs,
                                                       const double* __restrict__ x_bodies,
                                                       const double* __restrict__ v_bodies, int nblob,
                                                       int blobs_per_swim, int have_segs,
                                                       double* __restrict__ rhs) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= nblob) return;
    const int b = m / blobs_per_swim;
    const double r0 = x_blobs[3 * m] - x_bodies[3 * b];
    const double r1 = x_blobs[3 * m + 1] - x_bodies[3 * b + 1];
    const double r2 = x_blobs[3 * m + 2] - x_bodies[3 * b + 2];
    const double* V = v_bodies + 6 * b;
    const double k0 = V[0] + V[4] * r2 - V[5] * r1;
    const double k1 = V[1] + V[5] * r0 - V[3] * r2;
    const double k2 = V[2] + V[3] * r1 - V[4] * r0;
    rhs[3 * m] = (have_segs ? -v_blobs[3 * m] : 0.0) + k0;
    rhs[3 * m + 1] = (have_segs ? -v_blobs[3 * m + 1] : 0.0) + k1;
    rhs[3 * m + 2] = (have_segs ? -v_blobs[3 * m + 2] : 0.0) + k2;
}

// Free bodies: rhs[3 nb + 6 b + k] = -swimmer.f = sum over the body's segments
// of (F, T + (x - X) x F) minus the external force and torque
// (swimmer.cpp:469-533, mobility_solver.cpp:551-556); blob rows -v_blobs.
// One workgroup per body.
__global__ __launch_bounds__(256) void free_rhs_kernel(const double* __restrict__ x_segs,
                                                       const double* __restrict__ f_segs,
                                                       const double* __restrict__ v_blobs, int64_t nblob,
                                                       int segs_per_swim, const double* __restrict__ x_bodies,
                                                       const double* __restrict__ f_ext, double* __restrict__ rhs) {
    const int b = blockIdx.x;
    // blob rows, spread over every workgroup
    for (int64_t i = (int64_t)b * 256 + threadIdx.x; i < 3 * nblob; i += (int64_t)gridDim.x * 256)
        rhs[i] = segs_per_swim > 0 ? -v_blobs[i] : 0.0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double X0 = x_bodies[3 * b], X1 = x_bodies[3 * b + 1], X2 = x_bodies[3 * b + 2];
    for (int i = threadIdx.x; i < segs_per_swim; i += 256) {
        const int64_t g = (int64_t)b * segs_per_swim + i;
        const double* f = f_segs + 6 * g;
        const double r0 = x_segs[3 * g] - X0, r1 = x_segs[3 * g + 1] - X1, r2 = x_segs[3 * g + 2] - X2;
        acc[0] += f[0];
        acc[1] += f[1];
        acc[2] += f[2];
        acc[3] += f[3] + r1 * f[2] - r2 * f[1];
        acc[4] += f[4] + r2 * f[0] - r0 * f[2];
        acc[5] += f[5] + r0 * f[1] - r1 * f[0];
    }
    __shared__ double sm[6][256];
#pragma unroll
    for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) rhs[3 * nblob + 6 * b + threadIdx.x] = sm[threadIdx.x][0] - (f_ext ? f_ext[6 * b + threadIdx.x] : 0.0);
}

// ---- direct solve with one free body --------------------------------------
// The blob mobility at pose (X, R) is (I x R) M_ref (I x R^T), M_ref that of
// the body-frame reference positions; with A = M_ref^{-1}, Z = A K_ref and
// N = (K_ref^T Z)^{-1} (the body mobility) the free-body system solves in the
// body frame as y = A r_ref, V_ref = N (f_ref - K_ref^T y), f_b = R (y + Z V_ref).
__device__ __forceinline__ void rot_of(const double* q, double* R) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1.0 - 2.0 * (y * y + z * z);
    R[1] = 2.0 * (x * y - w * z);
    R[2] = 2.0 * (x * z + w * y);
    R[3] = 2.0 * (x * y + w * z);
    R[4] = 1.0 - 2.0 * (x * x + z * z);
    R[5] = 2.0 * (y * z - w * x);
    R[6] = 2.0 * (x * z - w * y);
    R[7] = 2.0 * (y * z + w * x);
    R[8] = 1.0 - 2.0 * (x * x + y * y);
}

// Column k of K_ref at every blob: e_k (k < 3) or e_{k-3} x ref.
__global__ __launch_bounds__(256) void kref_column_kernel(const double* __restrict__ refs, int nb, int k,
                                                          double* __restrict__ x) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= nb) return;
    const double* r = refs + 3 * m;
    double v[3] = {0.0, 0.0, 0.0};
    if (k < 3) {
        v[k] = 1.0;
    } else {
        const int a = k - 3;
        // e_a x r
        v[0] = a == 1 ? r[2] : (a == 2 ? -r[1] : 0.0);
        v[1] = a == 2 ? r[0] : (a == 0 ? -r[2] : 0.0);
        v[2] = a == 0 ? r[1] : (a == 1 ? -r[0] : 0.0);
    }
    for (int c = 0; c < 3; ++c) x[3 * m + c] = v[c];
}

// out = K_ref^T y = (sum y_m, sum ref_m x y_m), one workgroup.
__device__ void kref_transpose(const double* refs, int nb, const double* y, double* out, double (*sm)[256]) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int m = threadIdx.x; m < nb; m += 256) {
        const double* r = refs + 3 * m;
        const double* f = y + 3 * m;
        acc[0] += f[0];
        acc[1] += f[1];
        acc[2] += f[2];
        acc[3] += r[1] * f[2] - r[2] * f[1];
        acc[4] += r[2] * f[0] - r[0] * f[2];
        acc[5] += r[0] * f[1] - r[1] * f[0];
    }
    for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = sm[threadIdx.x][0];
    __syncthreads();
}

__global__ __launch_bounds__(256) void kref_transpose_kernel(const double* __restrict__ refs, int nb,
                                                             const double* __restrict__ y, double* __restrict__ out) {
    __shared__ double sm[6][256];
    kref_transpose(refs, nb, y, out, sm);
}

// Blob rows into the body frame, r_ref = R^T r; f_ref = R^T f from the body
// rows (which hold -f).
__global__ __launch_bounds__(256) void to_body_frame_kernel(const double* __restrict__ rhs, int nb,
                                                            const double* __restrict__ q, double* __restrict__ xpad,
                                                            double* __restrict__ fref) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    double R[9];
    rot_of(q, R);
    if (m < nb) {
        const double* r = rhs + 3 * m;
        for (int c = 0; c < 3; ++c) xpad[3 * m + c] = R[c] * r[0] + R[3 + c] * r[1] + R[6 + c] * r[2];
    }
    if (m == 0) {
        const double* b = rhs + 3 * nb;
        for (int c = 0; c < 3; ++c) {
            fref[c] = -(R[c] * b[0] + R[3 + c] * b[1] + R[6 + c] * b[2]);
            fref[3 + c] = -(R[c] * b[3] + R[3 + c] * b[4] + R[6 + c] * b[5]);
        }
    }
}

// V_ref = N (f_ref - K_ref^T y); the body velocity R V_ref into sol_body.
__global__ __launch_bounds__(256) void free_direct_velocity_kernel(const double* __restrict__ refs, int nb,
                                                                   const double* __restrict__ y,
                                                                   const double* __restrict__ fref,
                                                                   const double* __restrict__ nref,
                                                                   const double* __restrict__ q,
                                                                   double* __restrict__ vref,
                                                                   double* __restrict__ sol_body) {
    __shared__ double sm[6][256];
    __shared__ double kt[6];
    kref_transpose(refs, nb, y, kt, sm);
    if (threadIdx.x == 0) {
        double g[6], v[6];
        for (int k = 0; k < 6; ++k) g[k] = fref[k] - kt[k];
        for (int k = 0; k < 6; ++k) {
            double a = 0.0;
            for (int j = 0; j < 6; ++j) a += nref[6 * k + j] * g[j];
            v[k] = a;
            vref[k] = a;
        }
        double R[9];
        rot_of(q, R);
        for (int c = 0; c < 3; ++c) {
            sol_body[c] = R[3 * c] * v[0] + R[3 * c + 1] * v[1] + R[3 * c + 2] * v[2];
            sol_body[3 + c] = R[3 * c] * v[3] + R[3 * c + 1] * v[4] + R[3 * c + 2] * v[5];
        }
    }
}

// f_b = R (y + Z V_ref) per blob.
__global__ __launch_bounds__(256) void free_direct_forces_kernel(const double* __restrict__ y,
                                                                 const double* __restrict__ Z, int64_t n_pad,
                                                                 const double* __restrict__ vref,
                                                                 const double* __restrict__ q, int nb,
                                                                 double* __restrict__ out) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= nb) return;
    double R[9];
    rot_of(q, R);
    double b[3];
    for (int c = 0; c < 3; ++c) {
        double a = y[3 * m + c];
        for (int k = 0; k < 6; ++k) a += Z[k * n_pad + 3 * m + c] * vref[k];
        b[c] = a;
    }
    for (int c = 0; c < 3; ++c) out[3 * m + c] = R[3 * c] * b[0] + R[3 * c + 1] * b[1] + R[3 * c + 2] * b[2];
}

// Up to eight device-to-device copies of doubles in one launch.
struct CopyList {
    const double* src[8];
    double* dst[8];
    int64_t n[8];
    int count = 0;
    int64_t longest = 0;
    void add(const double* s, double* d, int64_t len) {
        if (s == d || len <= 0) return;
        src[count] = s;
        dst[count] = d;
        n[count] = len;
        ++count;
        longest = len > longest ? len : longest;
    }
};

__global__ __launch_bounds__(256) void copy_list_kernel(CopyList L) {
    for (int k = 0; k < L.count; ++k)
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < L.n[k]; i += (int64_t)gridDim.x * 256)
            L.dst[k][i] = L.src[k][i];
}

hipError_t copy_list(const CopyList& L, hipStream_t st) {
    if (L.count == 0) return hipSuccess;
    const int64_t g = (L.longest + 255) / 256;
    hipLaunchKernelGGL(copy_list_kernel, dim3((unsigned)(g < 1024 ? g : 1024)), dim3(256), 0, st, L);
    return hipGetLastError();
}

// flag = 1 if a[i] != b[i] for some i (every writer stores the same value).
__global__ __launch_bounds__(256) void differs_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                     int64_t n, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && a[i] != b[i]) *flag = 1;
}

int check(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }

}  // namespace

struct tfs_velocities {
    tfs_velocities_desc desc{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    tfs_rpy* rpy = nullptr;       // segments (6 DOF) and blobs, BS overwriting
    tfs_solver* blobs = nullptr;  // GMRES on M_bb
    int64_t ns = 0, nb = 0, nswim = 0;
    double* d_rhs = nullptr;      // 3*nb
    double* d_bodies = nullptr;   // 3*nswim positions then 6*nswim velocities (or external forces)
    double* d_sol = nullptr;      // free bodies: [blob forces | body velocities]
    // The blob solve's geometry (and so its preconditioner blocks) is renewed
    // only when the blob positions change: flag in device memory, its host
    // copy in pinned memory.
    int32_t* d_flag = nullptr;
    int32_t* h_flag = nullptr;
    bool have_blobs = false;
    // Direct blob solve (TFS_BLOB_DIRECT): the inverse of the dense blob
    // mobility, formed once per blob geometry, applied as one dense product.
    bool direct = false, direct_valid = false;
    int n_pad = 0, chunks = 32;
    double *Ainv = nullptr, *xpad = nullptr, *ypad = nullptr, *sym_partial = nullptr;
    void* dense_arena = nullptr;
    tfs::DenseWork dw;
    // one free body, direct: body-frame blob positions, Z = A K_ref (6
    // columns of n_pad), N (6 x 6), and V_ref, f_ref, K^T Z scratch
    double *body_refs = nullptr, *Z = nullptr, *nref = nullptr, *small = nullptr, *d_q = nullptr;
    bool free_direct_valid = false;
    int32_t blob_motion = TFS_BLOBS_CHECK;  // tfs_velocities_set_blob_motion
    // FP32 entry points: staging, and the FP64 images of the caller's arrays
    tfs::F32Stage f32;
    double* f32_arena = nullptr;
};

namespace {

// Forms v->Ainv = M_bb^{-1} for the blob positions x (device, 3 per blob).
int build_direct(tfs_velocities* v, const double* x) {
    const int nb = (int)v->nb;
    const int n = 3 * nb;
    const int n_pad = (n + kDenseTile - 1) / kDenseTile * kDenseTile;
    const int nt = n_pad / kDenseTile;
    hipStream_t st = v->stream;
    if (!v->Ainv) {
        v->n_pad = n_pad;
        v->chunks = nt;  // k_symv_dense's partials: one row of tiles per 64 rows
        if (hipMalloc(&v->Ainv, 8 * (size_t)n_pad * n_pad) != hipSuccess)
            return fail(TFS_ERR_OUT_OF_MEMORY, "direct blob solve: " + std::to_string(8.0 * n_pad * n_pad / 1e9) +
                                                   " GB for the dense blob mobility");
        const size_t bytes = 8 * ((size_t)2 * n_pad + (size_t)v->chunks * n_pad + kDenseTile * kDenseTile) + 8 + 4 +
                             4 * kDenseTile + 4 + 4 * (size_t)(nt + 1) + 64;
        if (hipMalloc(&v->dense_arena, bytes) != hipSuccess) return fail(TFS_ERR_OUT_OF_MEMORY, "direct blob solve: work space");
        char* p = (char*)v->dense_arena;
        auto take = [&](size_t b) { char* q = p; p += (b + 7) / 8 * 8; return q; };
        v->xpad = (double*)take(8 * (size_t)n_pad);
        v->ypad = (double*)take(8 * (size_t)n_pad);
        v->sym_partial = (double*)take(8 * (size_t)v->chunks * n_pad);
        v->dw.pivot = (double*)take(8 * kDenseTile * kDenseTile);
        v->dw.off = (int64_t*)take(8);
        v->dw.size = (int32_t*)take(4);
        v->dw.piv = (int32_t*)take(4 * kDenseTile);
        v->dw.piv_off = (int32_t*)take(4);
        v->dw.status = (int32_t*)take(4 * (size_t)(nt + 1));
        const int64_t zero = 0;
        const int32_t sz = kDenseTile, zero32 = 0;
        int rc = check(hipMemsetAsync(v->xpad, 0, 8 * (size_t)n_pad, st), "direct blob solve");
        if (!rc) rc = check(hipMemcpyAsync(v->dw.off, &zero, 8, hipMemcpyHostToDevice, st), "direct blob solve");
        if (!rc) rc = check(hipMemcpyAsync(v->dw.size, &sz, 4, hipMemcpyHostToDevice, st), "direct blob solve");
        if (!rc) rc = check(hipMemcpyAsync(v->dw.piv_off, &zero32, 4, hipMemcpyHostToDevice, st), "direct blob solve");
        if (!rc) rc = check(hipStreamSynchronize(st), "direct blob solve");
        if (rc) return rc;
    }
    int rc = check(k_blob_mobility_dense(x, nb, v->desc.blob_radius, v->desc.mu, v->Ainv, n_pad, st),
                   "dense blob mobility");
    if (!rc) rc = check(k_spd_inverse_blocked(v->Ainv, n_pad, v->dw, st), "dense blob inverse");
    std::vector<int32_t> status(nt);
    if (!rc) rc = check(hipMemcpyAsync(status.data(), v->dw.status, 4 * (size_t)nt, hipMemcpyDeviceToHost, st), "status");
    if (!rc) rc = check(hipStreamSynchronize(st), "dense blob inverse");
    if (rc) return rc;
    for (int k = 0; k < nt; ++k)
        if (status[k]) return fail(TFS_ERR_NON_FINITE, "direct blob solve: singular pivot block " + std::to_string(k));
    v->direct_valid = true;
    return TFS_OK;
}

// The inverse at the body-frame reference positions, Z = A K_ref and N.
int build_free_direct(tfs_velocities* v) {
    const int nb = (int)v->nb;
    hipStream_t st = v->stream;
    int rc = build_direct(v, v->body_refs);
    v->direct_valid = false;  // not the lab-frame inverse
    if (rc) return rc;
    const int64_t n_pad = v->n_pad;
    if (!v->Z) {
        if (hipMalloc(&v->Z, 8 * (6 * n_pad + 36 + 32)) != hipSuccess)
            return fail(TFS_ERR_OUT_OF_MEMORY, "direct free-body solve: work space");
        v->nref = v->Z + 6 * n_pad;
        v->small = v->nref + 36;
    }
    double kz[36];
    for (int k = 0; k < 6 && !rc; ++k) {
        hipLaunchKernelGGL(kref_column_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                           (const double*)v->body_refs, nb, k, v->xpad);
        rc = check(hipGetLastError(), "K_ref column");
        if (!rc) rc = check(k_symv_dense(v->Ainv, v->n_pad, v->xpad, v->ypad, v->sym_partial, v->chunks, st), "A K_ref");
        if (!rc) rc = check(hipMemcpyAsync(v->Z + k * n_pad, v->ypad, 8 * n_pad, hipMemcpyDeviceToDevice, st), "Z");
        if (!rc) {
            hipLaunchKernelGGL(kref_transpose_kernel, dim3(1), dim3(256), 0, st, (const double*)v->body_refs, nb,
                               (const double*)v->ypad, v->small);
            rc = check(hipGetLastError(), "K_ref^T Z");
        }
        double col[6];
        if (!rc) rc = check(hipMemcpyAsync(col, v->small, 48, hipMemcpyDeviceToHost, st), "K_ref^T Z");
        if (!rc) rc = check(hipStreamSynchronize(st), "K_ref^T Z");
        for (int r = 0; r < 6; ++r) kz[6 * r + k] = col[r];
    }
    if (rc) return rc;
    double a[6][12];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 12; ++c) a[r][c] = c < 6 ? kz[6 * r + c] : (c - 6 == r ? 1.0 : 0.0);
    for (int k = 0; k < 6; ++k) {
        int piv = k;
        for (int r = k + 1; r < 6; ++r)
            if (std::fabs(a[r][k]) > std::fabs(a[piv][k])) piv = r;
        if (!(std::fabs(a[piv][k]) > 0.0) || !std::isfinite(a[piv][k]))
            return fail(TFS_ERR_NON_FINITE, "direct free-body solve: singular body resistance");
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
    double nr[36];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) nr[6 * r + c] = a[r][6 + c];
    rc = check(hipMemcpyAsync(v->nref, nr, 288, hipMemcpyHostToDevice, st), "N");
    if (!rc) rc = check(hipStreamSynchronize(st), "N");
    if (!rc) v->free_direct_valid = true;
    return rc;
}

}  // namespace

extern "C" {

int tfs_velocities_create(const tfs_velocities_desc* desc, tfs_velocities** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: null argument");
    *out = nullptr;
    const tfs_velocities_desc& D = *desc;
    if (D.num_swim < 1 || D.segs_per_swim < 0 || D.blobs_per_swim < 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: bad sizes");
    if (D.memory != TFS_MEMORY_HOST && D.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: unknown memory space");
    if (D.free_bodies != 0 && D.free_bodies != 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: free_bodies is 0 or 1");
    if (D.mode != TFS_SOLVE_FP64 && D.mode != TFS_SOLVE_MIXED)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: unknown mode");
    const int64_t ns = (int64_t)D.num_swim * D.segs_per_swim, nb = (int64_t)D.num_swim * D.blobs_per_swim;
    if (ns > INT32_MAX / 8 || nb > INT32_MAX / 8)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_create: too many particles");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TFS_ERR_NO_DEVICE, "tfs_velocities_create: no HIP device visible");
    tfs_velocities* v = new tfs_velocities();
    v->desc = D;
    v->ns = ns;
    v->nb = nb;
    v->nswim = D.num_swim;
    hipError_t e = hipSuccess;
    if (D.stream) {
        v->stream = (hipStream_t)D.stream;
    } else {
        e = hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking);
        v->own_stream = e == hipSuccess;
    }
    if (e != hipSuccess) {
        delete v;
        return hip_fail(e, "tfs_velocities_create: stream");
    }
    tfs_rpy_desc rd{};
    rd.num_segs = (int32_t)ns;
    rd.num_blobs = (int32_t)nb;
    rd.seg_dof = 6;
    rd.precision = D.mode == TFS_SOLVE_MIXED ? TFS_PRECISION_FP32_ACC64 : TFS_PRECISION_FP64;
    rd.bs_mode = TFS_BS_OVERWRITE;
    rd.memory = TFS_MEMORY_DEVICE;
    rd.seg_radius = D.seg_radius;
    rd.blob_radius = D.blob_radius;
    rd.mu = D.mu;
    rd.stream = v->stream;
    int rc = tfs_rpy_create(&rd, &v->rpy);
    tfs_solver_desc sd{};
    sd.num_fil = 0;
    sd.num_seg = 0;
    sd.num_blobs = (int32_t)nb;
    sd.mode = D.mode;
    sd.seg_radius = D.seg_radius;
    sd.blob_radius = D.blob_radius;
    sd.mu = D.mu;
    sd.restart = D.restart;
    sd.max_cycles = D.max_cycles;
    sd.tol = D.tol;
    sd.inner_tol = D.inner_tol;
    sd.max_refinements = D.max_refinements;
    sd.stream = v->stream;
    sd.num_swim = D.num_swim;
    sd.free_bodies = D.free_bodies;
    const int64_t nsys = 3 * nb + (D.free_bodies ? 6LL * D.num_swim : 0);
    if (!rc) rc = tfs_solver_create(&sd, &v->blobs);
    if (!rc) rc = check(hipMalloc(&v->d_rhs, sizeof(double) * nsys), "tfs_velocities_create: rhs");
    if (!rc && D.free_bodies) rc = check(hipMalloc(&v->d_sol, sizeof(double) * nsys), "tfs_velocities_create: solution");
    if (!rc && D.free_bodies) rc = check(hipMalloc(&v->d_q, 32LL * D.num_swim), "tfs_velocities_create: orientations");
    if (!rc) rc = check(hipMalloc(&v->d_bodies, sizeof(double) * 9 * v->nswim), "tfs_velocities_create: bodies");
    if (!rc) rc = check(hipMalloc(&v->d_flag, sizeof(int32_t)), "tfs_velocities_create: flag");
    if (!rc) rc = check(hipHostMalloc(&v->h_flag, sizeof(int32_t)), "tfs_velocities_create: flag");
    if (rc) {
        const std::string msg = g_last_error;
        tfs_velocities_destroy(v);
        g_last_error = msg;
        return rc;
    }
    *out = v;
    return TFS_OK;
}

int tfs_velocities_destroy(tfs_velocities* v) {
    if (!v) return TFS_OK;
    if (v->stream) (void)hipStreamSynchronize(v->stream);
    if (v->blobs) tfs_solver_destroy(v->blobs);
    if (v->rpy) tfs_rpy_destroy(v->rpy);
    if (v->d_rhs) (void)hipFree(v->d_rhs);
    if (v->d_bodies) (void)hipFree(v->d_bodies);
    if (v->d_sol) (void)hipFree(v->d_sol);
    if (v->body_refs) (void)hipFree(v->body_refs);
    if (v->Z) (void)hipFree(v->Z);
    if (v->d_q) (void)hipFree(v->d_q);
    if (v->d_flag) (void)hipFree(v->d_flag);
    if (v->Ainv) (void)hipFree(v->Ainv);
    if (v->dense_arena) (void)hipFree(v->dense_arena);
    if (v->h_flag) (void)hipHostFree(v->h_flag);
    if (v->f32_arena) (void)hipFree(v->f32_arena);
    free_stage(v->f32);
    if (v->own_stream) (void)hipStreamDestroy(v->stream);
    delete v;
    return TFS_OK;
}

}  // extern "C"

namespace {

// The shared body of tfs_velocities_compute (prescribed body velocities:
// `bodies` = V, Omega) and tfs_velocities_compute_free (`bodies` = external
// force and torque or NULL; v_bodies_io the body velocities' warm start and
// result).
int compute(tfs_velocities* v, const double* x_segs, const double* f_segs, const double* x_blobs,
            const double* x_bodies, const double* bodies, const double* q_bodies, double* v_segs, double* f_blobs,
            double* v_bodies_io, tfs_solve_stats* stats, bool free_bodies, bool sync_end = true) {
    const hipMemcpyKind kin = v->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    const hipMemcpyKind kout = v->desc.memory == TFS_MEMORY_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    tfs_rpy* h = v->rpy;
    const int64_t ns = v->ns, nb = v->nb;
    hipStream_t st = v->stream;
    // read_positions_and_forces (mobility_solver.cpp:176-236).  Blob positions
    // that did not change keep the blob solve's geometry.
    int rc = TFS_OK;
    bool blobs_moved = !v->have_blobs;
    if (v->have_blobs && v->blob_motion == TFS_BLOBS_MOVING) {
        blobs_moved = true;
    } else if (v->have_blobs && v->blob_motion == TFS_BLOBS_CHECK) {
        if (v->desc.memory == TFS_MEMORY_HOST) {
            std::vector<double> old(3 * nb);
            rc = check(hipMemcpyAsync(old.data(), h->x_blobs, 24 * nb, hipMemcpyDeviceToHost, st), "blob positions");
            if (!rc) rc = check(hipStreamSynchronize(st), "blob positions");
            blobs_moved = !rc && std::memcmp(old.data(), x_blobs, 24 * nb) != 0;
        } else {
            *v->h_flag = 0;
            rc = check(hipMemsetAsync(v->d_flag, 0, sizeof(int32_t), st), "blob flag");
            if (!rc) {
                hipLaunchKernelGGL(differs_kernel, dim3((unsigned)((3 * nb + 255) / 256)), dim3(256), 0, st, x_blobs,
                                   h->x_blobs, 3 * nb, v->d_flag);
                rc = check(hipGetLastError(), "blob flag");
            }
            if (!rc) rc = check(hipMemcpyAsync(v->h_flag, v->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st), "blob flag");
            if (!rc) rc = check(hipStreamSynchronize(st), "blob flag");
            blobs_moved = *v->h_flag != 0;
        }
    }
    // With device memory the segment forces are read in place, the segment
    // velocities are summed straight into the caller's array, and the small
    // inputs arrive in one copy launch; a direct solve needs no warm start.
    const bool dev = v->desc.memory == TFS_MEMORY_DEVICE;
    const bool direct_fixed = v->direct && !free_bodies;
    const double* fs_in = dev && ns ? f_segs : h->f_segs;
    double* vs_out = dev && ns ? v_segs : h->v_segs;
    double* sol = free_bodies ? v->d_sol : h->f_blobs;
    if (!rc && dev) {
        CopyList L;
        L.add(x_blobs, h->x_blobs, 3 * nb);
        if (ns) L.add(x_segs, h->x_segs, 3 * ns);
        if (!v->direct) L.add(f_blobs, sol, 3 * nb);
        if (free_bodies && !v->direct) L.add(v_bodies_io, sol + 3 * nb, 6 * v->nswim);
        L.add(x_bodies, v->d_bodies, 3 * v->nswim);
        if (bodies) L.add(bodies, v->d_bodies + 3 * v->nswim, 6 * v->nswim);
        if (q_bodies) L.add(q_bodies, v->d_q, 4 * v->nswim);
        rc = check(copy_list(L, st), "inputs");
    } else if (!rc) {
        rc = check(hipMemcpyAsync(h->x_blobs, x_blobs, 24 * nb, kin, st), "blob positions");
        if (!rc && ns) rc = check(hipMemcpyAsync(h->x_segs, x_segs, 24 * ns, kin, st), "segment positions");
        if (!rc && ns) rc = check(hipMemcpyAsync(h->f_segs, f_segs, 48 * ns, kin, st), "segment forces");
        if (!rc) rc = check(hipMemcpyAsync(sol, f_blobs, 24 * nb, kin, st), "blob warm start");
        if (!rc && free_bodies)
            rc = check(hipMemcpyAsync(sol + 3 * nb, v_bodies_io, 48 * v->nswim, kin, st), "body warm start");
        if (!rc) rc = check(hipMemcpyAsync(v->d_bodies, x_bodies, 24 * v->nswim, kin, st), "body positions");
        if (!rc && bodies)
            rc = check(hipMemcpyAsync(v->d_bodies + 3 * v->nswim, bodies, 48 * v->nswim, kin, st), "body velocities");
        if (!rc && q_bodies) rc = check(hipMemcpyAsync(v->d_q, q_bodies, 32 * v->nswim, kin, st), "body orientations");
    }
    if (!rc) rc = tfs_rpy_positions_changed(h);
    // Mss_mult '=', Mbs_mult '=' (one launch)
    if (!rc && ns) rc = tfs_rpy_apply(h, TFS_BLOCK_SS | TFS_BLOCK_BS, fs_in, nullptr, vs_out, h->v_blobs);
    // the blob solve's geometry (the direct free-body solve works in the body
    // frame and needs none)
    if (!rc && blobs_moved && !(free_bodies && v->direct))
        rc = tfs_solver_set_geometry(v->blobs, nullptr, h->x_blobs, nullptr, nullptr, TFS_MEMORY_DEVICE);
    if (!rc) v->have_blobs = true;
    if (!rc && blobs_moved) v->direct_valid = false;
    if (!rc && direct_fixed && !v->direct_valid) rc = build_direct(v, v->rpy->x_blobs);
    if (!rc && free_bodies) {
        hipLaunchKernelGGL(free_rhs_kernel, dim3((unsigned)v->nswim), dim3(256), 0, st, h->x_segs, fs_in,
                           h->v_blobs, nb, (int)v->desc.segs_per_swim, v->d_bodies,
                           bodies ? v->d_bodies + 3 * v->nswim : nullptr, v->d_rhs);
        rc = check(hipGetLastError(), "free-body rhs");
        if (!rc && !v->direct) rc = tfs_solver_set_bodies(v->blobs, v->d_bodies, TFS_MEMORY_DEVICE);
    } else if (!rc) {
        // the direct solve reads its right-hand side from the padded vector
        hipLaunchKernelGGL(blob_rhs_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, h->v_blobs,
                           h->x_blobs, v->d_bodies, v->d_bodies + 3 * v->nswim, (int)nb, v->desc.blobs_per_swim,
                           ns > 0 ? 1 : 0, direct_fixed ? v->xpad : v->d_rhs);
        rc = check(hipGetLastError(), "blob rhs");
    }
    int solve_rc = TFS_OK;
    if (!rc && v->direct && free_bodies) {
        // one free body: in its frame with the stored reference inverse
        if (!v->free_direct_valid) rc = build_free_direct(v);
        const unsigned g = (unsigned)((nb + 255) / 256);
        if (!rc) {
            hipLaunchKernelGGL(to_body_frame_kernel, dim3(g), dim3(256), 0, st, (const double*)v->d_rhs, (int)nb,
                               (const double*)v->d_q, v->xpad, v->small + 8);
            rc = check(hipGetLastError(), "body frame");
        }
        if (!rc) rc = check(k_symv_dense(v->Ainv, v->n_pad, v->xpad, v->ypad, v->sym_partial, v->chunks, st), "direct solve");
        if (!rc) {
            hipLaunchKernelGGL(free_direct_velocity_kernel, dim3(1), dim3(256), 0, st, (const double*)v->body_refs,
                               (int)nb, (const double*)v->ypad, (const double*)(v->small + 8),
                               (const double*)v->nref, (const double*)v->d_q, v->small, sol + 3 * nb);
            hipLaunchKernelGGL(free_direct_forces_kernel, dim3(g), dim3(256), 0, st, (const double*)v->ypad,
                               (const double*)v->Z, (int64_t)v->n_pad, (const double*)v->small,
                               (const double*)v->d_q, (int)nb, h->f_blobs);
            rc = check(hipGetLastError(), "direct free-body solve");
        }
        if (!rc) rc = check(hipMemcpyAsync(sol, h->f_blobs, 24 * nb, hipMemcpyDeviceToDevice, st), "blob forces");
        if (stats) {
            *stats = tfs_solve_stats{};
            stats->converged = 1;
        }
    } else if (!rc && v->direct) {
        // blob forces = M_bb^{-1} rhs, left in ypad (its first 3 nb entries)
        rc = check(k_symv_dense(v->Ainv, v->n_pad, v->xpad, v->ypad, v->sym_partial, v->chunks, st), "direct solve");
        if (stats) {
            *stats = tfs_solve_stats{};
            stats->converged = 1;
        }
    } else if (!rc) {
        solve_rc = tfs_solver_solve(v->blobs, v->d_rhs, sol, TFS_MEMORY_DEVICE, stats);
        if (solve_rc != TFS_OK && solve_rc != TFS_ERR_NOT_CONVERGED) rc = solve_rc;
        if (!rc && free_bodies)
            rc = check(hipMemcpyAsync(h->f_blobs, sol, 24 * nb, hipMemcpyDeviceToDevice, st), "blob forces");
    }
    const std::string solve_msg = solve_rc ? g_last_error : std::string();
    const double* fb_solved = direct_fixed ? v->ypad : h->f_blobs;
    // Msb_mult '+='
    if (!rc && ns) rc = tfs_rpy_apply(h, TFS_BLOCK_SB, nullptr, fb_solved, vs_out, nullptr);
    if (!rc && dev) {
        CopyList L;
        L.add(fb_solved, f_blobs, 3 * nb);
        if (free_bodies) L.add(sol + 3 * nb, v_bodies_io, 6 * v->nswim);
        rc = check(copy_list(L, st), "outputs");
    } else if (!rc) {
        if (ns) rc = check(hipMemcpyAsync(v_segs, h->v_segs, 48 * ns, kout, st), "segment velocities");
        if (!rc) rc = check(hipMemcpyAsync(f_blobs, fb_solved, 24 * nb, kout, st), "blob forces");
        if (!rc && free_bodies)
            rc = check(hipMemcpyAsync(v_bodies_io, sol + 3 * nb, 48 * v->nswim, kout, st), "body velocities");
    }
    if (!rc && sync_end) rc = check(hipStreamSynchronize(st), "tfs_velocities_compute");
    if (rc) return rc;
    if (solve_rc) {
        g_last_error = solve_msg;
        return solve_rc;
    }
    return TFS_OK;
}

}  // namespace

extern "C" {

int tfs_velocities_compute(tfs_velocities* v, const double* x_segs, const double* f_segs, const double* x_blobs,
                           const double* x_bodies, const double* v_bodies, double* v_segs, double* f_blobs,
                           tfs_solve_stats* stats) {
    if (!v || !x_blobs || !x_bodies || !v_bodies || !f_blobs || (v->ns && (!x_segs || !f_segs || !v_segs)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute: null argument");
    if (v->desc.free_bodies)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute: free bodies: use tfs_velocities_compute_free");
    return compute(v, x_segs, f_segs, x_blobs, x_bodies, v_bodies, nullptr, v_segs, f_blobs, nullptr, stats, false);
}

int tfs_velocities_compute_free(tfs_velocities* v, const double* x_segs, const double* f_segs, const double* x_blobs,
                                const double* x_bodies, const double* f_bodies, const double* q_bodies, double* v_segs,
                                double* f_blobs, double* v_bodies, tfs_solve_stats* stats) {
    if (!v || !x_blobs || !x_bodies || !v_bodies || !f_blobs || (v->ns && (!x_segs || !f_segs || !v_segs)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute_free: null argument");
    if (!v->desc.free_bodies)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute_free: the handle has no free bodies");
    if (v->direct && (!q_bodies || !v->body_refs))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_velocities_compute_free: the direct solve needs the orientation and "
                    "tfs_velocities_set_body_frame");
    return compute(v, x_segs, f_segs, x_blobs, x_bodies, f_bodies, q_bodies, v_segs, f_blobs, v_bodies, stats, true);
}

}  // extern "C"

namespace {

// compute() on FP32 caller arrays: every input widened into an FP64 device
// image, the FP64 computation on those (device memory), the outputs narrowed
// back into the caller's arrays.
int compute_f32(tfs_velocities* v, const float* x_segs, const float* f_segs, const float* x_blobs,
                const float* x_bodies, const float* bodies, const float* q_bodies, float* v_segs, float* f_blobs,
                float* v_bodies_io, tfs_solve_stats* stats, bool free_bodies, const char* w) {
    const int64_t ns = v->ns, nb = v->nb, nsw = v->nswim;
    const int64_t lens[9] = {3 * ns, 6 * ns, 3 * nb, 3 * nsw, 6 * nsw, 4 * nsw, 6 * ns, 3 * nb, 6 * nsw};
    int64_t total = 0;
    for (int64_t l : lens) total += (std::max<int64_t>(l, 1) + 31) / 32 * 32;
    if (!v->f32_arena && hipMalloc(&v->f32_arena, 8 * total) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, std::string(w) + ": device memory");
    double* img[9];
    double* p = v->f32_arena;
    for (int i = 0; i < 9; ++i) {
        img[i] = p;
        p += (std::max<int64_t>(lens[i], 1) + 31) / 32 * 32;
    }
    const bool host = v->desc.memory == TFS_MEMORY_HOST;
    hipStream_t st = v->stream;
    const float* in[9] = {x_segs, f_segs, x_blobs, x_bodies, bodies, q_bodies, nullptr, f_blobs, v_bodies_io};
    int rc = TFS_OK;
    for (int i = 0; i < 9 && !rc; ++i)
        if (in[i] && lens[i] > 0) rc = widen_in(v->f32, in[i], img[i], lens[i], host, st, w);
    if (rc) return rc;
    const int32_t memory = v->desc.memory;
    v->desc.memory = TFS_MEMORY_DEVICE;
    const int crc = compute(v, ns ? img[0] : nullptr, ns ? img[1] : nullptr, img[2], img[3], bodies ? img[4] : nullptr,
                            q_bodies ? img[5] : nullptr, ns ? img[6] : nullptr, img[7], free_bodies ? img[8] : nullptr,
                            stats, free_bodies);
    v->desc.memory = memory;
    if (crc != TFS_OK && crc != TFS_ERR_NOT_CONVERGED) return crc;
    const std::string msg = g_last_error;
    if (ns) rc = narrow_out(v->f32, img[6], v_segs, 6 * ns, host, st, w);
    if (!rc) rc = narrow_out(v->f32, img[7], f_blobs, 3 * nb, host, st, w);
    if (!rc && free_bodies) rc = narrow_out(v->f32, img[8], v_bodies_io, 6 * nsw, host, st, w);
    if (!rc) rc = check(hipStreamSynchronize(st), w);
    if (rc) return rc;
    g_last_error = msg;
    return crc;
}

}  // namespace

int32_t tfs::velocities_blob_motion(const tfs_velocities* v) { return v->blob_motion; }

int tfs::velocities_compute_ordered(tfs_velocities* v, const double* x_segs, const double* f_segs,
                                    const double* x_blobs, const double* x_bodies, const double* bodies,
                                    const double* q_bodies, double* v_segs, double* f_blobs, double* v_bodies_io,
                                    tfs_solve_stats* stats, bool free_bodies) {
    return compute(v, x_segs, f_segs, x_blobs, x_bodies, bodies, q_bodies, v_segs, f_blobs, v_bodies_io, stats,
                   free_bodies, v->desc.memory != TFS_MEMORY_DEVICE);
}

hipStream_t tfs::velocities_stream(const tfs_velocities* v) { return v->stream; }

tfs_rpy* tfs::velocities_rpy(const tfs_velocities* v) { return v->rpy; }

bool tfs::velocities_graph_ready(const tfs_velocities* v) {
    return v->direct && v->desc.memory == TFS_MEMORY_DEVICE && v->have_blobs && v->blob_motion != TFS_BLOBS_CHECK &&
           (v->desc.free_bodies ? v->free_direct_valid : v->direct_valid);
}

extern "C" {

int tfs_velocities_compute_f32(tfs_velocities* v, const float* x_segs, const float* f_segs, const float* x_blobs,
                               const float* x_bodies, const float* v_bodies, float* v_segs, float* f_blobs,
                               tfs_solve_stats* stats) {
    if (!v || !x_blobs || !x_bodies || !v_bodies || !f_blobs || (v->ns && (!x_segs || !f_segs || !v_segs)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute_f32: null argument");
    if (v->desc.free_bodies)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_velocities_compute_f32: free bodies: use tfs_velocities_compute_free_f32");
    return compute_f32(v, x_segs, f_segs, x_blobs, x_bodies, v_bodies, nullptr, v_segs, f_blobs, nullptr, stats, false,
                       "tfs_velocities_compute_f32");
}

int tfs_velocities_compute_free_f32(tfs_velocities* v, const float* x_segs, const float* f_segs, const float* x_blobs,
                                    const float* x_bodies, const float* f_bodies, const float* q_bodies, float* v_segs,
                                    float* f_blobs, float* v_bodies, tfs_solve_stats* stats) {
    if (!v || !x_blobs || !x_bodies || !v_bodies || !f_blobs || (v->ns && (!x_segs || !f_segs || !v_segs)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute_free_f32: null argument");
    if (!v->desc.free_bodies)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_compute_free_f32: the handle has no free bodies");
    if (v->direct && (!q_bodies || !v->body_refs))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_velocities_compute_free_f32: the direct solve needs the orientation and "
                    "tfs_velocities_set_body_frame");
    return compute_f32(v, x_segs, f_segs, x_blobs, x_bodies, f_bodies, q_bodies, v_segs, f_blobs, v_bodies, stats, true,
                       "tfs_velocities_compute_free_f32");
}

int tfs_velocities_set_blob_motion(tfs_velocities* v, int32_t motion) {
    if (!v || (motion != TFS_BLOBS_CHECK && motion != TFS_BLOBS_FIXED && motion != TFS_BLOBS_MOVING))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_set_blob_motion: null handle or unknown mode");
    v->blob_motion = motion;
    return TFS_OK;
}

int tfs_velocities_set_body_frame(tfs_velocities* v, const double* blob_refs) {
    if (!v || !blob_refs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_set_body_frame: null argument");
    if (!v->desc.free_bodies) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_set_body_frame: no free bodies");
    if (!v->body_refs && hipMalloc(&v->body_refs, 24 * v->nb) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_velocities_set_body_frame: device memory");
    const hipMemcpyKind kin = v->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    int rc = check(hipMemcpyAsync(v->body_refs, blob_refs, 24 * v->nb, kin, v->stream), "tfs_velocities_set_body_frame");
    if (!rc) rc = check(hipStreamSynchronize(v->stream), "tfs_velocities_set_body_frame");
    v->free_direct_valid = false;
    return rc;
}

int tfs_velocities_set_preconditioner(tfs_velocities* v, int32_t kind) {
    if (!v) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_velocities_set_preconditioner: null handle");
    if (kind == TFS_BLOB_DIRECT) {
        if (v->desc.free_bodies && v->nswim != 1)
            return fail(TFS_ERR_INVALID_ARGUMENT,
                        "tfs_velocities_set_preconditioner: the direct solve takes one free body");
        v->direct = true;
        return TFS_OK;
    }
    const int rc = tfs_solver_set_preconditioner(v->blobs, kind);
    if (!rc) v->direct = false;
    return rc;
}

}  // extern "C"
