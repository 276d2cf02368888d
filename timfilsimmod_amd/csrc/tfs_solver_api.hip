// tfs_solver_api.hip -- device-resident per-timestep solve behind include/tfs_solver.h.
//
// The reference solves its saddle-point system on the host: every GMRES
// iteration copies the Krylov vector to the GPU, runs the four RPY kernels,
// copies velocities back and does modified Gram-Schmidt on host columns
// (mobility_solver.cpp:1806-2090, 2335-2419).  Here every vector lives in HBM
// and only one scalar -- the residual estimate -- crosses to the host per
// iteration.  Arnoldi stays in FP64 (CGS2 as GEMVs over the basis); the RPY
// matvec runs in FP64 tiles or, inside mixed-precision iterative refinement,
// in FP32 tiles with FP64 accumulation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tfs_solver.h"
#include "solver_kernels.h"
#include "dense_kernels.h"
#include "tfs_api_internal.h"

using namespace tfs;

namespace {

constexpr int kMaxRestart = 350;
// GMRES copies its device stop flag back every kEstBatch iterations and the
// host looks one batch behind; up to 2*kEstBatch - 1 iterations enqueued past
// convergence run as empty launches.
constexpr int kEstBatch = 4;
// Mixed-precision refinement: the inner FP32-tile solve of a correction d
// stops at max(inner_tol ||r||, kInnerFloor abs_tol).  The FP64 residual of
// x + d then exceeds the inner residual only by the FP32 operator's error on
// d, ~1e-7 ||A d|| ~ 1e-7 ||r||, far inside the remaining (1 - kInnerFloor)
// abs_tol, so one refinement meets the outer tolerance.
constexpr double kInnerFloor = 0.8;

__global__ void fill_kernel(double* p, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t k_fill_int(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t k_fill(double* p, double v, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace

// Independent bodies solved in lock-step (tfs_solver_set_independent_bodies):
// one GMRES per body, every batched launch covering all bodies.  The operator
// (matvec, couplings, preconditioner) works on the usual layout over all
// bodies -- [lambda_s | lambda_b | phi_dot | theta_dot], each part body after
// body -- and the Krylov vectors live in the bodies' own layout, body b's
// unknowns contiguous at koff[b]; map[i] is the operator index of Krylov
// index i.
struct EnsState {
    int nbody = 0;
    std::vector<int32_t> fil, blob;   // per body
    std::vector<int64_t> koff;        // nbody + 1
    int64_t ntot = 0;
    int G = 1;                        // workgroups per body of the 2-D launches
    void* arena = nullptr;
    int64_t* d_koff = nullptr;
    int32_t *d_map = nullptr, *d_fil_body = nullptr;
    double *V = nullptr, *w = nullptr, *tk = nullptr, *rk = nullptr;
    int64_t ldv = 0;
    double *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *est = nullptr, *y = nullptr, *scal = nullptr;
    double *abs_tol = nullptr, *norms = nullptr, *pa = nullptr, *pb = nullptr;
    int64_t hstride = 0, pstride = 0;
    int mstride = 0;
    int32_t *stop = nullptr, *nstop = nullptr, *kdone = nullptr, *flag = nullptr;
    int64_t* stamps = nullptr;        // nbody + 1: per-body last stop, then the step's start
    int32_t* h_i32 = nullptr;         // pinned: 2 nbody + 8
    double* h_f64 = nullptr;          // pinned: nbody
    int64_t* h_stamps = nullptr;      // pinned: nbody + 1
    // per-body results of the last solve
    std::vector<int32_t> iters, conv;
    std::vector<double> relres;
    // captured cycles (one per precision) and the pointers / values they hold
    bool use_graphs = true;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<uintptr_t> key[2];
    int64_t graph_launches = 0, eager_cycles = 0;
};

struct tfs_solver {
    tfs_solver_desc desc{};
    SaddleDims dims{};
    tfs_rpy* rpy = nullptr;       // seg_dof 3, Mbs accumulating (prescribed cilia)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int m = 50;                   // restart length
    int G = 256;                  // reduction blocks
    // device memory
    double* arena = nullptr;
    double *kphi = nullptr, *kth = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr, *w = nullptr, *z = nullptr, *t = nullptr, *d = nullptr,
           *rin = nullptr;
    double* V = nullptr;          // n x (m+1), column-major, ld = ldv
    int64_t ldv = 0;
    double *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr, *h = nullptr;
    double *partial = nullptr, *scal = nullptr, *nrm = nullptr, *est = nullptr;
    double *pa = nullptr, *pb = nullptr;  // per-workgroup partial sums of the Arnoldi step
    int GA = 1;                           // workgroups of the Arnoldi step
    int32_t* stop = nullptr;              // device stop flag of the current GMRES cycle
    // The GMRES stop flag as the close kernel mirrors it: pinned host memory
    // mapped for the device (host_stop) and its device address.
    int32_t* host_stop = nullptr;
    int32_t* dev_stop_mirror = nullptr;
    // The residual estimate of the last step the close kernel closed, mirrored
    // the same way (the cycle's final estimate once its launches are done).
    double* host_est = nullptr;
    double* dev_est_mirror = nullptr;
    hipEvent_t flag_ev[2] = {nullptr, nullptr};
    double* host_scal = nullptr;  // pinned, kEstBatch + 8 doubles
    double gs = 0, gb = 0;        // 6 pi mu a for segments / blobs
    // Block preconditioner (kind 1): filament blocks rebuilt when the
    // segments move, blob patches when the blobs move.
    int precond_kind = 0;
    bool fil_pc_valid = false, blob_pc_valid = false;
    double *fil_inv = nullptr, *fil_dk = nullptr, *fil_sinv = nullptr, *patch_inv = nullptr;
    int64_t *fil_off = nullptr, *patch_off = nullptr;
    int32_t *fil_size = nullptr, *fil_piv = nullptr, *fil_pivoff = nullptr, *fil_status = nullptr;
    int32_t *patch_blob = nullptr, *patch_start = nullptr, *patch_size = nullptr, *patch_piv = nullptr,
            *patch_pivoff = nullptr, *patch_status = nullptr;
    int npatch = 0;
    // Free bodies: centres, Z = P0^{-1} E (6 columns of ldv), the 6 x 6
    // inverses of E^T Z per swimmer, and a 6-per-swimmer work vector.
    double *x_bodies = nullptr, *bodyZ = nullptr, *body_sinv = nullptr, *body_t = nullptr;
    bool bodies_set = false, body_pc_valid = false;
    tfs::F32Stage f32;  // FP32 entry points (tfs_solver_*_f32)
    EnsState* ens = nullptr;  // independent bodies in lock-step (null: one system)
};

namespace {

int check(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }

// Read one device double (synchronises the stream).
int read_scalar(tfs_solver* s, const double* dev, double* out) {
    hipError_t e = hipMemcpyAsync(s->host_scal, dev, sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "read_scalar");
    *out = *s->host_scal;
    return TFS_OK;
}

int norm(tfs_solver* s, const double* v, double* out) {
    int rc = check(k_norm(v, (int)s->dims.n, s->partial, s->G, s->nrm, s->stream), "norm");
    if (rc) return rc;
    return read_scalar(s, s->nrm, out);
}

// out = A in.  With `skip`, every launch does nothing while *skip != 0.
int apply_A(tfs_solver* s, const double* in, double* out, int precision, tfs_solve_stats* st,
            const int32_t* skip = nullptr) {
    const SaddleDims& dm = s->dims;
    if (dm.free_body && !s->bodies_set) return fail(TFS_ERR_INVALID_ARGUMENT, "free bodies: tfs_solver_set_bodies not called");
    // The K_phi / K_theta couplings are applied by the matvec's reduce kernel.
    SaddleEpilogue sad;
    const bool couplings = (dm.nfil > 0 && (dm.dyn_phase || dm.dyn_angle)) || dm.free_body;
    if (couplings) {
        sad.in = in;
        sad.out = out;
        sad.kphi = s->kphi;
        sad.kth = s->kth;
        sad.nfil = dm.nfil;
        sad.nseg = dm.nseg;
        sad.dyn_phase = dm.dyn_phase;
        sad.dyn_angle = dm.dyn_angle;
        sad.off_phase = dm.off_phase;
        sad.off_angle = dm.off_angle;
        if (s->ens) sad.fil_body = s->ens->d_fil_body;
        if (dm.free_body) {
            sad.x_bodies = s->x_bodies;
            sad.free_body = 1;
            sad.segs_per_swim = dm.segs_per_swim;
            sad.blobs_per_swim = dm.blobs_per_swim;
            sad.off_body = dm.off_body;
        }
    }
    if (dm.free_body) {
        // body rows: -K_b^T lambda (mobility_solver.cpp:1850-1908)
        int rc0 = check(k_body_moments(in, dm, s->rpy->x_segs, s->rpy->x_blobs, s->x_bodies, -1.0, out + dm.off_body,
                                       skip, s->stream),
                        "body rows");
        if (rc0) return rc0;
    }
    s->rpy->skip_flag = skip;
    int rc = run_matvec(s->rpy, TFS_BLOCK_ALL, precision, CView{in, 3, 1}, CView{in + dm.off_blob, 3, 1},
                        View{out, 3, 1}, View{out + dm.off_blob, 3, 1}, 3, couplings ? &sad : nullptr);
    s->rpy->skip_flag = nullptr;
    if (rc) return rc;
    if (st) {
        if (precision == TFS_PRECISION_FP64) st->fp64_matvecs++;
        else st->fp32_matvecs++;
    }
    return TFS_OK;
}

PrecondArgs precond_args(const tfs_solver* s) {
    PrecondArgs P;
    P.kind = s->precond_kind;
    P.gs = s->gs;
    P.gb = s->gb;
    P.kphi = s->kphi;
    P.kth = s->kth;
    P.fil_inv = s->fil_inv;
    P.fil_dk = s->fil_dk;
    P.fil_sinv = s->fil_sinv;
    P.patch_blob = s->patch_blob;
    P.patch_start = s->patch_start;
    P.patch_inv = s->patch_inv;
    P.patch_off = s->patch_off;
    P.npatch = s->npatch;
    return P;
}

// Device buffer of n elements of T (freed with the handle).
template <typename T>
int dev_alloc(T** p, size_t n, const char* what) {
    if (*p) return TFS_OK;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
    return e == hipSuccess ? TFS_OK : fail(TFS_ERR_OUT_OF_MEMORY, std::string("block preconditioner: ") + what);
}

// Status flags of a batched inverse: any singular or non-finite block is an error.
int check_inverse_status(tfs_solver* s, const int32_t* status, int n, const char* what) {
    std::vector<int32_t> st(n);
    hipError_t e = hipMemcpyAsync(st.data(), status, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, what);
    for (int i = 0; i < n; ++i)
        if (st[i]) return fail(TFS_ERR_NON_FINITE, std::string(what) + ": block " + std::to_string(i) + " is singular");
    return TFS_OK;
}

// Morton (Z-order) code of a point quantised to 10 bits per axis in [lo, hi].
uint32_t morton3(const double* x, const double* lo, const double* span) {
    uint32_t code = 0;
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
        const double t = span[d] > 0 ? (x[d] - lo[d]) / span[d] : 0.0;
        q[d] = (uint32_t)std::min(1023.0, std::max(0.0, t * 1023.0));
    }
    for (int b = 0; b < 10; ++b)
        for (int d = 0; d < 3; ++d) code |= ((q[d] >> b) & 1u) << (3 * b + d);
    return code;
}

// Assembles blocks (by `assemble`) and inverts them in place: the one-wave SPD
// elimination when they have at most 64 rows, else -- or if a pivot was not
// positive -- re-assembly and Gauss-Jordan with partial pivoting.  Any block
// still singular is an error.  Synchronises the stream (status read-back).
template <typename F>
int invert_blocks(tfs_solver* s, F assemble, double* blk, const int64_t* off, const int32_t* sizes, int n,
                  int max_size, int32_t* piv, const int32_t* pivoff, int32_t* status, const char* what) {
    hipError_t e = assemble();
    if (e != hipSuccess) return hip_fail(e, what);
    if (max_size <= 64) {
        e = k_batched_inverse_spd(blk, off, sizes, n, max_size, status, s->stream);
        if (e != hipSuccess) return hip_fail(e, what);
        std::vector<int32_t> st(n);
        e = hipMemcpyAsync(st.data(), status, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_fail(e, what);
        bool ok = true;
        for (int i = 0; i < n && ok; ++i) ok = st[i] == 0;
        if (ok) return TFS_OK;
        e = assemble();
        if (e != hipSuccess) return hip_fail(e, what);
    }
    e = k_batched_inverse(blk, off, sizes, n, max_size, piv, pivoff, status, s->stream);
    if (e != hipSuccess) return hip_fail(e, what);
    return check_inverse_status(s, status, n, what);
}

// Builds whatever part of the block preconditioner is stale.
int ensure_block_precond(tfs_solver* s) {
    if (s->precond_kind != 1) return TFS_OK;
    const SaddleDims& d = s->dims;
    const tfs_solver_desc& D = s->desc;
    int rc = TFS_OK;
    if (!s->blob_pc_valid && d.nblob > 0) {
        // patches: consecutive kPatchBlobs blobs in Morton order of the positions
        std::vector<double> xb(3 * (size_t)d.nblob);
        hipError_t e = hipMemcpyAsync(xb.data(), s->rpy->x_blobs, xb.size() * 8, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) return hip_fail(e, "block preconditioner: blob positions");
        // Morton order within each body (independent bodies: their own boxes,
        // and no patch holds blobs of two bodies)
        std::vector<int32_t> first(1, 0);
        if (s->ens)
            for (int32_t c : s->ens->blob) first.push_back(first.back() + c);
        else
            first.push_back(d.nblob);
        std::vector<std::pair<uint64_t, int32_t>> key(d.nblob);
        for (size_t bi = 0; bi + 1 < first.size(); ++bi) {
            const int i0 = first[bi], i1 = first[bi + 1];
            if (i1 <= i0) continue;
            double lo[3], hi[3], span[3];
            for (int k = 0; k < 3; ++k) lo[k] = hi[k] = xb[3 * (size_t)i0 + k];
            for (int i = i0; i < i1; ++i)
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], xb[3 * i + k]);
                    hi[k] = std::max(hi[k], xb[3 * i + k]);
                }
            for (int k = 0; k < 3; ++k) span[k] = hi[k] - lo[k];
            for (int i = i0; i < i1; ++i) key[i] = {((uint64_t)bi << 32) | morton3(&xb[3 * i], lo, span), i};
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
                             return a.first < b.first;
                         });
        // patches of kPatchBlobs consecutive blobs, a new one at every body
        std::vector<int32_t> blob(d.nblob), start, size, pivoff;
        std::vector<int64_t> off;
        int64_t total = 0;
        for (int i = 0; i < d.nblob; ++i) blob[i] = key[i].second;
        for (int i = 0; i < d.nblob;) {
            int j = i + 1;
            while (j < d.nblob && j - i < kPatchBlobs && (key[j].first >> 32) == (key[i].first >> 32)) ++j;
            start.push_back(i);
            size.push_back(3 * (j - i));
            off.push_back(total);
            pivoff.push_back(3 * i);
            total += (int64_t)size.back() * size.back();
            i = j;
        }
        const int np = (int)start.size();
        start.push_back(d.nblob);
        if (s->npatch != np) {
            for (void* q : {(void*)s->patch_inv, (void*)s->patch_off, (void*)s->patch_blob, (void*)s->patch_start,
                            (void*)s->patch_size, (void*)s->patch_piv, (void*)s->patch_pivoff, (void*)s->patch_status})
                if (q) (void)hipFree(q);
            s->patch_inv = nullptr;
            s->patch_off = nullptr;
            s->patch_blob = s->patch_start = s->patch_size = s->patch_piv = s->patch_pivoff = s->patch_status = nullptr;
        }
        rc = dev_alloc(&s->patch_inv, (size_t)total, "patch blocks");
        if (!rc) rc = dev_alloc(&s->patch_off, np, "patch offsets");
        if (!rc) rc = dev_alloc(&s->patch_blob, d.nblob, "patch blobs");
        if (!rc) rc = dev_alloc(&s->patch_start, np + 1, "patch starts");
        if (!rc) rc = dev_alloc(&s->patch_size, np, "patch sizes");
        if (!rc) rc = dev_alloc(&s->patch_piv, 3 * (size_t)d.nblob, "patch pivots");
        if (!rc) rc = dev_alloc(&s->patch_pivoff, np, "patch pivot offsets");
        if (!rc) rc = dev_alloc(&s->patch_status, np, "patch status");
        if (rc) return rc;
        s->npatch = np;
        e = hipMemcpyAsync(s->patch_blob, blob.data(), sizeof(int32_t) * d.nblob, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->patch_start, start.data(), sizeof(int32_t) * (np + 1), hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->patch_size, size.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->patch_off, off.data(), sizeof(int64_t) * np, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->patch_pivoff, pivoff.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s->stream);
        if (e != hipSuccess) return hip_fail(e, "block preconditioner: patch tables");
        rc = invert_blocks(s, [&]() {
            return k_patch_blocks(s->rpy->x_blobs, s->patch_blob, s->patch_start, s->patch_off, np, D.blob_radius,
                                  D.mu, s->patch_inv, s->stream);
        }, s->patch_inv, s->patch_off, s->patch_size, np, 3 * kPatchBlobs, s->patch_piv, s->patch_pivoff,
                           s->patch_status, "block preconditioner: blob patch");
        if (rc) return rc;
        s->blob_pc_valid = true;
    }
    if (!s->fil_pc_valid && d.nfil > 0) {
        const int len = 3 * d.nseg;
        rc = dev_alloc(&s->fil_inv, (size_t)d.nfil * len * len, "filament blocks");
        if (!rc) rc = dev_alloc(&s->fil_dk, (size_t)d.nfil * 2 * len, "filament couplings");
        if (!rc) rc = dev_alloc(&s->fil_sinv, (size_t)d.nfil * 4, "filament Schur complements");
        if (!rc) rc = dev_alloc(&s->fil_piv, (size_t)d.nfil * len, "filament pivots");
        if (!rc && !s->fil_off) {
            std::vector<int64_t> off(d.nfil);
            std::vector<int32_t> size(d.nfil, len), pivoff(d.nfil);
            for (int m = 0; m < d.nfil; ++m) {
                off[m] = (int64_t)m * len * len;
                pivoff[m] = m * len;
            }
            rc = dev_alloc(&s->fil_off, d.nfil, "filament offsets");
            if (!rc) rc = dev_alloc(&s->fil_size, d.nfil, "filament sizes");
            if (!rc) rc = dev_alloc(&s->fil_pivoff, d.nfil, "filament pivot offsets");
            if (!rc) rc = dev_alloc(&s->fil_status, d.nfil, "filament status");
            hipError_t e = rc ? hipSuccess
                              : hipMemcpyAsync(s->fil_off, off.data(), sizeof(int64_t) * d.nfil,
                                               hipMemcpyHostToDevice, s->stream);
            if (!rc && e == hipSuccess)
                e = hipMemcpyAsync(s->fil_size, size.data(), sizeof(int32_t) * d.nfil, hipMemcpyHostToDevice,
                                   s->stream);
            if (!rc && e == hipSuccess)
                e = hipMemcpyAsync(s->fil_pivoff, pivoff.data(), sizeof(int32_t) * d.nfil, hipMemcpyHostToDevice,
                                   s->stream);
            // the host vectors must outlive the copies
            if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->stream);
            if (!rc && e != hipSuccess) rc = hip_fail(e, "block preconditioner: filament tables");
        }
        if (rc) return rc;
        rc = invert_blocks(s, [&]() {
            return k_fil_blocks(s->rpy->x_segs, d.nfil, d.nseg, D.seg_radius, D.mu, s->fil_inv, s->stream);
        }, s->fil_inv, s->fil_off, s->fil_size, d.nfil, len, s->fil_piv, s->fil_pivoff, s->fil_status,
                           "block preconditioner: filament");
        if (rc) return rc;
        rc = check(k_fil_schur(precond_args(s), d, s->fil_dk, s->fil_sinv, s->stream), "block preconditioner: Schur");
        if (rc) return rc;
        s->fil_pc_valid = true;
    }
    return TFS_OK;
}

// The free-body part of P^{-1} after P0^{-1} wrote z (see k_body_correct).
int body_correct(tfs_solver* s, const double* in, double host_sc, const double* p_in, int G, double* v, double* z,
                 const int32_t* skip) {
    const SaddleDims& d = s->dims;
    if (!d.free_body) return TFS_OK;
    int rc = check(k_body_moments(z, d, s->rpy->x_segs, s->rpy->x_blobs, s->x_bodies, 1.0, s->body_t, skip, s->stream),
                   "body moments");
    if (!rc)
        rc = check(k_body_correct(in, host_sc, p_in, G, v, z, s->bodyZ, s->ldv, s->body_t, s->body_sinv, d, skip,
                                  s->stream),
                   "body correction");
    return rc;
}

int apply_P(tfs_solver* s, const double* in, double* out) {
    int rc = check(k_precond(in, out, precond_args(s), s->dims, s->stream), "preconditioner");
    if (!rc) rc = body_correct(s, in, 1.0, nullptr, 0, nullptr, out, nullptr);
    return rc;
}

// Z = P0^{-1} E and the inverses of E^T Z (one 6 x 6 per swimmer), rebuilt
// when the geometry or P0 changed.
int ensure_body_precond(tfs_solver* s) {
    const SaddleDims& d = s->dims;
    if (!d.free_body || s->body_pc_valid) return TFS_OK;
    if (!s->bodies_set) return fail(TFS_ERR_INVALID_ARGUMENT, "free bodies: tfs_solver_set_bodies not called");
    const int nsw = d.nswim;
    std::vector<double> sig(36 * (size_t)nsw);
    int rc = TFS_OK;
    for (int c = 0; c < 6 && !rc; ++c) {
        double* Zc = s->bodyZ + (int64_t)c * s->ldv;
        rc = check(k_body_column(s->t, c, d, s->rpy->x_segs, s->rpy->x_blobs, s->x_bodies, s->stream), "body column");
        if (!rc) rc = check(k_fill(Zc, 0.0, d.n, s->stream), "body Z");
        if (!rc) rc = check(k_precond(s->t, Zc, precond_args(s), d, s->stream), "body Z");
        if (!rc) rc = check(k_body_moments(Zc, d, s->rpy->x_segs, s->rpy->x_blobs, s->x_bodies, 1.0, s->body_t, nullptr,
                                           s->stream),
                            "body Schur");
        std::vector<double> col(6 * (size_t)nsw);
        if (!rc) rc = check(hipMemcpyAsync(col.data(), s->body_t, 48 * (size_t)nsw, hipMemcpyDeviceToHost, s->stream),
                            "body Schur");
        if (!rc) rc = check(hipStreamSynchronize(s->stream), "body Schur");
        for (int w = 0; w < nsw && !rc; ++w)
            for (int r = 0; r < 6; ++r) sig[36 * (size_t)w + 6 * r + c] = col[6 * (size_t)w + r];
    }
    if (rc) return rc;
    // invert each 6 x 6 (Gauss-Jordan, partial pivoting)
    std::vector<double> inv(36 * (size_t)nsw);
    for (int w = 0; w < nsw; ++w) {
        double a[6][12];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 12; ++c) a[r][c] = c < 6 ? sig[36 * (size_t)w + 6 * r + c] : (c - 6 == r ? 1.0 : 0.0);
        for (int k = 0; k < 6; ++k) {
            int piv = k;
            for (int r = k + 1; r < 6; ++r)
                if (std::fabs(a[r][k]) > std::fabs(a[piv][k])) piv = r;
            if (!(std::fabs(a[piv][k]) > 0.0) || !std::isfinite(a[piv][k]))
                return fail(TFS_ERR_NON_FINITE, "free bodies: singular body Schur complement");
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
            for (int c = 0; c < 6; ++c) inv[36 * (size_t)w + 6 * r + c] = a[r][6 + c];
    }
    rc = check(hipMemcpyAsync(s->body_sinv, inv.data(), 288 * (size_t)nsw, hipMemcpyHostToDevice, s->stream),
               "body Schur inverse");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "body Schur inverse");
    if (!rc) s->body_pc_valid = true;
    return rc;
}

// r = b - A x (x may be known to be zero: then r = b).
int residual(tfs_solver* s, const double* b, const double* x, bool x_zero, double* r, int precision,
             tfs_solve_stats* st) {
    const int n = (int)s->dims.n;
    if (x_zero) return check(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream), "copy");
    int rc = apply_A(s, x, s->t, precision, st);
    if (rc) return rc;
    return check(k_sub(b, s->t, r, n, s->stream), "residual");
}

// One or more GMRES(m) cycles on A x = b with matvecs in `precision`, until
// ||b - A x|| <= abs_tol or max_cycles.  x is updated in place.  On entry r
// holds b - A x and rnorm its norm.  Returns TFS_OK (converged or not: see
// *converged), or an error.
int gmres(tfs_solver* s, int precision, const double* b, double* x, double* r, double rnorm, double abs_tol,
          int max_cycles, tfs_solve_stats* st, int* converged, double* final_res) {
    const int n = (int)s->dims.n;
    const int m = s->m;
    const int ldh = m + 1;
    // FP64 solves orthogonalise twice (CGS2, orthogonal to working precision);
    // the inner solve of mixed refinement, whose result the FP64 residual
    // checks, once.
    const int gs_passes = precision == TFS_PRECISION_FP64 ? 2 : 1;
    *converged = rnorm <= abs_tol;
    *final_res = rnorm;
    for (int cyc = 0; cyc < max_cycles && !*converged; ++cyc) {
        if (cyc > 0) {
            int rc = residual(s, b, x, false, r, precision, st);
            if (!rc) rc = norm(s, r, &rnorm);
            if (rc) return rc;
            if (!std::isfinite(rnorm)) return fail(TFS_ERR_NON_FINITE, "GMRES: non-finite residual");
            *final_res = rnorm;
            if (rnorm <= abs_tol) {
                *converged = 1;
                break;
            }
        }
        st->cycles++;
        int rc = check(k_fill(s->H, 0.0, (int64_t)ldh * m, s->stream), "H");
        if (!rc) rc = check(k_fill(s->g, 0.0, ldh, s->stream), "g");
        if (!rc) rc = check(k_fill(s->g, rnorm, 1, s->stream), "g0");
        // v0 = r / ||r||
        if (!rc) rc = check(k_axpby(1.0 / rnorm, r, 0.0, s->V, n, s->stream), "v0");
        if (rc) return rc;
        // The iterations are enqueued ahead of the host: every kEstBatch of
        // them an event is recorded, and the host waits for the event of the
        // batch before (so one batch is always in flight) and then reads the
        // stop flag that the close kernel mirrors into host memory.  Launches
        // enqueued after the step that met the tolerance see the flag and do
        // nothing.  (The previous cycle's work has finished: the mirror is the
        // host's to clear.)
        if (!rc) rc = check(k_fill_int(s->stop, 0, 1, s->stream), "stop flag");
        if (rc) return rc;
        __atomic_store_n(s->host_stop, 0, __ATOMIC_SEQ_CST);
        int enq = 0, batches = 0;
        bool seen = false;
        while (enq < m && !seen) {
            const int j = enq;
            double* vj = s->V + (int64_t)j * s->ldv;
            // v_j = w_{j-1} / ||w_{j-1}|| (v_0 = r / ||r||), z = P^{-1} v_j; for
            // j > 0 the same launch closes step j - 1.
            if (j == 0) {
                rc = check(k_scale_precond(r, nullptr, 1.0 / rnorm, vj, s->z, precond_args(s), s->dims, nullptr,
                                           s->stream), "v0");
                if (!rc) rc = body_correct(s, r, 1.0 / rnorm, nullptr, 0, vj, s->z, nullptr);
            } else {
                rc = check(k_arnoldi_close_next(s->pa, s->GA, s->H, ldh, j - 1, s->cs, s->sn, s->g, s->scal, s->est,
                                                s->stop, s->dev_stop_mirror, s->dev_est_mirror, abs_tol, s->w, vj,
                                                s->z, precond_args(s),
                                                s->dims, s->stream),
                           "close and normalise");
                if (!rc) rc = body_correct(s, s->w, 1.0, s->pa, s->GA, vj, s->z, s->stop);
            }
            if (!rc) rc = apply_A(s, s->z, s->w, precision, nullptr, s->stop);
            if (!rc)
                rc = check(k_arnoldi_step(s->V, s->ldv, j, s->w, n, s->pa, s->pb, s->GA, s->H, ldh, s->cs, s->sn,
                                          s->g, s->scal, s->est, s->stop, abs_tol, gs_passes, s->stream),
                           "arnoldi");
            if (rc) return rc;
            ++enq;
            if (enq % kEstBatch != 0 && enq < m) continue;
            const int slot = batches & 1;
            hipError_t e = hipEventRecord(s->flag_ev[slot], s->stream);
            if (e == hipSuccess && batches > 0) {
                e = hipEventSynchronize(s->flag_ev[slot ^ 1]);
                if (e == hipSuccess && __atomic_load_n(s->host_stop, __ATOMIC_SEQ_CST) != 0) seen = true;
            }
            if (e != hipSuccess) return hip_fail(e, "GMRES stop flag");
            ++batches;
        }
        // close the last enqueued step
        rc = check(k_arnoldi_close_next(s->pa, s->GA, s->H, ldh, enq - 1, s->cs, s->sn, s->g, s->scal, s->est, s->stop,
                                        s->dev_stop_mirror, s->dev_est_mirror, abs_tol, s->w, nullptr, nullptr, precond_args(s), s->dims,
                                        s->stream),
                   "close");
        if (rc) return rc;
        // Once the launches are done, the mirrors hold the stop value and the
        // estimate of the last step closed: step k - 1, where the cycle
        // stopped at k or, unstopped, the last one enqueued (closes after a
        // stop do nothing).
        int32_t flag = 0;
        double est = rnorm;
        {
            const hipError_t e = hipStreamSynchronize(s->stream);
            if (e != hipSuccess) return hip_fail(e, "GMRES stop flag");
            flag = __atomic_load_n(s->host_stop, __ATOMIC_SEQ_CST);
            __atomic_load(s->host_est, &est, __ATOMIC_SEQ_CST);
        }
        if (flag < 0) return fail(TFS_ERR_NON_FINITE, "GMRES: non-finite residual estimate");
        const int k = flag > 0 ? flag : enq;
        st->iterations += k;
        if (precision == TFS_PRECISION_FP64) st->fp64_matvecs += k;
        else st->fp32_matvecs += k;
        // x += P^{-1} V(:, 0:k) y
        rc = check(k_trisolve(s->H, ldh, s->g, k, s->y, s->stream), "trisolve");
        if (!rc) rc = check(k_gemv(s->V, s->ldv, k, s->y, s->t, n, false, s->stream), "Vy");
        if (!rc) rc = apply_P(s, s->t, s->z);
        if (!rc) rc = check(k_axpby(1.0, s->z, 1.0, x, n, s->stream), "update");
        if (rc) return rc;
        *final_res = est;
        if (est <= abs_tol) *converged = 1;
    }
    return TFS_OK;
}

// ---- independent bodies in lock-step --------------------------------------

EnsArgs ens_args(tfs_solver* s) {
    EnsState& E = *s->ens;
    EnsArgs a;
    a.koff = E.d_koff;
    a.nbody = E.nbody;
    a.V = E.V;
    a.ld = E.ldv;
    a.w = E.w;
    a.H = E.H;
    a.ldh = s->m + 1;
    a.hstride = E.hstride;
    a.cs = E.cs;
    a.sn = E.sn;
    a.g = E.g;
    a.est = E.est;
    a.mstride = E.mstride;
    a.scal = E.scal;
    a.stop = E.stop;
    a.nstop = E.nstop;
    a.abs_tol = E.abs_tol;
    a.pa = E.pa;
    a.pb = E.pb;
    a.pstride = E.pstride;
    a.stamps = E.stamps;
    return a;
}

// Per-body norms of an operator-layout vector, read back into norms[].
int ens_norms(tfs_solver* s, const double* v, std::vector<double>& norms) {
    EnsState& E = *s->ens;
    int rc = check(k_ens_permute(v, E.rk, E.d_map, E.ntot, false, s->stream), "per-body norms");
    if (!rc) rc = check(k_ens_norms(E.rk, E.d_koff, E.nbody, E.norms, s->stream), "per-body norms");
    if (!rc) rc = check(hipMemcpyAsync(E.h_f64, E.norms, 8LL * E.nbody, hipMemcpyDeviceToHost, s->stream), "per-body norms");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "per-body norms");
    if (rc) return rc;
    norms.assign(E.h_f64, E.h_f64 + E.nbody);
    return TFS_OK;
}

// Lock-step steps j0 <= j < j1 of the current cycle (no host round trip):
// w = A z with the stopped bodies' work skipped, w to the Krylov layout, the
// Gram-Schmidt step (twice in FP64 solves, once in the inner solve of mixed
// refinement) and close of every running body, v_{j+1} and z = P^{-1} v_{j+1}.
int ens_enqueue_steps(tfs_solver* s, int precision, int j0, int j1) {
    EnsState& E = *s->ens;
    const EnsArgs a = ens_args(s);
    const int m = s->m;
    for (int j = j0; j < j1; ++j) {
        s->rpy->body_skip = E.stop;
        int rc = apply_A(s, s->z, s->w, precision, nullptr);
        s->rpy->body_skip = nullptr;
        if (!rc) rc = check(k_ens_permute(s->w, E.w, E.d_map, E.ntot, false, s->stream), "w to the Krylov layout");
        if (!rc)
            rc = check(k_ens_arnoldi_step(a, j, E.G, precision == TFS_PRECISION_FP64 ? 2 : 1, s->stream),
                       "lock-step Arnoldi");
        if (!rc) rc = check(k_ens_close(a, j, j + 1 < m, E.G, s->stream), "lock-step close");
        if (!rc && j + 1 < m) {
            rc = check(k_ens_permute(E.V + (int64_t)(j + 1) * E.ldv, s->t, E.d_map, E.ntot, true, s->stream),
                       "v to the operator layout");
            if (!rc) rc = check(k_precond(s->t, s->z, precond_args(s), s->dims, s->stream), "preconditioner");
        }
        if (rc) return rc;
    }
    return TFS_OK;
}

// The m steps of a cycle as one hipGraph, captured again only when something
// the launches hold by value or pointer changed (the work list, partial
// buffer, near-pair lists, fixed-point positions and their scale).
int ens_cycle_graph(tfs_solver* s, int precision) {
    EnsState& E = *s->ens;
    tfs_rpy* h = s->rpy;
    int rc = ensure_near(h, precision);
    if (rc) return rc;
    int err = TFS_OK;
    const TileShape sh = effective_shape(h, precision, TFS_BLOCK_ALL);
    const Plan* p = get_plan(h, TFS_BLOCK_ALL, sh, &err);
    if (!p) return err;
    uint64_t lam_bits = 0;
    std::memcpy(&lam_bits, &h->fx_lambda, 8);
    const std::vector<uintptr_t> key = {(uintptr_t)p->d_tiles,
                                        (uintptr_t)p->d_items,
                                        (uintptr_t)h->partial,
                                        (uintptr_t)h->near[precision].cnt,
                                        (uintptr_t)h->near[precision].idx,
                                        (uintptr_t)h->near[precision].cap,
                                        (uintptr_t)h->x_fx,
                                        (uintptr_t)(precision == 1 ? lam_bits : 0),
                                        (uintptr_t)sh.block,
                                        (uintptr_t)sh.per_thread,
                                        (uintptr_t)sh.chunk,
                                        (uintptr_t)sh.persistent,
                                        (uintptr_t)sh.split,
                                        (uintptr_t)sh.symmetric,
                                        (uintptr_t)s->precond_kind,
                                        (uintptr_t)s->npatch};
    if (E.exec[precision] && E.key[precision] == key) return TFS_OK;
    if (E.exec[precision]) (void)hipGraphExecDestroy(E.exec[precision]);
    if (E.graph[precision]) (void)hipGraphDestroy(E.graph[precision]);
    E.exec[precision] = nullptr;
    E.graph[precision] = nullptr;
    const bool prof = h->profiling, timing = h->body_timing;
    h->profiling = false;
    h->body_timing = false;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        h->profiling = prof;
        h->body_timing = timing;
        return hip_fail(e, "lock-step cycle: begin capture");
    }
    rc = ens_enqueue_steps(s, precision, 0, s->m);
    hipError_t e2 = hipStreamEndCapture(s->stream, &E.graph[precision]);
    h->profiling = prof;
    h->body_timing = timing;
    if (rc || e2 != hipSuccess) {
        if (E.graph[precision]) (void)hipGraphDestroy(E.graph[precision]);
        E.graph[precision] = nullptr;
        return rc ? rc : hip_fail(e2, "lock-step cycle: end capture");
    }
    e = hipGraphInstantiate(&E.exec[precision], E.graph[precision], nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(E.graph[precision]);
        E.graph[precision] = nullptr;
        E.exec[precision] = nullptr;
        return hip_fail(e, "lock-step cycle: instantiate");
    }
    E.key[precision] = key;
    return TFS_OK;
}

// GMRES(m) cycles of every body with active[b] on A x = rhs (matvecs in
// `precision`), each to its own abs_tol[b]; r holds rhs - A x on entry and is
// work space after.  x (operator layout) is updated in place; est[b] gets the
// last residual estimate of a body that ran, iters[b] its steps.
int gmres_batched(tfs_solver* s, int precision, const double* rhs, double* x, double* r, std::vector<char>& active,
                  const std::vector<double>& abs_tol, std::vector<double>& est, int max_cycles, tfs_solve_stats* st) {
    EnsState& E = *s->ens;
    const int nb = E.nbody, m = s->m;
    for (int b = 0; b < nb; ++b) E.h_f64[b] = abs_tol[b];
    int rc = check(hipMemcpyAsync(E.abs_tol, E.h_f64, 8LL * nb, hipMemcpyHostToDevice, s->stream), "tolerances");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "tolerances");
    if (rc) return rc;
    const EnsArgs a = ens_args(s);
    std::vector<double> rn;
    std::vector<double> est_all((size_t)nb * E.mstride);
    std::vector<int32_t> kd(nb), stp(nb);
    for (int cyc = 0; cyc < max_cycles; ++cyc) {
        if (cyc > 0) {
            // the true residual of the bodies still running
            for (int b = 0; b < nb; ++b) E.h_i32[b] = active[b] ? 0 : 1;
            rc = check(hipMemcpyAsync(E.flag, E.h_i32, 4LL * nb, hipMemcpyHostToDevice, s->stream), "active bodies");
            s->rpy->body_skip = E.flag;
            if (!rc) rc = residual(s, rhs, x, false, r, precision, st);
            s->rpy->body_skip = nullptr;
            if (rc) return rc;
        }
        rc = ens_norms(s, r, rn);
        if (rc) return rc;
        int n_active = 0;
        for (int b = 0; b < nb; ++b) {
            if (!active[b]) continue;
            if (!std::isfinite(rn[b])) return fail(TFS_ERR_NON_FINITE, "GMRES: non-finite residual of body " + std::to_string(b));
            est[b] = rn[b];
            if (rn[b] <= abs_tol[b]) active[b] = 0;
            else ++n_active;
        }
        if (!n_active) break;
        st->cycles++;
        // open the cycle: stop flags, H = 0, v_0 = r / ||r||, z = P^{-1} v_0
        for (int b = 0; b < nb; ++b) E.h_i32[b] = active[b] ? 0 : kEnsInactive;
        rc = check(hipMemcpyAsync(E.stop, E.h_i32, 4LL * nb, hipMemcpyHostToDevice, s->stream), "stop flags");
        if (!rc) rc = check(k_fill_int(E.nstop, 0, 1, s->stream), "stop count");
        if (!rc) rc = check(k_fill(E.H, 0.0, (int64_t)nb * E.hstride, s->stream), "H");
        if (!rc) rc = check(k_ens_open(a, E.rk, E.norms, E.G, s->stream), "v0");
        if (!rc) rc = check(k_ens_permute(E.V, s->t, E.d_map, E.ntot, true, s->stream), "v0 to the operator layout");
        if (!rc) rc = check(k_precond(s->t, s->z, precond_args(s), s->dims, s->stream), "preconditioner");
        if (rc) return rc;
        int launched = m;
        if (E.use_graphs) {
            rc = ens_cycle_graph(s, precision);
            if (!rc) rc = check(hipGraphLaunch(E.exec[precision], s->stream), "lock-step cycle graph");
            if (rc) return rc;
            E.graph_launches++;
        } else {
            // steps enqueued four at a time; stop once every running body stopped
            E.eager_cycles++;
            launched = 0;
            for (int j0 = 0; j0 < m; j0 += 4) {
                const int j1 = std::min(m, j0 + 4);
                rc = ens_enqueue_steps(s, precision, j0, j1);
                if (!rc) rc = check(hipMemcpyAsync(E.h_i32 + nb, E.nstop, 4, hipMemcpyDeviceToHost, s->stream), "stop count");
                if (!rc) rc = check(hipStreamSynchronize(s->stream), "stop count");
                if (rc) return rc;
                launched = j1;
                if (E.h_i32[nb] >= n_active) break;
            }
        }
        // x += P^{-1} V y per body
        rc = check(k_ens_update(a, m, E.y, E.kdone, E.tk, E.G, s->stream), "lock-step update");
        if (!rc) rc = check(k_ens_permute(E.tk, s->t, E.d_map, E.ntot, true, s->stream), "update to the operator layout");
        if (!rc) rc = check(k_precond(s->t, s->z, precond_args(s), s->dims, s->stream), "preconditioner");
        if (!rc) rc = check(k_axpby(1.0, s->z, 1.0, x, (int)s->dims.n, s->stream), "x += d");
        if (!rc) rc = check(hipMemcpyAsync(kd.data(), E.kdone, 4LL * nb, hipMemcpyDeviceToHost, s->stream), "cycle steps");
        if (!rc) rc = check(hipMemcpyAsync(stp.data(), E.stop, 4LL * nb, hipMemcpyDeviceToHost, s->stream), "cycle stops");
        if (!rc) rc = check(hipMemcpyAsync(est_all.data(), E.est, 8 * est_all.size(), hipMemcpyDeviceToHost, s->stream), "estimates");
        if (!rc) rc = check(hipStreamSynchronize(s->stream), "cycle results");
        if (rc) return rc;
        int steps = 0;
        for (int b = 0; b < nb; ++b) {
            if (!active[b]) continue;
            if (stp[b] < 0) return fail(TFS_ERR_NON_FINITE, "GMRES: non-finite residual estimate of body " + std::to_string(b));
            const int k = std::min(kd[b], launched);
            E.iters[b] += k;
            steps = std::max(steps, k);
            if (k > 0) est[b] = est_all[(size_t)b * E.mstride + k - 1];
            if (est[b] <= abs_tol[b]) active[b] = 0;
        }
        st->iterations += steps;
        if (precision == TFS_PRECISION_FP64) st->fp64_matvecs += launched;
        else st->fp32_matvecs += launched;
    }
    return TFS_OK;
}

// solve_device for independent bodies: every body's solve follows the
// single-system rules (zero rhs, good warm start, rubbish guess, mixed
// precision refinement) with its own norms and tolerances.
int solve_batched(tfs_solver* s, const double* b, double* x, tfs_solve_stats* st) {
    EnsState& E = *s->ens;
    const tfs_solver_desc& D = s->desc;
    const int nb = E.nbody;
    const int n = (int)s->dims.n;
    *st = tfs_solve_stats{};
    E.iters.assign(nb, 0);
    // stamps[0..nb) = each body's last Arnoldi stop, stamps[nb] = the start
    int rc0 = check(k_ens_stamp(E.stamps, nb + 1, s->stream), "solve start");
    if (rc0) return rc0;
    E.conv.assign(nb, 0);
    E.relres.assign(nb, 0.0);
    int rc = ensure_block_precond(s);
    if (rc) return rc;
    std::vector<double> bn, rn;
    rc = ens_norms(s, b, bn);
    if (rc) return rc;
    std::vector<char> active(nb, 1);
    std::vector<double> abs_tol(nb), est(nb, 0.0);
    for (int i = 0; i < nb; ++i) {
        if (!std::isfinite(bn[i])) return fail(TFS_ERR_NON_FINITE, "solve: non-finite right-hand side of body " + std::to_string(i));
        abs_tol[i] = D.tol * bn[i];
        E.h_i32[i] = bn[i] == 0.0 ? 1 : 0;  // zero rhs: zero solution (mobility_solver.cpp:2127-2194)
    }
    rc = check(hipMemcpyAsync(E.flag, E.h_i32, 4LL * nb, hipMemcpyHostToDevice, s->stream), "zero rhs");
    if (!rc) rc = check(k_ens_reset(E.flag, E.d_koff, nb, E.d_map, x, nullptr, b, E.G, s->stream), "zero rhs");
    if (!rc) rc = residual(s, b, x, false, s->r, TFS_PRECISION_FP64, st);
    if (!rc) rc = ens_norms(s, s->r, rn);
    if (rc) return rc;
    int any_reset = 0;
    for (int i = 0; i < nb; ++i) {
        E.h_i32[i] = 0;
        if (bn[i] == 0.0) {
            active[i] = 0;
            rn[i] = 0.0;
            continue;
        }
        if (!std::isfinite(rn[i])) return fail(TFS_ERR_NON_FINITE, "solve: non-finite initial residual of body " + std::to_string(i));
        if (rn[i] <= abs_tol[i]) active[i] = 0;        // good warm start (:2304-2308)
        else if (rn[i] > bn[i]) {                      // rubbish initial guess (:2310-2317)
            E.h_i32[i] = 1;
            rn[i] = bn[i];
            any_reset = 1;
        }
    }
    if (any_reset) {
        rc = check(hipMemcpyAsync(E.flag, E.h_i32, 4LL * nb, hipMemcpyHostToDevice, s->stream), "reset");
        if (!rc) rc = check(k_ens_reset(E.flag, E.d_koff, nb, E.d_map, x, s->r, b, E.G, s->stream), "reset");
        if (rc) return rc;
    }
    double init = 0.0;
    for (int i = 0; i < nb; ++i) init = std::max(init, bn[i] > 0 ? rn[i] / bn[i] : 0.0);
    st->initial_relres = init;
    if (D.mode == TFS_SOLVE_FP64) {
        rc = gmres_batched(s, TFS_PRECISION_FP64, b, x, s->r, active, abs_tol, est, D.max_cycles, st);
        if (rc) return rc;
        for (int i = 0; i < nb; ++i)
            if (bn[i] > 0.0 && E.iters[i] > 0) rn[i] = est[i];
    } else {
        std::vector<double> inner(nb);
        for (int it = 0; it < D.max_refinements; ++it) {
            std::vector<char> act(nb, 0);
            int n_act = 0;
            for (int i = 0; i < nb; ++i) {
                act[i] = bn[i] > 0.0 && rn[i] > abs_tol[i];
                n_act += act[i];
                inner[i] = std::max(D.inner_tol * rn[i], kInnerFloor * abs_tol[i]);
            }
            if (!n_act) break;
            st->refinements++;
            rc = check(hipMemcpyAsync(s->rin, s->r, (size_t)n * 8, hipMemcpyDeviceToDevice, s->stream), "inner rhs");
            if (!rc) rc = check(k_fill(s->d, 0.0, n, s->stream), "d = 0");
            if (!rc) rc = gmres_batched(s, TFS_PRECISION_FP32_ACC64, s->rin, s->d, s->r, act, inner, est, D.max_cycles, st);
            if (!rc) rc = check(k_axpby(1.0, s->d, 1.0, x, n, s->stream), "x += d");
            if (!rc) rc = residual(s, b, x, false, s->r, TFS_PRECISION_FP64, st);
            if (!rc) rc = ens_norms(s, s->r, rn);
            if (rc) return rc;
            for (int i = 0; i < nb; ++i)
                if (!std::isfinite(rn[i])) return fail(TFS_ERR_NON_FINITE, "solve: non-finite residual of body " + std::to_string(i));
        }
    }
    bool all = true;
    double worst = 0.0;
    for (int i = 0; i < nb; ++i) {
        E.relres[i] = bn[i] > 0.0 ? rn[i] / bn[i] : 0.0;
        E.conv[i] = bn[i] == 0.0 || rn[i] <= abs_tol[i];
        all = all && E.conv[i];
        worst = std::max(worst, E.relres[i]);
    }
    st->converged = all ? 1 : 0;
    st->relres = worst;
    return all ? TFS_OK : fail(TFS_ERR_NOT_CONVERGED, "solve: tolerance not met by every body");
}

void free_ens(tfs_solver* s) {
    EnsState* E = s->ens;
    if (!E) return;
    for (int p = 0; p < 2; ++p) {
        if (E->exec[p]) (void)hipGraphExecDestroy(E->exec[p]);
        if (E->graph[p]) (void)hipGraphDestroy(E->graph[p]);
    }
    if (E->arena) (void)hipFree(E->arena);
    if (E->h_i32) (void)hipHostFree(E->h_i32);
    if (E->h_f64) (void)hipHostFree(E->h_f64);
    if (E->h_stamps) (void)hipHostFree(E->h_stamps);
    delete E;
    s->ens = nullptr;
}

int alloc_solver(tfs_solver* s) {
    const int64_t n = s->dims.n;
    s->ldv = ((n + 31) / 32) * 32;
    const int m = s->m, ldh = m + 1;
    s->G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
    s->GA = arnoldi_blocks(n);
    const int64_t ns3 = 3LL * s->desc.num_fil * s->desc.num_seg;
    std::vector<std::pair<double**, int64_t>> parts = {
        {&s->kphi, std::max<int64_t>(ns3, 1)}, {&s->kth, std::max<int64_t>(ns3, 1)},
        {&s->b, s->ldv}, {&s->x, s->ldv}, {&s->r, s->ldv}, {&s->w, s->ldv}, {&s->z, s->ldv},
        {&s->t, s->ldv}, {&s->d, s->ldv}, {&s->rin, s->ldv}, {&s->V, s->ldv * (m + 1)}, {&s->H, (int64_t)ldh * m},
        {&s->cs, m}, {&s->sn, m}, {&s->g, ldh}, {&s->y, ldh}, {&s->h, ldh},
        {&s->partial, (int64_t)s->G * ldh}, {&s->scal, 8}, {&s->nrm, 8}, {&s->est, ldh},
        {&s->pa, (int64_t)s->GA * ldh}, {&s->pb, (int64_t)s->GA * ldh}, {(double**)&s->stop, 1}};
    if (s->dims.free_body) {
        const int64_t nsw = s->dims.nswim;
        parts.push_back({&s->x_bodies, 3 * nsw});
        parts.push_back({&s->bodyZ, 6 * s->ldv});
        parts.push_back({&s->body_sinv, 36 * nsw});
        parts.push_back({&s->body_t, 6 * nsw});
    }
    int64_t total = 0;
    for (auto& p : parts) total += ((p.second + 31) / 32) * 32;
    hipError_t e = hipMalloc(&s->arena, total * sizeof(double));
    if (e != hipSuccess) return fail(TFS_ERR_OUT_OF_MEMORY, "solver: device memory");
    double* p = s->arena;
    for (auto& q : parts) {
        *q.first = p;
        p += ((q.second + 31) / 32) * 32;
    }
    e = hipMemsetAsync(s->arena, 0, total * sizeof(double), s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "solver: clear");
    e = hipHostMalloc((void**)&s->host_scal, sizeof(double) * (kEstBatch + 8), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->host_stop, sizeof(int32_t) * 4, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&s->dev_stop_mirror, s->host_stop, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->host_est, sizeof(double), hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&s->dev_est_mirror, s->host_est, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->flag_ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->flag_ev[1], hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "solver: pinned scalar");
    return TFS_OK;
}

hipMemcpyKind kind_in(int32_t memory) {
    return memory == TFS_MEMORY_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
hipMemcpyKind kind_out(int32_t memory) {
    return memory == TFS_MEMORY_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

// Core solve on device vectors b (rhs) and x (warm start in, solution out).
int solve_device(tfs_solver* s, const double* b, double* x, tfs_solve_stats* st) {
    if (s->ens) return solve_batched(s, b, x, st);
    const int n = (int)s->dims.n;
    const tfs_solver_desc& D = s->desc;
    *st = tfs_solve_stats{};
    int rc = ensure_block_precond(s);
    if (!rc) rc = ensure_body_precond(s);
    if (rc) return rc;
    double bnorm = 0.0;
    rc = norm(s, b, &bnorm);
    if (rc) return rc;
    if (!std::isfinite(bnorm)) return fail(TFS_ERR_NON_FINITE, "solve: non-finite right-hand side");
    if (bnorm == 0.0) {  // mobility_solver.cpp:2127-2194
        st->converged = 1;
        return check(k_fill(x, 0.0, n, s->stream), "zero solution");
    }
    const int outer_prec = TFS_PRECISION_FP64;
    rc = residual(s, b, x, false, s->r, outer_prec, st);
    double rnorm = 0.0;
    if (!rc) rc = norm(s, s->r, &rnorm);
    if (rc) return rc;
    if (!std::isfinite(rnorm)) return fail(TFS_ERR_NON_FINITE, "solve: non-finite initial residual");
    st->initial_relres = rnorm / bnorm;
    const double abs_tol = D.tol * bnorm;
    if (rnorm <= abs_tol) {  // mobility_solver.cpp:2304-2308: the warm start is good enough
        st->converged = 1;
        st->relres = rnorm / bnorm;
        return TFS_OK;
    }
    if (rnorm > bnorm) {  // mobility_solver.cpp:2310-2317: drop a rubbish initial guess
        rc = check(k_fill(x, 0.0, n, s->stream), "reset");
        if (!rc) rc = check(hipMemcpyAsync(s->r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream), "r=b");
        if (rc) return rc;
        rnorm = bnorm;
    }
    int conv = 0;
    double res = rnorm;
    if (D.mode == TFS_SOLVE_FP64) {
        rc = gmres(s, TFS_PRECISION_FP64, b, x, s->r, rnorm, abs_tol, D.max_cycles, st, &conv, &res);
        if (rc) return rc;
        st->converged = conv;
        st->relres = res / bnorm;
    } else {
        // Mixed-precision iterative refinement: the residual of x is formed with
        // FP64 matvecs; each correction d solves A d = r with FP32-tile matvecs
        // (FP64 Arnoldi) to a relative reduction inner_tol.
        for (int it = 0; it < D.max_refinements && rnorm > abs_tol; ++it) {
            st->refinements++;
            rc = check(hipMemcpyAsync(s->rin, s->r, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream),
                       "inner rhs");
            if (!rc) rc = check(k_fill(s->d, 0.0, n, s->stream), "d=0");
            int inner_conv = 0;
            double inner_res = 0.0;
            if (!rc)
                rc = gmres(s, TFS_PRECISION_FP32_ACC64, s->rin, s->d, s->r, rnorm,
                           std::max(D.inner_tol * rnorm, kInnerFloor * abs_tol), D.max_cycles, st, &inner_conv,
                           &inner_res);
            if (!rc) rc = check(k_axpby(1.0, s->d, 1.0, x, n, s->stream), "x += d");
            if (!rc) rc = residual(s, b, x, false, s->r, TFS_PRECISION_FP64, st);
            if (!rc) rc = norm(s, s->r, &rnorm);
            if (rc) return rc;
            if (!std::isfinite(rnorm)) return fail(TFS_ERR_NON_FINITE, "solve: non-finite residual");
        }
        st->converged = rnorm <= abs_tol;
        st->relres = rnorm / bnorm;
    }
    if (!std::isfinite(st->relres)) return fail(TFS_ERR_NON_FINITE, "solve: non-finite residual");
    return st->converged ? TFS_OK : fail(TFS_ERR_NOT_CONVERGED, "solve: tolerance not met");
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI: solver
// ---------------------------------------------------------------------------
extern "C" {

int tfs_solver_create(const tfs_solver_desc* desc, tfs_solver** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: null argument");
    *out = nullptr;
    const tfs_solver_desc& D = *desc;
    if (D.num_fil < 0 || D.num_seg < 0 || D.num_blobs < 0 || (D.num_fil > 0 && D.num_seg < 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: bad sizes");
    if ((int64_t)D.num_fil * D.num_seg + D.num_blobs == 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: empty system");
    if (D.mode != TFS_SOLVE_FP64 && D.mode != TFS_SOLVE_MIXED)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: unknown mode");
    if (D.restart < 1 || D.restart > kMaxRestart)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: restart must be in [1, 350]");
    if (D.max_cycles < 1 || !(D.tol > 0) || (D.mode == TFS_SOLVE_MIXED && (!(D.inner_tol > 0) || D.max_refinements < 1)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: bad tolerances or iteration limits");
    if (!(D.seg_radius > 0) || !(D.blob_radius > 0) || !(D.mu > 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: radii and mu must be positive");
    tfs_solver* s = new tfs_solver();
    s->desc = D;
    SaddleDims& dm = s->dims;
    dm.nfil = D.num_fil;
    dm.nseg = D.num_seg;
    dm.nblob = D.num_blobs;
    dm.dyn_phase = D.dynamic_phase ? 1 : 0;
    dm.dyn_angle = D.dynamic_angle ? 1 : 0;
    const int64_t ns = (int64_t)D.num_fil * D.num_seg;
    dm.nswim = D.num_swim > 0 ? D.num_swim : 1;
    dm.free_body = D.free_bodies ? 1 : 0;
    if (D.num_fil % dm.nswim != 0 || D.num_blobs % dm.nswim != 0) {
        delete s;
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: filaments and blobs must split evenly between swimmers");
    }
    dm.fil_per_swim = D.num_fil / dm.nswim;
    dm.segs_per_swim = (int)(ns / dm.nswim);
    dm.blobs_per_swim = D.num_blobs / dm.nswim;
    dm.off_blob = 3 * ns;
    dm.off_body = dm.off_blob + 3LL * D.num_blobs;
    dm.off_phase = dm.off_body + (dm.free_body ? 6LL * dm.nswim : 0);
    dm.off_angle = dm.off_phase + (dm.dyn_phase ? D.num_fil : 0);
    dm.n = dm.off_angle + (dm.dyn_angle ? D.num_fil : 0);
    if (dm.n > INT32_MAX / 2) {
        delete s;
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_create: system too large");
    }
    s->m = (int)std::min<int64_t>(D.restart, dm.n);
    const double PI = 3.14159265358979323846;
    s->gs = 6.0 * PI * D.mu * D.seg_radius;
    s->gb = 6.0 * PI * D.mu * D.blob_radius;
    tfs_rpy_desc rd{};
    rd.num_segs = (int32_t)ns;
    rd.num_blobs = D.num_blobs;
    rd.seg_dof = 3;
    rd.precision = TFS_PRECISION_FP64;
    rd.bs_mode = TFS_BS_ACCUMULATE;
    rd.memory = TFS_MEMORY_DEVICE;
    rd.seg_radius = D.seg_radius;
    rd.blob_radius = D.blob_radius;
    rd.mu = D.mu;
    rd.stream = D.stream;
    int rc = tfs_rpy_create(&rd, &s->rpy);
    if (rc) {
        delete s;
        return rc;
    }
    s->stream = s->rpy->stream;
    rc = alloc_solver(s);
    if (rc) {
        tfs_solver_destroy(s);
        return rc;
    }
    *out = s;
    return TFS_OK;
}

int tfs_solver_destroy(tfs_solver* s) {
    if (!s) return TFS_OK;
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->arena) hipFree(s->arena);
    if (s->host_scal) hipHostFree(s->host_scal);
    if (s->host_stop) hipHostFree(s->host_stop);
    if (s->host_est) hipHostFree(s->host_est);
    free_stage(s->f32);
    free_ens(s);
    for (hipEvent_t ev : s->flag_ev)
        if (ev) hipEventDestroy(ev);
    if (s->rpy) tfs_rpy_destroy(s->rpy);
    for (void* q : {(void*)s->fil_inv, (void*)s->fil_dk, (void*)s->fil_sinv, (void*)s->patch_inv, (void*)s->fil_off,
                    (void*)s->patch_off, (void*)s->fil_size, (void*)s->fil_piv, (void*)s->fil_pivoff,
                    (void*)s->fil_status, (void*)s->patch_blob, (void*)s->patch_start, (void*)s->patch_size,
                    (void*)s->patch_piv, (void*)s->patch_pivoff, (void*)s->patch_status})
        if (q) (void)hipFree(q);
    delete s;
    return TFS_OK;
}

int tfs_solver_set_bodies(tfs_solver* s, const double* x_bodies, int32_t memory) {
    if (!s || !x_bodies) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_bodies: null argument");
    if (!s->dims.free_body) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_bodies: the system has no free bodies");
    if (memory != TFS_MEMORY_HOST && memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_bodies: unknown memory space");
    hipError_t e = hipMemcpyAsync(s->x_bodies, x_bodies, 24LL * s->dims.nswim, kind_in(memory), s->stream);
    if (e == hipSuccess && memory == TFS_MEMORY_HOST) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_solver_set_bodies");
    s->bodies_set = true;
    s->body_pc_valid = false;
    return TFS_OK;
}

int tfs_solver_set_preconditioner(tfs_solver* s, int32_t kind) {
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_preconditioner: null handle");
    if (kind != TFS_PRECOND_REFERENCE && kind != TFS_PRECOND_BLOCKS)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_preconditioner: unknown kind");
    if (kind == TFS_PRECOND_BLOCKS && s->desc.num_seg > kMaxBlockFilSeg)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_preconditioner: block form needs num_seg <= " +
                                                  std::to_string(kMaxBlockFilSeg));
    s->precond_kind = kind;
    s->body_pc_valid = false;
    return TFS_OK;
}

int tfs_solver_rpy(tfs_solver* s, tfs_rpy** out) {
    if (!s || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_rpy: null argument");
    *out = s->rpy;
    return TFS_OK;
}

int tfs_solver_size(tfs_solver* s, int64_t* n) {
    if (!s || !n) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_size: null argument");
    *n = s->dims.n;
    return TFS_OK;
}

int tfs_solver_set_geometry(tfs_solver* s, const double* x_segs, const double* x_blobs,
                            const double* vel_dir_phase, const double* vel_dir_angle, int32_t memory) {
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_geometry: null handle");
    if (memory != TFS_MEMORY_HOST && memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_geometry: unknown memory space");
    const size_t ns3 = (size_t)3 * s->desc.num_fil * s->desc.num_seg, nb3 = (size_t)3 * s->desc.num_blobs;
    const hipMemcpyKind k = kind_in(memory);
    hipError_t e = hipSuccess;
    if (x_segs || x_blobs) invalidate_near(s->rpy);
    if (x_segs || vel_dir_phase || vel_dir_angle) s->fil_pc_valid = false;
    if (x_blobs) s->blob_pc_valid = false;
    if (x_segs || x_blobs || vel_dir_phase || vel_dir_angle) s->body_pc_valid = false;
    if (x_segs && ns3) e = hipMemcpyAsync(s->rpy->x_segs, x_segs, ns3 * 8, k, s->stream);
    if (e == hipSuccess && x_blobs && nb3) e = hipMemcpyAsync(s->rpy->x_blobs, x_blobs, nb3 * 8, k, s->stream);
    if (e == hipSuccess && vel_dir_phase && ns3) e = hipMemcpyAsync(s->kphi, vel_dir_phase, ns3 * 8, k, s->stream);
    if (e == hipSuccess && vel_dir_angle && ns3) e = hipMemcpyAsync(s->kth, vel_dir_angle, ns3 * 8, k, s->stream);
    if (e == hipSuccess && memory == TFS_MEMORY_HOST) e = hipStreamSynchronize(s->stream);
    return check(e, "tfs_solver_set_geometry");
}

int tfs_solver_system_mult(tfs_solver* s, const double* in, double* out, int32_t precision, int32_t memory) {
    if (!s || !in || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_system_mult: null argument");
    if (precision != TFS_PRECISION_FP64 && precision != TFS_PRECISION_FP32_ACC64)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_system_mult: unknown precision");
    const size_t n = s->dims.n;
    const double* din = in;
    double* dout = out;
    int rc = TFS_OK;
    if (memory == TFS_MEMORY_HOST) {
        rc = check(hipMemcpyAsync(s->b, in, n * 8, hipMemcpyHostToDevice, s->stream), "upload");
        din = s->b;
        dout = s->x;
    }
    if (!rc) rc = apply_A(s, din, dout, precision, nullptr);
    if (!rc && memory == TFS_MEMORY_HOST)
        rc = check(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, s->stream), "download");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "sync");
    return rc;
}

int tfs_solver_precondition(tfs_solver* s, const double* in, double* out, int32_t memory) {
    if (!s || !in || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_precondition: null argument");
    const size_t n = s->dims.n;
    const double* din = in;
    double* dout = out;
    int rc = TFS_OK;
    if (memory == TFS_MEMORY_HOST) {
        rc = check(hipMemcpyAsync(s->b, in, n * 8, hipMemcpyHostToDevice, s->stream), "upload");
        din = s->b;
        dout = s->x;
    }
    if (!rc) rc = ensure_block_precond(s);
    if (!rc) rc = ensure_body_precond(s);
    if (!rc) rc = apply_P(s, din, dout);
    if (!rc && memory == TFS_MEMORY_HOST)
        rc = check(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, s->stream), "download");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "sync");
    return rc;
}

int tfs_solver_solve(tfs_solver* s, const double* rhs, double* x, int32_t memory, tfs_solve_stats* stats) {
    if (!s || !rhs || !x) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_solve: null argument");
    const size_t n = s->dims.n;
    tfs_solve_stats local{};
    tfs_solve_stats* st = stats ? stats : &local;
    int rc = check(hipMemcpyAsync(s->b, rhs, n * 8, kind_in(memory), s->stream), "rhs");
    if (!rc) rc = check(hipMemcpyAsync(s->x, x, n * 8, kind_in(memory), s->stream), "x0");
    if (rc) return rc;
    const int solve_rc = solve_device(s, s->b, s->x, st);
    if (solve_rc != TFS_OK && solve_rc != TFS_ERR_NOT_CONVERGED) return solve_rc;
    rc = check(hipMemcpyAsync(x, s->x, n * 8, kind_out(memory), s->stream), "x");
    if (!rc) rc = check(hipStreamSynchronize(s->stream), "sync");
    return rc ? rc : solve_rc;
}

int tfs_solver_residual(tfs_solver* s, const double* rhs, const double* x, int32_t precision, int32_t memory,
                        double* relres) {
    if (!s || !rhs || !x || !relres) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_residual: null argument");
    const size_t n = s->dims.n;
    int rc = check(hipMemcpyAsync(s->b, rhs, n * 8, kind_in(memory), s->stream), "rhs");
    if (!rc) rc = check(hipMemcpyAsync(s->d, x, n * 8, kind_in(memory), s->stream), "x");
    if (!rc) rc = residual(s, s->b, s->d, false, s->r, precision, nullptr);
    double rn = 0, bn = 0;
    if (!rc) rc = norm(s, s->r, &rn);
    if (!rc) rc = norm(s, s->b, &bn);
    if (rc) return rc;
    *relres = bn > 0 ? rn / bn : rn;
    return TFS_OK;
}

// ---- FP32 arrays (Real = float builds): widened into the FP64 device
// vectors the solver works on, results narrowed back. --------------------

static int memory_ok(int32_t memory, const char* where) {
    return memory == TFS_MEMORY_HOST || memory == TFS_MEMORY_DEVICE
               ? TFS_OK
               : fail(TFS_ERR_INVALID_ARGUMENT, std::string(where) + ": unknown memory space");
}

int tfs_solver_set_geometry_f32(tfs_solver* s, const float* x_segs, const float* x_blobs, const float* vel_dir_phase,
                                const float* vel_dir_angle, int32_t memory) {
    const char* w = "tfs_solver_set_geometry_f32";
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null handle");
    int rc = memory_ok(memory, w);
    if (rc) return rc;
    const int64_t ns3 = 3LL * s->desc.num_fil * s->desc.num_seg, nb3 = 3LL * s->desc.num_blobs;
    const bool host = memory == TFS_MEMORY_HOST;
    if (x_segs || x_blobs) invalidate_near(s->rpy);
    if (x_segs || vel_dir_phase || vel_dir_angle) s->fil_pc_valid = false;
    if (x_blobs) s->blob_pc_valid = false;
    if (x_segs || x_blobs || vel_dir_phase || vel_dir_angle) s->body_pc_valid = false;
    if (x_segs) rc = widen_in(s->f32, x_segs, s->rpy->x_segs, ns3, host, s->stream, w);
    if (!rc && x_blobs) rc = widen_in(s->f32, x_blobs, s->rpy->x_blobs, nb3, host, s->stream, w);
    if (!rc && vel_dir_phase) rc = widen_in(s->f32, vel_dir_phase, s->kphi, ns3, host, s->stream, w);
    if (!rc && vel_dir_angle) rc = widen_in(s->f32, vel_dir_angle, s->kth, ns3, host, s->stream, w);
    if (!rc && host) rc = check(hipStreamSynchronize(s->stream), w);
    return rc;
}

int tfs_solver_set_bodies_f32(tfs_solver* s, const float* x_bodies, int32_t memory) {
    const char* w = "tfs_solver_set_bodies_f32";
    if (!s || !x_bodies) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null argument");
    if (!s->dims.free_body) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": the system has no free bodies");
    int rc = memory_ok(memory, w);
    if (!rc) rc = widen_in(s->f32, x_bodies, s->x_bodies, 3LL * s->dims.nswim, memory == TFS_MEMORY_HOST, s->stream, w);
    if (!rc && memory == TFS_MEMORY_HOST) rc = check(hipStreamSynchronize(s->stream), w);
    if (rc) return rc;
    s->bodies_set = true;
    s->body_pc_valid = false;
    return TFS_OK;
}

int tfs_solver_system_mult_f32(tfs_solver* s, const float* in, float* out, int32_t precision, int32_t memory) {
    const char* w = "tfs_solver_system_mult_f32";
    if (!s || !in || !out) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null argument");
    if (precision != TFS_PRECISION_FP64 && precision != TFS_PRECISION_FP32_ACC64)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": unknown precision");
    int rc = memory_ok(memory, w);
    const bool host = memory == TFS_MEMORY_HOST;
    const int64_t n = s->dims.n;
    if (!rc) rc = widen_in(s->f32, in, s->b, n, host, s->stream, w);
    if (!rc) rc = apply_A(s, s->b, s->x, precision, nullptr);
    if (!rc) rc = narrow_out(s->f32, s->x, out, n, host, s->stream, w);
    if (!rc) rc = check(hipStreamSynchronize(s->stream), w);
    return rc;
}

int tfs_solver_precondition_f32(tfs_solver* s, const float* in, float* out, int32_t memory) {
    const char* w = "tfs_solver_precondition_f32";
    if (!s || !in || !out) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null argument");
    int rc = memory_ok(memory, w);
    const bool host = memory == TFS_MEMORY_HOST;
    const int64_t n = s->dims.n;
    if (!rc) rc = widen_in(s->f32, in, s->b, n, host, s->stream, w);
    if (!rc) rc = ensure_block_precond(s);
    if (!rc) rc = ensure_body_precond(s);
    if (!rc) rc = apply_P(s, s->b, s->x);
    if (!rc) rc = narrow_out(s->f32, s->x, out, n, host, s->stream, w);
    if (!rc) rc = check(hipStreamSynchronize(s->stream), w);
    return rc;
}

int tfs_solver_solve_f32(tfs_solver* s, const float* rhs, float* x, int32_t memory, tfs_solve_stats* stats) {
    const char* w = "tfs_solver_solve_f32";
    if (!s || !rhs || !x) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null argument");
    int rc = memory_ok(memory, w);
    if (rc) return rc;
    const bool host = memory == TFS_MEMORY_HOST;
    const int64_t n = s->dims.n;
    tfs_solve_stats local{};
    tfs_solve_stats* st = stats ? stats : &local;
    rc = widen_in(s->f32, rhs, s->b, n, host, s->stream, w);
    if (!rc) rc = widen_in(s->f32, x, s->x, n, host, s->stream, w);
    if (rc) return rc;
    const int solve_rc = solve_device(s, s->b, s->x, st);
    if (solve_rc != TFS_OK && solve_rc != TFS_ERR_NOT_CONVERGED) return solve_rc;
    const std::string msg = g_last_error;
    rc = narrow_out(s->f32, s->x, x, n, host, s->stream, w);
    if (!rc) rc = check(hipStreamSynchronize(s->stream), w);
    if (rc) return rc;
    g_last_error = msg;
    return solve_rc;
}

int tfs_solver_residual_f32(tfs_solver* s, const float* rhs, const float* x, int32_t precision, int32_t memory,
                            double* relres) {
    const char* w = "tfs_solver_residual_f32";
    if (!s || !rhs || !x || !relres) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null argument");
    int rc = memory_ok(memory, w);
    const bool host = memory == TFS_MEMORY_HOST;
    const int64_t n = s->dims.n;
    if (!rc) rc = widen_in(s->f32, rhs, s->b, n, host, s->stream, w);
    if (!rc) rc = widen_in(s->f32, x, s->d, n, host, s->stream, w);
    if (!rc) rc = residual(s, s->b, s->d, false, s->r, precision, nullptr);
    double rn = 0, bn = 0;
    if (!rc) rc = norm(s, s->r, &rn);
    if (!rc) rc = norm(s, s->b, &bn);
    if (rc) return rc;
    *relres = bn > 0 ? rn / bn : rn;
    return TFS_OK;
}

int tfs_solver_set_independent_bodies(tfs_solver* s, int32_t num_bodies, const int32_t* fil_per_body,
                                      const int32_t* blobs_per_body) {
    g_last_error.clear();
    if (!s || num_bodies < 1 || !fil_per_body || !blobs_per_body)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: bad arguments");
    const SaddleDims& d = s->dims;
    if (d.free_body) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: not with free bodies");
    int64_t nf = 0, nbl = 0;
    for (int b = 0; b < num_bodies; ++b) {
        if (fil_per_body[b] < 0 || blobs_per_body[b] < 0 || fil_per_body[b] + blobs_per_body[b] == 0)
            return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: every body needs particles");
        nf += fil_per_body[b];
        nbl += blobs_per_body[b];
    }
    if (nf != d.nfil || nbl != d.nblob)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: counts do not add up to the solver's sizes");
    const int m = s->m;
    // the per-body back substitution stages its triangle in LDS
    if (num_bodies > 1 && m > 120)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: restart above 120");
    int rc = check(hipStreamSynchronize(s->stream), "tfs_solver_set_independent_bodies");
    if (rc) return rc;
    free_ens(s);
    std::vector<int32_t> segs(num_bodies);
    for (int b = 0; b < num_bodies; ++b) segs[b] = fil_per_body[b] * d.nseg;
    rc = tfs_rpy_set_bodies(s->rpy, num_bodies, segs.data(), blobs_per_body);
    s->blob_pc_valid = false;
    if (rc || num_bodies == 1) return rc;  // one body: the single-system solve
    EnsState* E = new EnsState();
    s->ens = E;
    E->nbody = num_bodies;
    E->fil.assign(fil_per_body, fil_per_body + num_bodies);
    E->blob.assign(blobs_per_body, blobs_per_body + num_bodies);
    // Krylov layout of body b: its lambda_s rows, lambda_b rows, phase rows, angle rows
    std::vector<int32_t> map, fil_body(std::max(d.nfil, 1), 0);
    map.reserve(d.n);
    E->koff.assign(1, 0);
    int64_t max_n = 0;
    for (int b = 0, f0 = 0, b0 = 0; b < num_bodies; ++b) {
        for (int64_t i = 3LL * f0 * d.nseg; i < 3LL * (f0 + fil_per_body[b]) * d.nseg; ++i) map.push_back((int32_t)i);
        for (int64_t i = 3LL * b0; i < 3LL * (b0 + blobs_per_body[b]); ++i) map.push_back((int32_t)(d.off_blob + i));
        for (int f = f0; f < f0 + fil_per_body[b]; ++f) {
            fil_body[f] = b;
            if (d.dyn_phase) map.push_back((int32_t)(d.off_phase + f));
        }
        if (d.dyn_angle)
            for (int f = f0; f < f0 + fil_per_body[b]; ++f) map.push_back((int32_t)(d.off_angle + f));
        E->koff.push_back((int64_t)map.size());
        max_n = std::max(max_n, E->koff[b + 1] - E->koff[b]);
        f0 += fil_per_body[b];
        b0 += blobs_per_body[b];
    }
    E->ntot = (int64_t)map.size();
    if (E->ntot != d.n) {
        free_ens(s);
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_independent_bodies: layout size");
    }
    const int nb = num_bodies;
    E->G = arnoldi_blocks(max_n);
    E->ldv = ((E->ntot + 31) / 32) * 32;
    E->mstride = m + 1;
    E->hstride = (int64_t)(m + 1) * m;
    E->pstride = (int64_t)E->G * (m + 1);
    // one device arena, every part 256-byte aligned
    std::vector<std::pair<void**, int64_t>> parts = {
        {(void**)&E->d_koff, 8LL * (nb + 1)}, {(void**)&E->d_map, 4 * E->ntot},
        {(void**)&E->d_fil_body, 4LL * fil_body.size()}, {(void**)&E->V, 8 * E->ldv * (m + 1)},
        {(void**)&E->w, 8 * E->ldv}, {(void**)&E->tk, 8 * E->ldv}, {(void**)&E->rk, 8 * E->ldv},
        {(void**)&E->H, 8 * nb * E->hstride}, {(void**)&E->cs, 8LL * nb * E->mstride},
        {(void**)&E->sn, 8LL * nb * E->mstride}, {(void**)&E->g, 8LL * nb * E->mstride},
        {(void**)&E->est, 8LL * nb * E->mstride}, {(void**)&E->y, 8LL * nb * E->mstride},
        {(void**)&E->scal, 16LL * nb}, {(void**)&E->abs_tol, 8LL * nb}, {(void**)&E->norms, 8LL * nb},
        {(void**)&E->pa, 8 * nb * E->pstride}, {(void**)&E->pb, 8 * nb * E->pstride},
        {(void**)&E->stop, 4LL * nb}, {(void**)&E->nstop, 4}, {(void**)&E->kdone, 4LL * nb},
        {(void**)&E->flag, 4LL * nb}, {(void**)&E->stamps, 8LL * (nb + 1)}};
    int64_t total = 0;
    for (auto& q : parts) total += (q.second + 255) / 256 * 256;
    hipError_t e = hipMalloc(&E->arena, total);
    if (e != hipSuccess) {
        free_ens(s);
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_solver_set_independent_bodies: device memory");
    }
    char* p = (char*)E->arena;
    for (auto& q : parts) {
        *q.first = p;
        p += (q.second + 255) / 256 * 256;
    }
    e = hipHostMalloc((void**)&E->h_i32, 4 * (2LL * nb + 8), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&E->h_f64, 8LL * nb, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&E->h_stamps, 8LL * (nb + 1), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(E->arena, 0, total, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(E->d_koff, E->koff.data(), 8LL * (nb + 1), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(E->d_map, map.data(), 4 * E->ntot, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(E->d_fil_body, fil_body.data(), 4 * fil_body.size(), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        free_ens(s);
        return hip_fail(e, "tfs_solver_set_independent_bodies: set-up");
    }
    return TFS_OK;
}

int tfs_solver_set_graphs(tfs_solver* s, int32_t enable) {
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_graphs: null handle");
    if (!s->ens) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_set_graphs: no independent bodies set");
    s->ens->use_graphs = enable != 0;
    return TFS_OK;
}

int tfs_solver_body_stats(tfs_solver* s, int32_t* iterations, int32_t* converged, double* relres, double* solve_ms) {
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_body_stats: null handle");
    if (!s->ens) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_body_stats: no independent bodies set");
    EnsState& E = *s->ens;
    const int nb = E.nbody;
    if (E.iters.size() != (size_t)nb) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_body_stats: no solve yet");
    if (iterations) std::copy(E.iters.begin(), E.iters.end(), iterations);
    if (converged) std::copy(E.conv.begin(), E.conv.end(), converged);
    if (relres) std::copy(E.relres.begin(), E.relres.end(), relres);
    if (solve_ms) {
        int khz = 0;
        hipError_t e = hipMemcpyAsync(E.h_stamps, E.stamps, 8LL * (nb + 1), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->rpy->device);
        if (e != hipSuccess) return hip_fail(e, "tfs_solver_body_stats");
        if (khz <= 0) return fail(TFS_ERR_HIP, "tfs_solver_body_stats: no wall clock rate");
        for (int b = 0; b < nb; ++b) solve_ms[b] = double(E.h_stamps[b] - E.h_stamps[nb]) / khz;
    }
    return TFS_OK;
}

int tfs_solver_cycle_counts(tfs_solver* s, int64_t* graph_cycles, int64_t* eager_cycles) {
    if (!s) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_solver_cycle_counts: null handle");
    if (graph_cycles) *graph_cycles = s->ens ? s->ens->graph_launches : 0;
    if (eager_cycles) *eager_cycles = s->ens ? s->ens->eager_cycles : 0;
    return TFS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Prescribed-beat timestep driver
// ---------------------------------------------------------------------------
struct tfs_beat {
    tfs_beat_desc desc{};
    tfs_solver* solver = nullptr;
    int64_t step = 0;
    double* arena = nullptr;
    double *frames = nullptr, *bases = nullptr, *s_values = nullptr, *phase = nullptr, *angle = nullptr,
           *omega0 = nullptr, *gen_phase = nullptr, *gen_angle = nullptr, *body = nullptr, *rhs = nullptr,
           *sol = nullptr;
    // Warm-start extrapolation: order 0 = the previous solution (the
    // reference), 1 = linear, 2 = quadratic in time from the last solutions;
    // hist1/hist2 hold the two solutions before the one in `sol`.
    int warm_order = 0;
    double *hist1 = nullptr, *hist2 = nullptr;
    std::vector<double> beat_A, beat_B;
    // Velocities of the last solution's forces (tfs_beat_velocities), made on
    // first use.
    double* vel = nullptr;
    // Swimming body (solver.free_bodies): pose (X, q) and body-frame
    // references of the filament bases, frames and blobs.
    bool free_body = false;
    double *pose = nullptr, *base_ref = nullptr, *frame_ref = nullptr, *blob_ref = nullptr;
};

extern "C" {

int tfs_beat_create(const tfs_beat_desc* desc, tfs_beat** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: null argument");
    *out = nullptr;
    const tfs_beat_desc& D = *desc;
    const int nf = D.solver.num_fil, nseg = D.solver.num_seg;
    if (nf < 1 || nseg < 2) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: need filaments of >= 2 segments");
    if (!D.frames || !D.bases || !D.omega0 || !D.phase0 || !D.beat_A || !D.beat_B)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: missing geometry or beat tables");
    if (D.solver.num_blobs > 0 && !D.x_blobs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: missing blobs");
    if (D.solver.num_swim > 1) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: one body");
    if (D.solver.dynamic_phase && (!D.gen_phase_forces || D.n_gen_phase < 2))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: missing generalised phase forces");
    if (D.solver.dynamic_angle && (!D.gen_angle_forces || D.n_gen_angle < 2))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: missing generalised angle forces");
    if (!(D.dt > 0) || !(D.fil_length > 0)) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_create: dt, length > 0");
    tfs_beat* b = new tfs_beat();
    b->desc = D;
    int rc = tfs_solver_create(&D.solver, &b->solver);
    if (rc) {
        delete b;
        return rc;
    }
    tfs_solver* s = b->solver;
    const int ngp = D.solver.dynamic_phase ? D.n_gen_phase : 1, nga = D.solver.dynamic_angle ? D.n_gen_angle : 1;
    const int64_t n = s->dims.n;
    std::vector<std::pair<double**, int64_t>> parts = {
        {&b->frames, 9LL * nf}, {&b->bases, 3LL * nf}, {&b->s_values, nseg}, {&b->phase, nf}, {&b->angle, nf},
        {&b->omega0, nf}, {&b->gen_phase, ngp}, {&b->gen_angle, nga}, {&b->body, 9}, {&b->rhs, n}, {&b->sol, n},
        {&b->hist1, n}, {&b->hist2, n}};
    if (D.solver.free_bodies) {
        parts.push_back({&b->pose, 8});
        parts.push_back({&b->base_ref, 3LL * nf});
        parts.push_back({&b->frame_ref, 9LL * nf});
        parts.push_back({&b->blob_ref, std::max<int64_t>(3LL * D.solver.num_blobs, 1)});
    }
    int64_t total = 0;
    for (auto& p : parts) total += ((p.second + 31) / 32) * 32;
    if (hipMalloc(&b->arena, total * 8) != hipSuccess) {
        tfs_beat_destroy(b);
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_beat_create: device memory");
    }
    double* p = b->arena;
    for (auto& q : parts) {
        *q.first = p;
        p += ((q.second + 31) / 32) * 32;
    }
    std::vector<double> sv(nseg), zeros(std::max(nf, 1), 0.0), body(9, 0.0);
    for (int i = 0; i < nseg; ++i) sv[i] = D.s_values ? D.s_values[i] : double(i) / double(nseg - 1);
    for (int i = 0; i < 3; ++i) body[i] = D.body_x[i];
    // a swimming body's velocity is solved for: no prescribed motion in the rhs
    for (int i = 0; i < 6; ++i) body[3 + i] = D.solver.free_bodies ? 0.0 : D.body_v[i];
    b->free_body = D.solver.free_bodies != 0;
    hipStream_t st = s->stream;
    hipError_t e = hipMemsetAsync(b->arena, 0, total * 8, st);
    auto up = [&](double* dst, const double* src, int64_t cnt) {
        if (e == hipSuccess && cnt > 0) e = hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyHostToDevice, st);
    };
    up(b->frames, D.frames, 9LL * nf);
    up(b->bases, D.bases, 3LL * nf);
    up(b->s_values, sv.data(), nseg);
    up(b->phase, D.phase0, nf);
    up(b->angle, D.angle0 ? D.angle0 : zeros.data(), nf);
    up(b->omega0, D.omega0, nf);
    if (D.solver.dynamic_phase) up(b->gen_phase, D.gen_phase_forces, D.n_gen_phase);
    if (D.solver.dynamic_angle) up(b->gen_angle, D.gen_angle_forces, D.n_gen_angle);
    up(b->body, body.data(), 9);
    if (e == hipSuccess && D.solver.num_blobs > 0)
        e = hipMemcpyAsync(s->rpy->x_blobs, D.x_blobs, 24LL * D.solver.num_blobs, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->free_body) {
        // body-frame references: the given lab positions relative to the
        // initial centre, initial orientation the identity
        std::vector<double> pose = {D.body_x[0], D.body_x[1], D.body_x[2], 1.0, 0.0, 0.0, 0.0, 0.0};
        std::vector<double> bref(3 * (size_t)nf), xref(3 * (size_t)std::max(D.solver.num_blobs, 0));
        for (int m = 0; m < nf; ++m)
            for (int k = 0; k < 3; ++k) bref[3 * m + k] = D.bases[3 * m + k] - D.body_x[k];
        for (int i = 0; i < D.solver.num_blobs; ++i)
            for (int k = 0; k < 3; ++k) xref[3 * i + k] = D.x_blobs[3 * i + k] - D.body_x[k];
        e = hipMemcpyAsync(b->pose, pose.data(), 64, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->base_ref, bref.data(), 24LL * nf, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(b->frame_ref, D.frames, 72LL * nf, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && D.solver.num_blobs > 0)
            e = hipMemcpyAsync(b->blob_ref, xref.data(), 24LL * D.solver.num_blobs, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    invalidate_near(s->rpy);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    b->beat_A.assign(D.beat_A, D.beat_A + 24);
    b->beat_B.assign(D.beat_B, D.beat_B + 18);
    if (e != hipSuccess) {
        tfs_beat_destroy(b);
        return hip_fail(e, "tfs_beat_create: upload");
    }
    *out = b;
    return TFS_OK;
}

int tfs_beat_destroy(tfs_beat* b) {
    if (!b) return TFS_OK;
    if (b->solver && b->solver->stream) hipStreamSynchronize(b->solver->stream);
    if (b->arena) hipFree(b->arena);
    if (b->vel) hipFree(b->vel);
    tfs_solver_destroy(b->solver);
    delete b;
    return TFS_OK;
}

int tfs_beat_step(tfs_beat* b, tfs_solve_stats* stats) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_step: null handle");
    tfs_solver* s = b->solver;
    const tfs_beat_desc& D = b->desc;
    hipStream_t st = s->stream;
    // The beat coefficients live in constant memory shared by every handle.
    int rc = check(k_beat_coeffs(b->beat_A.data(), b->beat_B.data()), "beat coefficients");
    if (rc) return rc;
    // filament::initial_guess: explicit phase/angle update from the last solve
    // (and, for a swimming body, its position and orientation).
    if (b->step > 0) {
        rc = check(k_phase_update(b->sol, s->dims, D.dt, b->phase, b->angle, st), "phase update");
        if (!rc && b->free_body) rc = check(k_body_advance(b->sol, s->dims.off_body, D.dt, b->pose, st), "body update");
        if (rc) return rc;
    }
    if (b->free_body) {
        rc = check(k_body_pose(b->pose, b->base_ref, b->frame_ref, b->blob_ref, D.solver.num_fil, D.solver.num_blobs,
                               b->bases, b->frames, s->rpy->x_blobs, b->body, st),
                   "body pose");
        if (!rc) rc = tfs_solver_set_bodies(s, b->body, TFS_MEMORY_DEVICE);
        if (rc) return rc;
        s->blob_pc_valid = false;
    }
    BeatArgs ba;
    ba.nfil = D.solver.num_fil;
    ba.nseg = D.solver.num_seg;
    ba.length = D.fil_length;
    ba.frames = b->frames;
    ba.bases = b->bases;
    ba.s_values = b->s_values;
    ba.phase = b->phase;
    ba.angle = D.solver.dynamic_angle ? b->angle : nullptr;
    rc = check(k_beat(ba, s->rpy->x_segs, s->kphi, s->kth, st), "beat kinematics");
    if (rc) return rc;
    invalidate_near(s->rpy);
    s->fil_pc_valid = false;
    RhsArgs ra;
    ra.d = s->dims;
    ra.phase = b->phase;
    ra.angle = b->angle;
    ra.omega0 = b->omega0;
    ra.gen_phase = b->gen_phase;
    ra.n_gen_phase = D.n_gen_phase;
    ra.gen_angle = b->gen_angle;
    ra.n_gen_angle = D.n_gen_angle;
    ra.gen_force_factor = D.gen_force_factor;
    ra.torsion_factor = D.torsion_factor;
    ra.length = D.fil_length;
    ra.body_x = b->body;
    ra.body_v = b->body + 3;
    rc = check(k_rhs(ra, s->rpy->x_segs, s->rpy->x_blobs, b->rhs, st), "rhs");
    if (rc) return rc;
    tfs_solve_stats local{};
    tfs_solve_stats* sst = stats ? stats : &local;
    // Warm start: the previous step's solution (mobility_solver.cpp:2201-2293),
    // or, if asked for, its extrapolation from the solutions before it.
    if (b->step > 0) {
        const int order = (int)std::min<int64_t>(b->warm_order, b->step - 1);
        rc = check(k_extrapolate(b->sol, b->hist1, b->hist2, (int)s->dims.n, order, st), "warm start");
        if (rc) return rc;
    }
    const int solve_rc = solve_device(s, b->rhs, b->sol, sst);
    b->step++;
    return solve_rc;
}

int tfs_beat_set_warm_start(tfs_beat* b, int32_t order) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_set_warm_start: null handle");
    if (order < 0 || order > 2) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_set_warm_start: order must be 0, 1 or 2");
    b->warm_order = order;
    return TFS_OK;
}

int tfs_beat_state(tfs_beat* b, double* phase, double* angle, double* x_segs, double* solution) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_state: null handle");
    tfs_solver* s = b->solver;
    const int nf = b->desc.solver.num_fil;
    hipError_t e = hipSuccess;
    auto down = [&](double* dst, const double* src, int64_t cnt) {
        if (e == hipSuccess && dst && cnt > 0) e = hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyDeviceToHost, s->stream);
    };
    down(phase, b->phase, nf);
    down(angle, b->angle, nf);
    down(x_segs, s->rpy->x_segs, 3LL * nf * b->desc.solver.num_seg);
    down(solution, b->sol, s->dims.n);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return check(e, "tfs_beat_state");
}

int tfs_beat_restore(tfs_beat* b, const double* phase, const double* angle, const double* solution) {
    if (!b || !phase) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_restore: null handle or phases");
    tfs_solver* s = b->solver;
    const int nf = b->desc.solver.num_fil;
    const int64_t n = s->dims.n;
    std::vector<double> zeros(std::max<int64_t>(n, nf), 0.0);
    hipError_t e = hipMemcpyAsync(b->phase, phase, 8LL * nf, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->angle, angle ? angle : zeros.data(), 8LL * nf, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->sol, solution ? solution : zeros.data(), 8 * n, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_beat_restore");
    // The next step solves at the restored phases (no update) and starts from
    // `solution`; the extrapolation history begins again.
    b->step = 0;
    return TFS_OK;
}

int tfs_beat_velocities(tfs_beat* b, double* v_segs, double* v_blobs) {
    if (!b || !v_segs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_velocities: null handle or segment array");
    tfs_solver* s = b->solver;
    const SaddleDims& dm = s->dims;
    const int64_t ns = int64_t(dm.nfil) * dm.nseg, nb = dm.nblob;
    if (nb > 0 && !v_blobs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_velocities: null blob array");
    if (!b->vel && hipMalloc(&b->vel, 8 * std::max<int64_t>(3 * (ns + nb), 1)) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_beat_velocities: device memory");
    // M lambda in FP64 at the current positions (mobility_solver.cpp:3053-3058).
    int rc = run_matvec(s->rpy, TFS_BLOCK_ALL, TFS_PRECISION_FP64, CView{b->sol, 3, 1},
                        CView{b->sol + dm.off_blob, 3, 1}, View{b->vel, 3, 1}, View{b->vel + 3 * ns, 3, 1}, 3);
    if (rc) return rc;
    std::vector<double> v(3 * (ns + nb));
    hipError_t e = hipMemcpyAsync(v.data(), b->vel, 8 * v.size(), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_beat_velocities");
    for (int64_t i = 0; i < ns; ++i)
        for (int c = 0; c < 6; ++c) v_segs[6 * i + c] = c < 3 ? v[3 * i + c] : 0.0;
    if (nb > 0) std::copy(v.begin() + 3 * ns, v.end(), v_blobs);
    return TFS_OK;
}

int tfs_beat_body(tfs_beat* b, double* x, double* q, double* v) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_body: null handle");
    tfs_solver* s = b->solver;
    double pose[8] = {b->desc.body_x[0], b->desc.body_x[1], b->desc.body_x[2], 1.0, 0.0, 0.0, 0.0, 0.0};
    double vel[6] = {0, 0, 0, 0, 0, 0};
    hipError_t e = hipSuccess;
    if (b->free_body) {
        e = hipMemcpyAsync(pose, b->pose, 64, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vel, b->sol + s->dims.off_body, 48, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    } else {
        for (int k = 0; k < 6; ++k) vel[k] = b->desc.body_v[k];
    }
    if (e != hipSuccess) return hip_fail(e, "tfs_beat_body");
    if (x) std::copy(pose, pose + 3, x);
    if (q) std::copy(pose + 3, pose + 7, q);
    if (v) std::copy(vel, vel + 6, v);
    return TFS_OK;
}

int tfs_beat_solver(tfs_beat* b, tfs_solver** out) {
    if (!b || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_solver: null argument");
    *out = b->solver;
    return TFS_OK;
}

int tfs_beat_set_mode(tfs_beat* b, int32_t mode) {
    if (!b) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_set_mode: null handle");
    if (mode != TFS_SOLVE_FP64 && mode != TFS_SOLVE_MIXED)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_set_mode: unknown mode");
    if (mode == TFS_SOLVE_MIXED && (!(b->solver->desc.inner_tol > 0) || b->solver->desc.max_refinements < 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_beat_set_mode: mixed mode needs inner_tol and max_refinements");
    b->solver->desc.mode = mode;
    b->desc.solver.mode = mode;
    return TFS_OK;
}

}  // extern "C"
