// solver_kernels.h -- launchers of solver_kernels.hip (library-internal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tfs_internal.h"

namespace tfs {

// Offsets into the saddle-point vector
//   [ lambda_s (3*nfil*nseg) | lambda_b (3*nblob) | V (6*nswim, free bodies) |
//     phi_dot (nfil) | theta_dot (nfil) ]
// (mobility_solver.cpp:107-140).  Swimmer s owns segments
// [s*segs_per_swim, ...), blobs [s*blobs_per_swim, ...), filaments
// [s*fil_per_swim, ...).
struct SaddleDims {
    int nfil = 0, nseg = 0, nblob = 0;
    int dyn_phase = 0, dyn_angle = 0;
    int64_t off_blob = 0, off_phase = 0, off_angle = 0, n = 0;
    int free_body = 0, nswim = 1;
    int segs_per_swim = 0, blobs_per_swim = 0, fil_per_swim = 0;
    int64_t off_body = 0;
};

// Right preconditioner P = [D -K; -K^T 0], applied as its exact inverse.
// kind 0 (mobility_solver.cpp:1040-1116): D = I/(6 pi mu a), gs / gb = 6 pi mu a
// for segments / blobs.  kind 1: D = block-diagonal of each filament's own RPY
// mobility (3 nseg square) and of the blobs' mobility within patches of up to
// kPatchBlobs blobs (neighbours in Morton order); inverses stored row-major
// (they are symmetric, so column c is read as row c).  Per filament,
// fil_dk = D^-1 [K_phi K_theta] (two columns of 3 nseg) and fil_sinv the
// inverse of the 2 x 2 (or 1 x 1) Schur complement K^T D^-1 K.
constexpr int kPatchBlobs = 16;
constexpr int kMaxBlockFilSeg = 128;
struct PrecondArgs {
    int kind = 0;
    double gs = 0, gb = 0;
    const double* kphi = nullptr;
    const double* kth = nullptr;
    const double* fil_inv = nullptr;      // nfil blocks of (3 nseg)^2
    const double* fil_dk = nullptr;       // nfil x 2 x 3 nseg
    const double* fil_sinv = nullptr;     // nfil x 4
    const int32_t* patch_blob = nullptr;  // blob indices, patch after patch
    const int32_t* patch_start = nullptr; // npatch + 1 offsets into patch_blob
    const double* patch_inv = nullptr;    // npatch blocks of (3 * patch size)^2 ...
    const int64_t* patch_off = nullptr;   // ... at these offsets
    int npatch = 0;
};

struct BeatArgs {
    int nfil, nseg;
    double length;            // filament length L = DL*(NSEG-1)
    const double* frames;     // 9 per filament: e1 (normal), e2 (beat direction), e3
    const double* bases;      // 3 per filament: position of segment 0
    const double* s_values;   // nseg arclength parameters in [0, 1]
    const double* phase;      // nfil
    const double* angle;      // nfil or null
};

struct RhsArgs {
    SaddleDims d;
    const double* phase;
    const double* angle;
    const double* omega0;     // nfil natural frequencies
    const double* gen_phase;  // reference generalised phase forces over one period
    int n_gen_phase;
    const double* gen_angle;  // reference generalised angle forces over one period
    int n_gen_angle;
    double gen_force_factor;  // GEN_FORCE_MAGNITUDE_FACTOR
    double torsion_factor;    // TORSIONAL_SPRING_MAGNITUDE_FACTOR
    double length;
    const double* body_x;     // 3
    const double* body_v;     // 6: V, Omega
};

hipError_t k_dot_cols(const double* V, int64_t ld, int ncols, const double* w, int n, double* partial, int G,
                      hipStream_t s);
hipError_t k_reduce_cols(const double* partial, int G, int ncols, double* h, double* Hcol, hipStream_t s);
hipError_t k_gemv(const double* V, int64_t ld, int ncols, const double* h, double* w, int n, bool subtract,
                  hipStream_t s);
hipError_t k_norm(const double* w, int n, double* partial, int G, double* out, hipStream_t s);
hipError_t k_givens(double* H, int ldh, int j, double* cs, double* sn, double* g, const double* wnorm,
                    double* scal, hipStream_t s);
hipError_t k_scale(double* v, const double* w, const double* scal, int idx, int n, hipStream_t s);
hipError_t k_trisolve(const double* H, int ldh, const double* g, int k, double* y, hipStream_t s);
hipError_t k_axpby(double a, const double* x, double b, double* y, int n, hipStream_t s);
// x <- extrapolated warm start of the given order from x, h1, h2; h2 <- h1, h1 <- old x.
hipError_t k_extrapolate(double* x, double* h1, double* h2, int n, int order, hipStream_t s);
hipError_t k_sub(const double* x, const double* y, double* out, int n, hipStream_t s);
// One Arnoldi step on w against V(:, 0:j] by classical Gram-Schmidt, with
// passes = 2 (CGS2: three launches of G = arnoldi_blocks(n) workgroups,
// H(:, j) = h1 + h2) or passes = 1 (CGS: two launches, H(:, j) = h1); pa and
// pb hold G x (j+1) partial sums each.  w is left orthogonalised and G partial
// sums of ||w||^2 in pa for k_arnoldi_close_next.  Every launch does nothing
// while *stop is nonzero.
hipError_t k_arnoldi_step(const double* V, int64_t ld, int j, double* w, int n, double* pa, double* pb, int G,
                          double* H, int ldh, double* cs, double* sn, double* g, double* scal, double* est,
                          int32_t* stop, double abs_tol, int passes, hipStream_t s);
// Closes step j -- ||w|| from pa, H(j+1, j) rotated, Givens rotations
// updated, scal[0] = est[j] = |g(j+1)|, scal[1] = 1/||w||, *stop = j + 1 when
// est[j] <= abs_tol (-(j + 1) if it is not finite), the same value to
// *stop_mirror and est[j] to *est_mirror (host memory mapped for the device,
// either may be null) -- and, when v is not null, writes v = w / ||w|| and
// out = P^{-1} v in the same launch.  Nothing happens while *stop is nonzero.
hipError_t k_arnoldi_close_next(const double* pa, int G, double* H, int ldh, int j, double* cs, double* sn,
                                double* g, double* scal, double* est, int32_t* stop, int32_t* stop_mirror,
                                double* est_mirror, double abs_tol,
                                const double* w, double* v, double* out, const PrecondArgs& P,
                                const SaddleDims& d, hipStream_t s);
int arnoldi_blocks(int64_t n);
// v = w * scale; out = P^{-1} v.  scale = dev_scale[1] if dev_scale, else host_scale.
// Nothing happens while *skip (if given) is nonzero.
hipError_t k_scale_precond(const double* w, const double* dev_scale, double host_scale, double* v, double* out,
                           const PrecondArgs& P, const SaddleDims& d, const int32_t* skip, hipStream_t s);
hipError_t k_precond(const double* in, double* out, const PrecondArgs& P, const SaddleDims& d, hipStream_t s);
// Block preconditioner set-up (kind 1).  fil_blk / patch_blk receive the
// blocks of D (to be inverted in place by k_batched_inverse); k_fil_schur then
// forms fil_dk and fil_sinv from the inverted filament blocks.
hipError_t k_fil_blocks(const double* x_segs, int nfil, int nseg, double a_seg, double mu, double* fil_blk,
                        hipStream_t s);
hipError_t k_patch_blocks(const double* x_blobs, const int32_t* patch_blob, const int32_t* patch_start,
                          const int64_t* patch_off, int npatch, double a_blob, double mu, double* patch_blk,
                          hipStream_t s);
hipError_t k_fil_schur(const PrecondArgs& P, const SaddleDims& d, double* fil_dk, double* fil_sinv, hipStream_t s);
hipError_t k_beat_coeffs(const double* A, const double* B);
hipError_t k_beat(const BeatArgs& a, double* x_segs, double* kphi, double* kth, hipStream_t s);
hipError_t k_rhs(const RhsArgs& a, const double* x_segs, const double* x_blobs, double* rhs, hipStream_t s);
hipError_t k_phase_update(const double* x, const SaddleDims& d, double dt, double* phase, double* angle,
                          hipStream_t s);

// ---- independent bodies in lock-step ----------------------------------------
// One GMRES per body, every body's Krylov vectors contiguous in the Krylov
// layout (body b at [koff[b], koff[b+1])).  Per-body arrays: H at b*hstride
// (ldh x m), cs / sn / g / est / y at b*mstride (m + 1 each), scal at 2b,
// partial sums at b*pstride (G x (m+1)).  stop[b]: 0 running, j + 1 stopped
// at step j, -(j + 1) non-finite, kEnsInactive not taking part.
constexpr int32_t kEnsInactive = 0x40000000;
struct EnsArgs {
    const int64_t* koff = nullptr;
    int nbody = 0;
    double* V = nullptr;
    int64_t ld = 0;
    double* w = nullptr;
    double* H = nullptr;
    int ldh = 0;
    int64_t hstride = 0;
    double *cs = nullptr, *sn = nullptr, *g = nullptr, *est = nullptr;
    int mstride = 0;
    double* scal = nullptr;
    int32_t* stop = nullptr;
    int32_t* nstop = nullptr;         // bodies that stopped since it was cleared
    const double* abs_tol = nullptr;  // per body
    double *pa = nullptr, *pb = nullptr;
    int64_t pstride = 0;
    int64_t* stamps = nullptr;        // per body: wall clock of its last stop (may be null)
};
// Gram-Schmidt step j of every running body, G x nbody workgroups per launch:
// passes = 2 (CGS2, three launches) or 1 (CGS, two launches).
hipError_t k_ens_arnoldi_step(const EnsArgs& e, int j, int G, int passes, hipStream_t s);
// Close step j of every running body; with write_next, v_{j+1} into the basis.
hipError_t k_ens_close(const EnsArgs& e, int j, bool write_next, int G, hipStream_t s);
// v_0 = r / ||r|| and g = (||r||, 0, ...) for every running body (r in the Krylov layout).
hipError_t k_ens_open(const EnsArgs& e, const double* r, const double* rnorm, int G, hipStream_t s);
// Between layouts: map[i] is the operator index of Krylov index i.
hipError_t k_ens_permute(const double* src, double* dst, const int32_t* map, int64_t n, bool to_operator,
                         hipStream_t s);
// out[b] = ||v|| over body b's rows of a Krylov-layout vector.
hipError_t k_ens_norms(const double* v, const int64_t* koff, int nbody, double* out, hipStream_t s);
// Per body: the back substitution of the cycle's k_b steps (kdone[b] = k_b)
// and out_b = V_b y_b (Krylov layout).
hipError_t k_ens_update(const EnsArgs& e, int m, double* y, int32_t* kdone, double* out, int G, hipStream_t s);
// x = 0 (and r = rhs) over the operator rows of every body with flag[b] set.
hipError_t k_ens_reset(const int32_t* flag, const int64_t* koff, int nbody, const int32_t* map, double* x, double* r,
                       const double* rhs, int G, hipStream_t s);
// stamps[b] = the device wall clock, every body.
hipError_t k_ens_stamp(int64_t* stamps, int nbody, hipStream_t s);

// ---- free bodies ----------------------------------------------------------
// out[6 s + c] = sign * (sum over the particles i of swimmer s of
// (vec_i, (x_i - x_body_s) x vec_i))_c: the body rows K_b^T vec (one
// workgroup per swimmer).
hipError_t k_body_moments(const double* vec, const SaddleDims& d, const double* x_segs, const double* x_blobs,
                          const double* x_bodies, double sign, double* out, const int32_t* skip, hipStream_t s);
// Column c of E = [K_b; 0] for every swimmer: particle rows e_c (c < 3) or
// e_{c-3} x (x_i - x_body), zero elsewhere.
hipError_t k_body_column(double* col, int c, const SaddleDims& d, const double* x_segs, const double* x_blobs,
                         const double* x_bodies, hipStream_t s);
// The body part of P^{-1} on top of z0 = P0^{-1} r (P0: the preconditioner
// without body columns): y_s = -Sinv_s (r_body_s + t_s) with t = E^T z0,
// z body rows = y, other rows z += Z y (Z = P0^{-1} E, 6 columns of ldz).
// r is in * sc with sc = 1 / sqrt(sum of the G partials p_in) (as the Arnoldi
// close computes it) or host_sc when p_in is null; v (may be null) gets the
// body rows of r.
hipError_t k_body_correct(const double* in, double host_sc, const double* p_in, int G, double* v, double* z,
                          const double* Z, int64_t ldz, const double* t, const double* sinv, const SaddleDims& d,
                          const int32_t* skip, hipStream_t s);

// Dense blob-blob mobility of equal spheres into a row-major n_pad x n_pad
// matrix (identity past 3 nb).
hipError_t k_blob_mobility_dense(const double* x_blobs, int nb, double a, double mu, double* A, int n_pad,
                                 hipStream_t s);

// Swimming body: pose = (X[3], q[4]) on the device.  Lab-frame bases,
// frames, blob positions and body centre from body-frame references; and the
// explicit pose update from the solution's body velocity.
hipError_t k_body_pose(const double* pose, const double* base_ref, const double* frame_ref, const double* blob_ref,
                       int nfil, int nblob, double* bases, double* frames, double* x_blobs, double* body_x,
                       hipStream_t s);
hipError_t k_body_advance(const double* sol, int64_t off_body, double dt, double* pose, hipStream_t s);

}  // namespace tfs
