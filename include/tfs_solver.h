/*
 * tfs_solver.h -- C-ABI of the per-timestep mobility solve on the MI355X.
 *
 * Replaces the generic solver layer of the reference's `mobility_solver`
 * (src/cilia/mobility/mobility_solver.cpp) for prescribed cilia on a body whose
 * velocity is prescribed (config.hpp: CILIA_TYPE 3, PRESCRIBED_BODY_VELOCITIES):
 *   system_matrix_mult_new   mobility_solver.cpp:1806-2090 -> tfs_solver_system_mult
 *   apply_preconditioner_new mobility_solver.cpp:1040-1513 -> tfs_solver_precondition
 *   solve_linear_system      mobility_solver.cpp:2092-2570 -> tfs_solver_solve
 *   assemble_rhs + compute_velocities + filament::initial_guess
 *                            mobility_solver.cpp:240-440, 2574-2659,
 *                            filament.cpp:1083-1273       -> tfs_beat_step
 *
 * Unknown vector (mobility_solver.cpp:107-140), n = 3*Ns + 3*Nb + [Nfil] + [Nfil]:
 *   [ segment forces (xyz per segment) | blob forces (xyz per blob) |
 *     phase speeds (if dynamic_phase) | shape-rotation speeds (if dynamic_angle) ]
 *
 * The Krylov solver is GMRES(m) with right preconditioning and FP64 Arnoldi
 * (classical Gram-Schmidt applied twice; once in the inner solve of
 * TFS_SOLVE_MIXED, whose result the FP64 residual checks).  restart = min(n, 350) with
 * max_cycles = 1 and tol = 1e-4 reproduces the reference's single-cycle solve
 * (config.hpp:312-321); restarted cycles and mixed-precision iterative
 * refinement are additions.  All statuses are TFS_* codes of tfs_rpy.h.
 */
#ifndef TFS_SOLVER_H
#define TFS_SOLVER_H

#include <stddef.h>
#include <stdint.h>

#include "tfs_rpy.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TFS_SOLVE_FP64 0   /* every matvec in FP64 tiles                               */
#define TFS_SOLVE_MIXED 1  /* inner GMRES on FP32-tile matvecs, FP64 residuals outside */

#define TFS_PRECOND_REFERENCE 0  /* D = I/(6 pi mu a) in P = [D -K; -K^T 0]              */
#define TFS_PRECOND_BLOCKS 1     /* D = each filament's own mobility + blob patches     */

typedef struct tfs_solver_desc {
    int32_t num_fil;         /* NSWIM*NFIL                                          */
    int32_t num_seg;         /* NSEG (segments per filament)                        */
    int32_t num_blobs;       /* NSWIM*NBLOB                                         */
    int32_t dynamic_phase;   /* DYNAMIC_PHASE_EVOLUTION (config.hpp:100)            */
    int32_t dynamic_angle;   /* DYNAMIC_SHAPE_ROTATION (config.hpp:104)             */
    int32_t mode;            /* TFS_SOLVE_*                                          */
    double seg_radius, blob_radius, mu;
    int32_t restart;         /* Krylov vectors per cycle (reference: min(n, 350))   */
    int32_t max_cycles;      /* restart cycles per (inner) solve                     */
    double tol;              /* target ||b - A x|| / ||b|| (reference: 1e-4)         */
    double inner_tol;        /* mixed: reduction asked of each inner solve           */
    int32_t max_refinements; /* mixed: iterative-refinement steps                    */
    void *stream;            /* hipStream_t or NULL                                   */
    int32_t num_swim;        /* NSWIM (0 is taken as 1): filaments, segments and
                                blobs split evenly between the swimmers             */
    int32_t free_bodies;     /* 1: each swimmer's velocity V (6) is an unknown, with
                                force- and torque-balance rows (not
                                PRESCRIBED_BODY_VELOCITIES); body centres from
                                tfs_solver_set_bodies                                */
} tfs_solver_desc;

typedef struct tfs_solve_stats {
    int32_t iterations;       /* Krylov iterations (all cycles, all refinements)     */
    int32_t cycles;           /* GMRES cycles                                        */
    int32_t refinements;      /* mixed: outer refinement steps taken                 */
    int32_t fp64_matvecs;     /* RPY matvecs in FP64 tiles                           */
    int32_t fp32_matvecs;     /* RPY matvecs in FP32 tiles                           */
    int32_t converged;        /* 1 if the tolerance was met                          */
    double relres;            /* last relative residual (true residual in mixed mode,
                                 GMRES estimate of the last cycle otherwise)         */
    double initial_relres;    /* ||b - A x0|| / ||b|| of the warm start              */
} tfs_solve_stats;

typedef struct tfs_solver tfs_solver;

int tfs_solver_create(const tfs_solver_desc *desc, tfs_solver **out);
int tfs_solver_destroy(tfs_solver *s);
/* System size n. */
int tfs_solver_size(tfs_solver *s, int64_t *n);
/* Geometry of the current step: segment and blob positions (3 per particle)
 * and the phase / angle velocity directions vel_dir_phase, vel_dir_angle
 * (3 per segment, filament.cpp:1083-1273); NULL keeps what was set before.
 * `memory` is TFS_MEMORY_HOST or TFS_MEMORY_DEVICE. */
int tfs_solver_set_geometry(tfs_solver *s, const double *x_segs, const double *x_blobs,
                            const double *vel_dir_phase, const double *vel_dir_angle,
                            int32_t memory);
/* Body centres [3 per swimmer] of a free-body system (the x_body of
 * mobility_solver.cpp:1858-1870 and the blob displacements R b_ref = x_blob -
 * x_body).  Needed before the first product when free_bodies is set. */
int tfs_solver_set_bodies(tfs_solver *s, const double *x_bodies, int32_t memory);
/* out = A in (system_matrix_mult_new); precision is TFS_PRECISION_*.  With
 * free bodies the vector is [lambda_s | lambda_b | V | phi_dot | theta_dot]
 * (mobility_solver.cpp:107-140) and
 *   particle rows  M lambda - (V + Omega x (x - x_body)) - K_phi phi_dot - K_theta theta_dot,
 *   body rows      -sum lambda,  -sum (x - x_body) x lambda   (:1850-1908). */
int tfs_solver_system_mult(tfs_solver *s, const double *in, double *out, int32_t precision,
                           int32_t memory);
/* out = P^{-1} in (apply_preconditioner_new).  With free bodies, the exact
 * inverse of [D -K; -K^T 0] with K = [K_body K_phi K_theta]: the
 * preconditioner without body columns, P0, then a rank-6-per-swimmer
 * correction, y = -(E^T P0^-1 E)^-1 (r_body + E^T P0^-1 r), with E the body
 * columns; the 6 x 6 matrices play the part of the reference's KTMinvK_inv
 * (mobility_solver.cpp:1186-1330, swimmer.cpp:1154-1326). */
int tfs_solver_precondition(tfs_solver *s, const double *in, double *out, int32_t memory);
/* Right preconditioner of later solves.  TFS_PRECOND_REFERENCE (default):
 * P = [D -K; -K^T 0] with D = I/(6 pi mu a), apply_preconditioner_new
 * (mobility_solver.cpp:1040-1116).  TFS_PRECOND_BLOCKS: the same saddle form
 * with D the block-diagonal of every filament's own RPY mobility (3 NSEG
 * square, rebuilt when the segments move) and of the blobs' RPY mobility
 * within patches of 16 blobs neighbouring in Morton order (rebuilt when the
 * blobs move), each block inverted on the device.  No counterpart in the
 * reference: solves meet the same tolerance in fewer iterations.  Needs
 * num_seg <= 128. */
int tfs_solver_set_preconditioner(tfs_solver *s, int32_t kind);
/* Solve A x = rhs; x holds the warm start on entry (solve_linear_system,
 * including its zero-rhs and rubbish-guess rules, mobility_solver.cpp:2123-2322).
 * Returns TFS_ERR_NOT_CONVERGED (with x updated) if the tolerance is not met,
 * TFS_ERR_NON_FINITE if a residual is not finite. */
int tfs_solver_solve(tfs_solver *s, const double *rhs, double *x, int32_t memory,
                     tfs_solve_stats *stats);
/* ||rhs - A x|| / ||rhs|| with the given matvec precision. */
int tfs_solver_residual(tfs_solver *s, const double *rhs, const double *x, int32_t precision,
                        int32_t memory, double *relres);

/* Float twins of the array entry points above, for a Real = float build
 * whose matrix class holds floats (mobility_solver.cpp works on Real): the
 * arrays are widened exactly to FP64 on the device, the solve runs as its
 * double twin (FP64 Arnoldi, residuals and solution), and outputs are rounded
 * once to float.  Complete on return. */
int tfs_solver_set_geometry_f32(tfs_solver *s, const float *x_segs, const float *x_blobs,
                                const float *vel_dir_phase, const float *vel_dir_angle, int32_t memory);
int tfs_solver_set_bodies_f32(tfs_solver *s, const float *x_bodies, int32_t memory);
int tfs_solver_system_mult_f32(tfs_solver *s, const float *in, float *out, int32_t precision,
                               int32_t memory);
int tfs_solver_precondition_f32(tfs_solver *s, const float *in, float *out, int32_t memory);
int tfs_solver_solve_f32(tfs_solver *s, const float *rhs, float *x, int32_t memory,
                         tfs_solve_stats *stats);
int tfs_solver_residual_f32(tfs_solver *s, const float *rhs, const float *x, int32_t precision,
                            int32_t memory, double *relres);

/* The RPY matvec handle the solver evaluates with (for tfs_rpy_profile,
 * tfs_rpy_set_tile_shape, ...).  Owned by the solver. */
int tfs_solver_rpy(tfs_solver *s, tfs_rpy **out);

/* Independent bodies solved in lock-step.  The solver's arrays hold
 * num_bodies bodies one after another (fil_per_body[i] filaments and
 * blobs_per_body[i] blobs each, summing to the solver's sizes); the matvec
 * couples particles of the same body only (tfs_rpy_set_bodies), and every
 * tfs_solver_solve runs one GMRES per body: each body has its own Krylov
 * basis, Hessenberg matrix, tolerance (tol times its own ||rhs||), warm-start
 * rules and refinement count, and one batched launch serves all bodies that
 * are still iterating -- a body that has met its tolerance drops out of the
 * matvec's work list and of the Arnoldi launches.  With graphs on (the
 * default) each GMRES cycle of `restart` steps is one captured hipGraph,
 * replayed while the positions' work list and scales are unchanged.  Fixed
 * bodies only (no free_bodies); restart <= 120.  No counterpart in the
 * reference, whose NSWIM bodies are all coupled (mobility_solver.cpp:1806-1908);
 * a solve here equals num_bodies separate one-body solves.  num_bodies = 1
 * returns to the single-system solve. */
int tfs_solver_set_independent_bodies(tfs_solver *s, int32_t num_bodies, const int32_t *fil_per_body,
                                      const int32_t *blobs_per_body);
/* Captured cycles on (1) or off (0: the cycle's steps are enqueued four at a
 * time with a device stop count read back between). */
int tfs_solver_set_graphs(tfs_solver *s, int32_t enable);
/* Per body, of the last solve: GMRES steps, whether its tolerance was met,
 * its final ||rhs - A x|| / ||rhs||, and the device time from the solve's
 * start to the body's last Arnoldi stop in ms (any argument may be null). */
int tfs_solver_body_stats(tfs_solver *s, int32_t *iterations, int32_t *converged, double *relres,
                          double *solve_ms);
/* GMRES cycles run as captured graphs and as eager launches since creation. */
int tfs_solver_cycle_counts(tfs_solver *s, int64_t *graph_cycles, int64_t *eager_cycles);

/* ---- Prescribed-beat timestep driver (one body, body velocity prescribed) ---- */
typedef struct tfs_beat_desc {
    tfs_solver_desc solver;
    double dt;                       /* DT                                             */
    double fil_length;               /* FIL_LENGTH = DL*(NSEG-1) (config.hpp:512)       */
    const double *frames;            /* 9*num_fil: e1 normal, e2 beat direction, e3     */
    const double *bases;             /* 3*num_fil: segment-0 positions                  */
    const double *x_blobs;           /* 3*num_blobs                                     */
    const double *s_values;          /* num_seg shape parameters, NULL: n/(NSEG-1)      */
    const double *omega0;            /* num_fil natural frequencies                     */
    const double *phase0;            /* num_fil initial phases                          */
    const double *angle0;            /* num_fil initial shape-rotation angles or NULL   */
    const double *beat_A;            /* 24: Fourier cosine coefficients [x|y][4][3]     */
    const double *beat_B;            /* 18: Fourier sine coefficients  [x|y][3][3]      */
    const double *gen_phase_forces;  /* reference generalised phase forces, one period  */
    int32_t n_gen_phase;
    const double *gen_angle_forces;  /* reference generalised angle forces, one period  */
    int32_t n_gen_angle;
    double gen_force_factor;         /* GEN_FORCE_MAGNITUDE_FACTOR                      */
    double torsion_factor;           /* TORSIONAL_SPRING_MAGNITUDE_FACTOR               */
    double body_x[3];                /* body centre                                     */
    double body_v[6];                /* prescribed body velocity and angular velocity  */
} tfs_beat_desc;

typedef struct tfs_beat tfs_beat;

int tfs_beat_create(const tfs_beat_desc *desc, tfs_beat **out);
int tfs_beat_destroy(tfs_beat *b);
/* One timestep: phase update from the previous solution (step > 0), beat
 * kinematics, right-hand side, warm-started solve. */
int tfs_beat_step(tfs_beat *b, tfs_solve_stats *stats);
/* Host copies of the current state (any pointer may be NULL). */
int tfs_beat_state(tfs_beat *b, double *phase, double *angle, double *x_segs, double *solution);
/* Resume from saved state (READ_INITIAL_CONDITIONS_FROM_BACKUP,
 * cilia_sim_main.cu:118-160, filament.cpp:85-130): the phases, the angles
 * (NULL: zero) and the solution vector the next solve starts from (NULL:
 * zero).  The next tfs_beat_step solves at these phases without advancing
 * them first, as the reference's first step after a restart re-solves the
 * saved step. */
int tfs_beat_restore(tfs_beat *b, const double *phase, const double *angle, const double *solution);
/* Host velocities of the last solution's forces, M lambda in FP64 at the
 * current positions: v_segs [6*Ns] (rotational slots 0, as the prescribed
 * build writes them) and v_blobs [3*Nb] (may be NULL without blobs).  This is
 * the product the reference makes before saving velocities
 * (mobility_solver.cpp:3053-3058, written at :3149). */
int tfs_beat_velocities(tfs_beat *b, double *v_segs, double *v_blobs);
/* Body state: centre x [3], orientation q [4] (w, x, y, z) and velocity v [6]
 * (V, Omega; the last solve's for a swimming body).  With
 * desc.solver.free_bodies the body swims: its velocity is solved for with
 * force- and torque-free rows, and each step first moves it explicitly,
 * x += dt V, q = exp(dt Omega) q (cilia_sim_main.cu:677-693), carrying the
 * filament bases, frames and blobs with it (rigid_body::blob_positions,
 * rigid_body.cpp:482-500); the positions given at creation are the body frame
 * at orientation 1. */
int tfs_beat_body(tfs_beat *b, double *x, double *q, double *v);
/* The solver handle inside the driver (to call tfs_solver_* on its state). */
int tfs_beat_solver(tfs_beat *b, tfs_solver **out);
/* Switch between TFS_SOLVE_FP64 and TFS_SOLVE_MIXED for later steps. */
int tfs_beat_set_mode(tfs_beat *b, int32_t mode);
/* Initial guess of later solves.  0 (default): the previous step's solution,
 * as the reference (mobility_solver.cpp:2201-2293).  1 / 2: its linear /
 * quadratic extrapolation in time from the last two / three solutions (used
 * once that many exist).  No counterpart in the reference: the solve meets the
 * same tolerance from a closer start. */
int tfs_beat_set_warm_start(tfs_beat *b, int32_t order);

/* ---- Elastic-filament velocities (non-prescribed cilia, body velocity prescribed) ----
 * mobility_solver::compute_velocities without PRESCRIBED_CILIA
 * (mobility_solver.cpp:2598-2655): the velocities of given segment forces and
 * torques, with the blob forces that hold the body surfaces to the rigid-body
 * motion V + Omega x (x_blob - x_body) (assemble_rhs, mobility_solver.cpp:439-570,
 * PRESCRIBED_BODY_VELOCITIES):
 *   v_segs  = M_ss f_segs                              (Mss_mult '=')
 *   v_blobs = M_bs f_segs                              (Mbs_mult '=')
 *   M_bb f_blobs = -v_blobs + K V_body                 (GMRES on M_bb, preconditioner
 *                                                       6 pi mu RBLOB, mobility_solver.cpp:1439-1443)
 *   v_segs += M_sb f_blobs                             (Msb_mult '+=')
 * Segments carry 6 values (force and torque; velocity and angular velocity). */
typedef struct tfs_velocities_desc {
    int32_t num_swim;          /* NSWIM: bodies, each with the same segment and blob counts */
    int32_t segs_per_swim;     /* NFIL*NSEG                                                 */
    int32_t blobs_per_swim;    /* NBLOB (>= 1)                                              */
    int32_t mode;              /* TFS_SOLVE_* of the blob solve; the segment blocks are
                                  FP64 in TFS_SOLVE_FP64 and FP32-tile in TFS_SOLVE_MIXED   */
    double seg_radius, blob_radius, mu;
    int32_t restart;           /* GMRES restart of the blob solve                           */
    int32_t max_cycles;
    double tol;                /* relative residual of the blob solve                       */
    double inner_tol;          /* mixed mode                                                */
    int32_t max_refinements;   /* mixed mode                                                */
    int32_t memory;            /* TFS_MEMORY_* of the arrays passed to tfs_velocities_compute */
    void *stream;              /* hipStream_t or NULL                                        */
    int32_t free_bodies;       /* 1: bodies free, use tfs_velocities_compute_free           */
} tfs_velocities_desc;

typedef struct tfs_velocities tfs_velocities;

int tfs_velocities_create(const tfs_velocities_desc *desc, tfs_velocities **out);
int tfs_velocities_destroy(tfs_velocities *v);
/* x_segs 3 and f_segs 6 per segment, x_blobs 3 per blob, x_bodies 3 and
 * v_bodies 6 (V, Omega) per body; out: v_segs 6 per segment.  f_blobs (3 per
 * blob) holds the warm start on entry and the solved blob forces on return.
 * Complete on return.  TFS_ERR_NOT_CONVERGED if the blob solve misses its
 * tolerance (outputs still written). */
/* How the blob forces are solved for: GMRES preconditioned with
 * TFS_PRECOND_REFERENCE (6 pi mu a, the reference's, default) or
 * TFS_PRECOND_BLOCKS (the blobs' mobility inverted within 16-blob patches),
 * or TFS_BLOB_DIRECT: the dense blob mobility (3 NBLOB square, equal blob radii)
 * inverted on the device by blocked Gauss-Jordan, then one dense product per
 * solve.  Both the patches and the inverse are rebuilt only when the blob
 * positions change; the inverse holds 8 (3 NBLOB)^2 bytes of device memory
 * (0.29 GB at 2,000 blobs, 5.8 GB at 9,000).  No counterpart in the
 * reference, whose GMRES runs on every call. */
#define TFS_BLOB_DIRECT 2
int tfs_velocities_set_preconditioner(tfs_velocities *v, int32_t kind);
int tfs_velocities_compute(tfs_velocities *v, const double *x_segs, const double *f_segs,
                           const double *x_blobs, const double *x_bodies, const double *v_bodies,
                           double *v_segs, double *f_blobs, tfs_solve_stats *stats);
/* Free bodies (a handle created with free_bodies; GMRES blob solves only):
 * the blob forces and each body's velocity V, Omega solve the blob-only
 * system with body rows (mobility_solver.cpp:544-560 and the operator of
 * :1806-2090): M_bb f_b - K V = -v_blobs, -K^T f_b = -(f_ext - sum over the
 * body's segments of (F, T + (x - X) x F)), the force- and torque-free
 * condition (swimmer.cpp:469-533).  f_bodies: external force and torque per
 * body (6 each) or NULL for none; v_bodies (6 per body) holds the warm start
 * on entry and the solved velocities on return; the rest as above.
 * q_bodies (4 per body, the orientation of the pose the blobs are at) is used
 * by the direct solve only and may be NULL otherwise.
 * Direct solve with one free body (TFS_BLOB_DIRECT, num_swim 1, after
 * tfs_velocities_set_body_frame): the blob mobility at pose (X, R) is
 * (I x R) M_ref (I x R^T) with M_ref that of the body-frame positions, so its
 * inverse A, Z = A K_ref and the body mobility N = (K_ref^T Z)^{-1} are
 * formed once; each call is then one dense product: y = A R^T r,
 * V_ref = N (R^T f - K_ref^T y), f_b = R (y + Z V_ref).  No counterpart in
 * the reference, which runs GMRES on every call. */
int tfs_velocities_compute_free(tfs_velocities *v, const double *x_segs, const double *f_segs,
                                const double *x_blobs, const double *x_bodies, const double *f_bodies,
                                const double *q_bodies, double *v_segs, double *f_blobs, double *v_bodies,
                                tfs_solve_stats *stats);
/* Float twins of tfs_velocities_compute / tfs_velocities_compute_free
 * (compute_velocities of a Real = float build): inputs widened exactly to
 * FP64 on the device, the FP64 computation, outputs (and the warm starts they
 * overwrite) rounded once to float.  Complete on return. */
int tfs_velocities_compute_f32(tfs_velocities *v, const float *x_segs, const float *f_segs,
                               const float *x_blobs, const float *x_bodies, const float *v_bodies,
                               float *v_segs, float *f_blobs, tfs_solve_stats *stats);
int tfs_velocities_compute_free_f32(tfs_velocities *v, const float *x_segs, const float *f_segs,
                                    const float *x_blobs, const float *x_bodies, const float *f_bodies,
                                    const float *q_bodies, float *v_segs, float *f_blobs, float *v_bodies,
                                    tfs_solve_stats *stats);
/* The body-frame blob positions (3 per blob, relative to the centre, at
 * orientation 1) for the direct free-body solve. */
int tfs_velocities_set_body_frame(tfs_velocities *v, const double *blob_refs);
/* How a call learns whether the blob positions changed since the last one
 * (the blob solve's geometry, or the dense inverse, is renewed when they did):
 * TFS_BLOBS_CHECK compares them on the device (one flag read back per call,
 * the default), TFS_BLOBS_FIXED takes them as unchanged after the first call
 * (a body held still), TFS_BLOBS_MOVING as changed on every call (a free
 * body).  No counterpart in the reference. */
#define TFS_BLOBS_CHECK 0
#define TFS_BLOBS_FIXED 1
#define TFS_BLOBS_MOVING 2
int tfs_velocities_set_blob_motion(tfs_velocities *v, int32_t motion);

#ifdef __cplusplus
}
#endif

#endif /* TFS_SOLVER_H */
