/*
 * tfs_filament.h -- C-ABI of the Cosserat filament model on the device: the
 * per-filament work of the elastic implicit timestep around the Broyden solver
 * (tfs_broyden.h) and the mobility (tfs_solver.h tfs_velocities_*).
 *
 * Replaces, for every filament of every swimmer at once:
 *   filament::initial_guess        filament.cpp:1083-1385 (with segment::initial_guess,
 *                                  and accept_state_from_rigid_body :649-672 for a body
 *                                  held still)
 *   filament::internal_forces_and_torques   filament.cpp:1467-1575
 *   filament::invert_approx_jacobian        filament.cpp:1761-1894 (the matrix; its
 *                                  inversion is tfs_broyden_set_jacobian) and
 *   filament::jacobian_lie_algebra_block    filament.cpp:1577-1759
 *   filament::update, robot_arm    filament.cpp:1896-1930, 625-647
 * for the reference's clamped-base build with a follower end force
 * (INSTABILITY_CILIA): the base segment is clamped by a multiplier, a force
 * -END_FORCE_MAGNITUDE t acts on the last segment, and an optional kick acts
 * on two segments during the first step (filament.cpp:1533-1552); and, with
 * tfs_filament_set_base, for its GEOMETRIC_CILIA (driven free base) and
 * CONSTANT_BASE_ROTATION (turning clamped base) builds.
 *
 * Layouts: segments body -> filament -> segment (as x_segs, f_segs of
 * tfs_rpy.h); quaternions (w, x, y, z); per filament the 3 (NSEG + 1)
 * multipliers [tether, lambda_0 .. lambda_{NSEG-2}, clamp].  The Broyden
 * vector and J0 blocks follow swimmer ordering: per swimmer its NFIL filament
 * blocks of 6 NSEG unknowns ([tether, lambda..., clamp, u_1 ...]) and then 6
 * body unknowns; J0 as tfs_broyden_set_jacobian takes it, per swimmer NFIL
 * (6 NSEG)^2 row-major blocks then a 6 x 6 body block (left to the caller).
 */
#ifndef TFS_FILAMENT_H
#define TFS_FILAMENT_H

#include <stddef.h>
#include <stdint.h>

#include "tfs_rpy.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tfs_filament_desc {
    int32_t num_swim;          /* NSWIM                                        */
    int32_t fil_per_swim;      /* NFIL                                         */
    int32_t num_seg;           /* NSEG (>= 2)                                  */
    double dl;                 /* DL                                           */
    double kb, kt;             /* KB, KT                                       */
    double mu, radius;         /* MU, RSEG                                     */
    double dt;                 /* DT                                           */
    double end_force;          /* END_FORCE_MAGNITUDE                          */
    double strain_twist[3];    /* intrinsic strain and twist (zero: straight)  */
    int32_t memory;            /* TFS_MEMORY_* of the arrays passed later      */
    void *stream;              /* hipStream_t or NULL                          */
} tfs_filament_desc;

typedef struct tfs_filament tfs_filament;

int tfs_filament_create(const tfs_filament_desc *desc, tfs_filament **out);
int tfs_filament_destroy(tfs_filament *h);
/* The base segment (the reference's CILIA_TYPE, config.hpp:346-348):
 *   TFS_BASE_CLAMPED (default): held by the clamp multiplier, the Broyden
 *     unknowns' clamp slot (INSTABILITY_CILIA, CILIA_TYPE 0).  values, if
 *     given, are per filament the Lie-algebra element [3] the base turns by
 *     every step (CONSTANT_BASE_ROTATION, CILIA_TYPE 2: BASE_ROTATION_RATE DT
 *     about the base's tangent, filament.cpp:487-492); NULL keeps it still.
 *   TFS_BASE_DRIVEN: free, with its Lie-algebra element in the clamp's slot of
 *     the unknowns (filament.cpp:1918-1927), its rotation moving every later
 *     segment in J0 (:1789-1800), and values per filament the lab-frame
 *     torque [3] on segment 0 (GEOMETRIC_CILIA, CILIA_TYPE 1: the reference's
 *     base_torque_magnitude_factor DEFAULT_BASE_TORQUE_MAGNITUDE R(body_q)
 *     (my_vertical x fast_beating_direction), filament.cpp:1523-1531, whose
 *     factor keeps its initial value because find_my_angle() returns 0,
 *     :1433-1447); NULL: no torque.  Bodies held still only.
 * The end force and first-step kick stay as set (zero for CILIA_TYPE 1, 2).
 * The reference's CILIA_TYPE 2 branch does not compile as shipped
 * (clamp_lambda exists only for CILIA_TYPE 0, filament.hpp:66-70) and would
 * reset the base's rotation to the body's every step (filament.cpp:660-665);
 * this follows its evident intent. */
#define TFS_BASE_CLAMPED 0
#define TFS_BASE_DRIVEN 1
int tfs_filament_set_base(tfs_filament *h, int32_t mode, const double *values);
/* Initial state (filament::initial_setup): segment-0 positions [3 per
 * filament] and quaternions [4 per segment]; multipliers and Lie-algebra
 * elements zero, the positions from the robot arm, their history equal. */
int tfs_filament_set_state(tfs_filament *h, const double *bases, const double *q);
/* First-step kick: forces [6 per filament: on seg_a, then on seg_b] added
 * while nt == 0 (the reference's perturbation1/2 on segments 4 and 12). */
int tfs_filament_set_perturbation(tfs_filament *h, int32_t seg_a, int32_t seg_b, const double *forces);
/* Start of step nt: history shift, extrapolated Lie-algebra elements and
 * multipliers, robot arm. */
int tfs_filament_initial_guess(tfs_filament *h, int32_t nt);
/* Internal forces and torques at the current state, f_segs [6 per segment]
 * (overwritten). */
int tfs_filament_forces(tfs_filament *h, int32_t nt, double *f_segs);
/* J0 filament blocks of step nt at the current state, with f_segs the
 * internal forces there; written into `blocks` at the swimmer layout above
 * (body blocks untouched). */
int tfs_filament_jacobian(tfs_filament *h, int32_t nt, const double *f_segs, double *blocks);
/* Adds a Broyden update (the whole vector, swimmer layout) and re-runs the
 * robot arm. */
int tfs_filament_update(tfs_filament *h, const double *update);

/* ---- Free bodies (PRESCRIBED_BODY_VELOCITIES false) ----
 * Makes each swimmer's body a state of its own: position, Lie-algebra element
 * and orientation (rigid_body.cpp:395-447, 506-535).  Filament f is then
 * anchored at X + R(q) fil_refs[f] and its base segment turns with the body
 * (swimmer::initial_guess, swimmer::update and
 * filament::accept_state_from_rigid_body, swimmer.cpp:433-467, 1072-1098,
 * filament.cpp:649-672): tfs_filament_initial_guess advances the body first,
 * tfs_filament_update applies the body's 6 entries of the update first.
 * x_bodies [3], q_bodies [4] per swimmer (the starting pose, at rest);
 * fil_refs [3 per filament] in the body frame; blob_refs [3 per blob],
 * blobs_per_swim per swimmer, for tfs_filament_blob_positions. */
int tfs_filament_set_body(tfs_filament *h, const double *x_bodies, const double *q_bodies, const double *fil_refs,
                          int32_t blobs_per_swim, const double *blob_refs);
/* rigid_body::blob_positions (rigid_body.cpp:482-500) at the current pose:
 * x_blobs [3 per blob], swimmer by swimmer. */
int tfs_filament_blob_positions(tfs_filament *h, double *x_blobs);
/* The body's part of J0 (swimmer::prepare_jacobian_inv, GMRES branch with a
 * free body, swimmer.cpp:688-880), after tfs_filament_jacobian at the same
 * state: per filament A (6 NSEG x 6, its rows and the body's columns) and B
 * (6 x 6 NSEG, the body's rows and its columns), row-major, swimmer by
 * swimmer; per swimmer S0 (6 x 6), so that J0 = [[diag J_i, A], [B, S0]] (see
 * tfs_broyden_set_body_coupling).  body_mobility: per swimmer the 6 x 6 body
 * mobility in the body frame (make_body_reference_matrices,
 * mobility_solver.cpp:2950-3047); sphere_radius > 0 adds the RPY couplings of
 * the segments with the body as a sphere of that radius (the reference does
 * so for spherical bodies, swimmer.cpp:818-843); 0 leaves them out. */
int tfs_filament_body_coupling(tfs_filament *h, int32_t nt, const double *body_mobility, double sphere_radius,
                               double *A, double *B, double *S0);
/* The same for the Broyden-for-everything form with free bodies
 * (swimmer::prepare_jacobian_inv, swimmer.cpp:565-676), after
 * tfs_filament_jacobian and with the blobs given to tfs_filament_set_body: A
 * as above (the filaments' rows, the body's columns); B per filament the body's
 * force and torque balance against its unknowns (6 x 6 NSEG; the reference's
 * D); S0 per swimmer its 6 x 6 matrix before the Schur corrections -- the
 * blobs' terms with the diagonal blob mobility -6 pi mu blob_radius / DT (or
 * -9 pi mu a / DT in BDF2 steps) and the blob forces f_blobs [3 per blob], and
 * the filaments' terms; blob_disp [3 per blob] the blobs' displacements
 * Q_init ref.  For tfs_broyden_set_body_coupling_blobs. */
int tfs_filament_body_coupling_blobs(tfs_filament *h, int32_t nt, const double *f_blobs, double blob_radius,
                                     double *A, double *B, double *S0, double *blob_disp);

/* Waits for the handle's work (device-memory outputs are then complete). */
int tfs_filament_sync(tfs_filament *h);

/* Device arrays of the state (owned by the handle), e.g. for
 * tfs_broyden_compute_errors: */
#define TFS_FIL_X 0     /* 3 per segment, positions at n+1 */
#define TFS_FIL_XM1 1   /* at n                            */
#define TFS_FIL_XM2 2   /* at n-1                          */
#define TFS_FIL_U 3     /* 3 per segment, Lie algebra at n+1 */
#define TFS_FIL_UM1 4
#define TFS_FIL_Q 5     /* 4 per segment                   */
#define TFS_FIL_MULT 6  /* 3 (NSEG+1) per filament          */
/* free bodies only (tfs_filament_set_body), per swimmer:          */
#define TFS_FIL_BODY_X 7    /* 3, position at n+1                 */
#define TFS_FIL_BODY_XM1 8  /* at n                               */
#define TFS_FIL_BODY_XM2 9  /* at n-1                             */
#define TFS_FIL_BODY_U 10   /* 3, Lie-algebra element at n+1      */
#define TFS_FIL_BODY_UM1 11
#define TFS_FIL_BODY_Q 12   /* 4, orientation                     */
int tfs_filament_device_array(tfs_filament *h, int32_t which, double **ptr);
/* Host copy of one of those arrays (synchronises). */
int tfs_filament_read(tfs_filament *h, int32_t which, double *out);

/* ---- One elastic implicit timestep on the device ----
 * The time loop body of cilia_sim_main.cu:353-502 for filaments on bodies
 * held still (PRESCRIBED_BODY_VELOCITIES, GMRES for the blob forces) or, when
 * the filament handle has free bodies (tfs_filament_set_body), on free bodies:
 *   initial_guess; forces; velocities; errors;
 *   if too large: J0 = approximate Jacobian (body blocks I), inverted on the device;
 *   Broyden: update -> forces -> velocities -> errors -> end_of_iter, until the
 *   residual meets TOL or max_iter iterations.
 * Free bodies: the velocities handle is created with free_bodies; blob
 * positions follow the bodies (x_blobs, x_bodies and v_bodies of the
 * descriptor are not used), the blob forces and body velocities come from
 * tfs_velocities_compute_free, the residual has the body rows, and J0^{-1}
 * takes the body's Schur complement (tfs_filament_body_coupling,
 * tfs_broyden_set_body_coupling) with body_mobility and sphere_radius below.
 * Broyden for everything (USE_BROYDEN_FOR_EVERYTHING, bodies held still):
 * broyden_blobs = NBLOB makes the blob forces Broyden unknowns, 3 NBLOB per
 * swimmer after its body's 6 (the tfs_broyden handle is made with the blocks
 * [6 NSEG] * NFIL, [6], [3] * NBLOB per swimmer).  The velocities are one
 * plain matvec (Mss, Msb, Mbb, Mbs '+=', mobility_solver.cpp:2603-2640; the
 * velocities handle is not used and may be NULL), the residual has the blob
 * rows, f_blobs holds the blob forces, extrapolated at each step's start and
 * updated with the Broyden update (rigid_body.cpp:447-478, 506-520), and
 * J0^{-1} scales the blob rows by -6 pi mu RBLOB (swimmer.cpp:931-945).
 * The three handles are the caller's, created with TFS_MEMORY_DEVICE and sizes
 * that agree (tfs_velocities: num_swim bodies of fil_per_swim*num_seg segments;
 * tfs_broyden: the block sizes of the Broyden layout above).  Only the
 * too-large flag of each residual crosses to the host.  During each
 * tfs_elastic_step the velocities handle's blob motion
 * (tfs_velocities_set_blob_motion) is TFS_BLOBS_MOVING for free bodies and,
 * for bodies held still, TFS_BLOBS_CHECK at the step's first evaluation and
 * TFS_BLOBS_FIXED after it; the handle's own mode is restored on return. */
struct tfs_velocities;
struct tfs_broyden;
typedef struct tfs_elastic_desc {
    tfs_filament *filaments;
    struct tfs_velocities *velocities;
    struct tfs_broyden *broyden;
    const double *x_blobs;     /* device, 3 per blob (fixed: the bodies do not move)  */
    const double *x_bodies;    /* device, 3 per body                                  */
    const double *v_bodies;    /* device, 6 per body (zero: held still)               */
    double *f_blobs;           /* device, 3 per blob: warm start in, blob forces out  */
    double tol;                /* TOL (config.hpp:302)                                 */
    int32_t max_iter;          /* MAX_BROYDEN_ITER                                     */
    const double *body_mobility; /* free bodies: device, 36 per swimmer (body frame)   */
    double sphere_radius;      /* free bodies: see tfs_filament_body_coupling          */
    int32_t broyden_blobs;     /* > 0: NBLOB, the Broyden-for-everything form (below)  */
    double blob_radius;        /* RBLOB, for that form's matvec                         */
} tfs_elastic_desc;

typedef struct tfs_elastic_stats {
    int32_t broyden_iterations;
    int32_t gmres_iterations;  /* summed over the blob solves of the step             */
    int32_t converged;         /* 1 if the residual met TOL                            */
    int32_t velocity_solves;
    double error_norm;         /* ||residual|| of the last evaluation                  */
} tfs_elastic_stats;

typedef struct tfs_elastic tfs_elastic;
int tfs_elastic_create(const tfs_elastic_desc *desc, tfs_elastic **out);
int tfs_elastic_destroy(tfs_elastic *e);
/* Step nt.  TFS_ERR_NOT_CONVERGED if max_iter is reached (state kept),
 * TFS_ERR_NON_FINITE if a residual is not finite (cilia_sim_main.cu:469-487). */
int tfs_elastic_step(tfs_elastic *e, int32_t nt, tfs_elastic_stats *stats);
/* Captured Broyden iterations (no counterpart in the reference, whose loop
 * is host code, cilia_sim_main.cu:394-502): with graphs on (the default),
 * when the three handles share one stream, the pair blocks are FP64 and the
 * blob forces come from the dense direct solve or the Broyden-for-everything
 * matvec, each step captures one iteration (the engine's end of the previous
 * iteration, find_update, update, forces, velocities, residual) as a hipGraph
 * with its bookkeeping on the device, and replays it in batches of up to 8
 * per host wait (the state's update is a no-op once the iterations have
 * stopped; every other launch recomputes the same values); results equal the
 * launch-by-launch loop's bit for bit.  0: always launch by launch. */
int tfs_elastic_set_graphs(tfs_elastic *e, int32_t enable);
/* Broyden iterations run as graph replays and launch by launch since creation. */
int tfs_elastic_iteration_counts(tfs_elastic *e, int64_t *graph_iterations, int64_t *eager_iterations);

/* The state of the step in the reference's backup record, after its step
 * number (swimmer::write_backup, filament::write_backup and
 * rigid_body::write_backup, swimmer.cpp:1139-1150, filament.cpp:1952-2020,
 * rigid_body.cpp:547-572; read back by the initial_setup functions): per
 * swimmer, per filament per segment q (4), x, xm1, xm2, u, um1 (3 each), then
 * the tether multiplier and its two previous values, the clamp's, and per
 * link lambda and its two previous values (28 NSEG + 9 values), then the
 * body: q (4), x, xm1, xm2, u, um1 (19), then in the Broyden-for-everything
 * form per blob component the force and its two previous values (9 NBLOB).
 * Host memory, *count values (tfs_elastic_backup_size).  held_bodies: for
 * bodies held still, x (3) and q (4) per swimmer to record, or NULL for the
 * origin and no rotation; free bodies record their own state.  Restoring
 * sets each previous orientation to the current one, as the reference's
 * reader does, and leaves the velocity solve's warm start alone. */
int tfs_elastic_backup_size(tfs_elastic *e, int64_t *count);
int tfs_elastic_backup(tfs_elastic *e, const double *held_bodies, double *out);
int tfs_elastic_restore(tfs_elastic *e, const double *in);

#ifdef __cplusplus
}
#endif

#endif /* TFS_FILAMENT_H */
