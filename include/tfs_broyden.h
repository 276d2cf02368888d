/*
 * tfs_broyden.h -- C-ABI of the device-resident Broyden solver.
 *
 * Replaces broyden_solver (src/cilia/broyden_solver.hpp:18-43,
 * src/cilia/broyden_solver.cpp:9-129): the limited-memory inverse-Jacobian
 * update of the implicit timestep, with FP64 storage of the C and D columns
 * (NBROY x MAX_BROYDEN_ITER each) in HBM.
 *
 *   find_update  (broyden_solver.cpp:26-67):
 *       update = -c J0^{-1} error - sum_{i<iter} (error . D_i) C_i
 *   end_of_iter  (broyden_solver.cpp:69-129), when the error is still too large:
 *       C_iter = -c J0^{-1} new_error - sum_{i<iter} (new_error . D_i) C_i
 *       D_iter = new_error - error;  C_iter, D_iter /= ||D_iter||;  error = new_error
 *     and always iter += 1.
 *
 * c is JACOBIAN_CONFIDENCE_FACTOR (config.hpp:298).  J0^{-1} -- the model's
 * approximate Jacobian inverse, swimmer::jacobian_inv_mult (swimmer.cpp:915-1070),
 * per-filament dense blocks of filament::invert_approx_jacobian
 * (filament.cpp:1761-1894) plus body blocks -- is given as block-diagonal dense
 * blocks, applied on the device.
 */
#ifndef TFS_BROYDEN_H
#define TFS_BROYDEN_H

#include <stddef.h>
#include <stdint.h>

#include "tfs_rpy.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tfs_broyden_desc {
    int32_t n;                    /* NBROY                                                  */
    int32_t max_iter;             /* MAX_BROYDEN_ITER (config.hpp:300)                      */
    double confidence;            /* JACOBIAN_CONFIDENCE_FACTOR                             */
    int32_t num_blocks;           /* diagonal blocks of J0^{-1}                             */
    const int32_t *block_sizes;   /* host array, num_blocks sizes in [1, 2048] summing to n */
    int32_t memory;               /* TFS_MEMORY_* of the vectors passed later               */
    void *stream;                 /* hipStream_t or NULL                                     */
} tfs_broyden_desc;
/* Device vectors passed to the engine must be ready when a call is made.  With
 * a caller stream, device outputs are ordered on it; with stream == NULL the
 * engine runs on a stream of its own and its device outputs are complete when
 * the call returns.  Host vectors are always complete on return. */

typedef struct tfs_broyden tfs_broyden;

int tfs_broyden_create(const tfs_broyden_desc *desc, tfs_broyden **out);
int tfs_broyden_destroy(tfs_broyden *b);
/* J0^{-1}: the blocks one after another, each row-major (sum of size^2 values),
 * in the desc's memory space (swimmer::prepare_jacobian_inv, swimmer.cpp:535-913). */
int tfs_broyden_set_jacobian_inverse(tfs_broyden *b, const double *blocks);
/* J0 itself, same layout: every block is inverted on the device by Gauss-Jordan
 * elimination with partial pivoting, one workgroup per block (replaces the
 * per-filament LAPACK getrf/getri of filament::invert_approx_jacobian,
 * filament.cpp:1761-1894).  TFS_ERR_NON_FINITE if a block is singular. */
int tfs_broyden_set_jacobian(tfs_broyden *b, const double *blocks);
/* Free bodies (swimmer::jacobian_inv_mult, swimmer.cpp:1012-1066): after
 * tfs_broyden_set_jacobian, couples each swimmer's filament blocks J_i with
 * its 6 x 6 body block through the body's Schur complement, so that J0^{-1} is
 * the inverse of [[diag J_i, A], [B, S0]].  The blocks must be, per swimmer,
 * fil_per_swim filament blocks of one size and then one block of 6.  A, B, S0
 * as tfs_filament_body_coupling writes them (desc's memory space).  A later
 * tfs_broyden_set_jacobian or _inverse removes the coupling. */
int tfs_broyden_set_body_coupling(tfs_broyden *b, int32_t num_swim, int32_t fil_per_swim, const double *A,
                                  const double *B, const double *S0);
/* Free bodies in the Broyden-for-everything form (swimmer::jacobian_inv_mult,
 * swimmer.cpp:931-1010): the blocks are, per swimmer, fil_per_swim filament
 * blocks, one of 6 and blobs_per_swim blob blocks of 3 (their J0 blocks
 * -I / (6 pi mu a), inverse blob_mob_fac I); J0^{-1} eliminates the blob rows
 * with that diagonal mobility and couples the body through the Schur matrix
 * S0 - sum B J^{-1} A, where S0 carries the blob terms (as
 * tfs_filament_body_coupling_blobs writes it).  blob_disp [3 per blob]: the
 * blobs' displacements Q_init ref; blob_mob_fac = -6 pi mu a; dt_fac = 1/DT
 * (backward Euler) or 1.5/DT (BDF2). */
int tfs_broyden_set_body_coupling_blobs(tfs_broyden *b, int32_t num_swim, int32_t fil_per_swim,
                                        int32_t blobs_per_swim, const double *A, const double *B,
                                        const double *S0, const double *blob_disp, double blob_mob_fac,
                                        double dt_fac);
/* Start a timestep: iter = 0 and error = the residual of the initial guess
 * (compute_errors before the loop, cilia_sim_main.cu:394-411).  error == NULL
 * keeps the residual that tfs_broyden_compute_errors left in slot 0. */
int tfs_broyden_start(tfs_broyden *b, const double *error);
int tfs_broyden_find_update(tfs_broyden *b, double *update);
/* new_error == NULL takes the residual that tfs_broyden_compute_errors left in slot 1. */
int tfs_broyden_end_of_iter(tfs_broyden *b, const double *new_error, int32_t error_is_too_large);
/* Current iteration count (broyden_solver::iter) and ||error||, the engine's
 * error (slot 0): like broyden_solver::error it is not replaced when
 * end_of_iter finds the new error small enough or the memory full (that
 * residual stays in slot 1). */
int tfs_broyden_status(tfs_broyden *b, int32_t *iter, double *error_norm);

/* State of the implicit step for mobility_solver::compute_errors
 * (mobility_solver.cpp:2661-2946; free bodies only outside the
 * Broyden-for-everything form).  Pointers live in the desc's memory space; 3 values per segment or
 * body unless stated. */
typedef struct tfs_broyden_state {
    int32_t num_swim, num_fil, num_seg;  /* NSWIM, NFIL, NSEG: n = 6*(NSWIM*NFIL*NSEG + NSWIM) */
    int32_t bdf2;                        /* nt >= NUM_EULER_STEPS (config.hpp:326)             */
    double dt;                           /* DT                                                  */
    double dl;                           /* DL: the position tolerance is DL*TOL                */
    double tol;                          /* TOL (config.hpp:302)                                */
    const double *seg_x, *seg_xm1, *seg_xm2;   /* segment positions at n+1, n, n-1            */
    const double *seg_u, *seg_um1;             /* segment Lie-algebra elements at n+1, n       */
    const double *v_segs;                      /* 6 per segment: velocity, angular velocity   */
    const double *body_x, *body_xm1, *body_xm2, *body_u, *body_um1;
    const double *v_bodies;                    /* 6 per body                                  */
    /* blobs_per_swim > 0: the Broyden-for-everything form (USE_BROYDEN_FOR_EVERYTHING,
     * bodies of prescribed velocity, mobility_solver.cpp:2674-2684, 2774-2850): per
     * swimmer its 3 NBLOB rows V + Omega x (x_blob - x_body) - v_blob follow the body's
     * rows (tolerance TOL) and n = 6*(NSWIM*NFIL*NSEG + NSWIM) + 3*NSWIM*NBLOB.  0: the
     * form above. */
    int32_t blobs_per_swim;
    const double *x_blobs, *v_blobs;           /* 3 per blob                                  */
    /* free_bodies (with blobs_per_swim > 0): free bodies in the Broyden-for-everything
     * form (PRESCRIBED_BODY_VELOCITIES false, mobility_solver.cpp:2676-2681, 2786-2847):
     * the body rows are its force and torque balance -- the segment forces f_segs and
     * their torques about body_x (swimmer::forces_and_torques, swimmer.cpp:469-499) plus
     * the blob forces f_blobs and theirs, tolerance TOL -- and the blob rows take the
     * velocities the body's positions imply (backward Euler, or BDF2 with dexp);
     * v_bodies is then unused. */
    int32_t free_bodies;
    const double *f_segs;                      /* 6 per segment                               */
    const double *f_blobs;                     /* 3 per blob                                  */
} tfs_broyden_state;

/* Residuals of the BDF2 (or backward-Euler) equations for positions and
 * Lie-algebra elements (dexpinv, quaternion.cpp:610-633), written into the
 * engine: slot 0 = error (before tfs_broyden_start), slot 1 = new_error (before
 * tfs_broyden_end_of_iter).  *error_is_too_large is the reference's test
 * (|e| > DL*TOL for positions, > TOL for Lie-algebra elements).  Returns
 * TFS_ERR_NON_FINITE if a residual is not finite (cilia_sim_main.cu:469-487). */
int tfs_broyden_compute_errors(tfs_broyden *b, const tfs_broyden_state *st, int32_t slot,
                               int32_t *error_is_too_large);
/* Copy of slot 0 or 1 into `out` (desc's memory space). */
int tfs_broyden_read_error(tfs_broyden *b, int32_t slot, double *out);

#ifdef __cplusplus
}
#endif

#endif /* TFS_BROYDEN_H */
