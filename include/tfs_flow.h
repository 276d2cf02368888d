/*
 * tfs_flow.h -- C-ABI of the flow-field evaluator on the MI355X.
 *
 * Replaces flow_field_evaluator (src/flow_field/flow_field_evaluator.hpp:8-63):
 *   setup / read_positions_and_forces  (flow_field_evaluator.cu:265-544)
 *                                       -> tfs_flow_create, tfs_flow_set_sources
 *   velocities + flow_velocities kernel (flow_field_evaluator.cu:13-228, 546-579)
 *                                       -> tfs_flow_evaluate
 * The tracer points are the caller's (the reference hard-codes a 551-wide
 * plane at x = 0, flow_field_evaluator.cu:387-389).  Layouts are the
 * reference's: x_segs [3*Ns], f_segs [6*Ns] (force, torque), x_blobs / f_blobs
 * [3*Nb], x_bodies [3*NBOD], points and velocities [3*Nf], bodies stored
 * contiguously (Ns = NBOD*NFIL*NSEG, Nb = NBOD*NBLOB).
 */
#ifndef TFS_FLOW_H
#define TFS_FLOW_H

#include <stddef.h>
#include <stdint.h>

#include "tfs_rpy.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TFS_FLOW_STOKESLET_ROTLET 0 /* the reference evaluator: (F/r + (r.F) r/r^3 + tau x r/r^3)/(8 pi mu),
                                        Blakelet images when there are no blobs                    */
#define TFS_FLOW_RPY 1              /* RPY velocity of a point tracer (radius 0)                     */

typedef struct tfs_flow_desc {
    int32_t num_bodies;       /* NBOD                                                  */
    int32_t segs_per_body;    /* NFIL*NSEG                                             */
    int32_t blobs_per_body;   /* NBLOB (0: half-space with Blakelet images)            */
    int32_t mode;             /* TFS_FLOW_*                                            */
    int32_t precision;        /* TFS_PRECISION_*                                       */
    int32_t remove_net_force; /* REMOVE_NET_FORCE_STOKESLET (Stokeslet mode only)       */
    double seg_radius, blob_radius, mu;
    int32_t memory;           /* TFS_MEMORY_* of every pointer passed later            */
    void *stream;             /* hipStream_t or NULL                                    */
} tfs_flow_desc;

typedef struct tfs_flow tfs_flow;

int tfs_flow_create(const tfs_flow_desc *desc, tfs_flow **out);
int tfs_flow_destroy(tfs_flow *h);
/* Source state of one save (read_positions_and_forces). x_bodies may be NULL
 * unless remove_net_force is set. */
int tfs_flow_set_sources(tfs_flow *h, const double *x_segs, const double *f_segs,
                         const double *x_blobs, const double *f_blobs, const double *x_bodies);
/* Velocities at num_points tracer points (velocities()).  Asynchronous for
 * device memory; with host memory it returns when v_points is written. */
int tfs_flow_evaluate(tfs_flow *h, int64_t num_points, const double *x_points, double *v_points);
/* The same with float arrays (flow_field_evaluator of a Real = float build,
 * flow_field_evaluator.hpp:8-63): widened to FP64 on the device, velocities
 * rounded once to float. */
int tfs_flow_set_sources_f32(tfs_flow *h, const float *x_segs, const float *f_segs,
                             const float *x_blobs, const float *f_blobs, const float *x_bodies);
int tfs_flow_evaluate_f32(tfs_flow *h, int64_t num_points, const float *x_points, float *v_points);
int tfs_flow_sync(tfs_flow *h);
/* Threads per workgroup (64, 128 or 256) of the evaluation kernel. */
int tfs_flow_set_block(tfs_flow *h, int32_t block);
/* Milliseconds of the last tfs_flow_evaluate's kernel (synchronises). */
int tfs_flow_last_kernel_ms(tfs_flow *h, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* TFS_FLOW_H */
