/*
 * tfs_rpy.h -- C-ABI of the MI355X (gfx950) RPY mobility library.
 *
 * Drop-in for the RPY mobility back-end of TimFilSim: the abstract class
 * `mobility_solver` (src/cilia/mobility/mobility_solver.hpp:19-127) and its
 * GPU implementation `rpy_mobility_solver`
 * (src/cilia/mobility/rpy_mobility_solver.cu:1-337).  Plain pointers and
 * sizes only; no C++ or torch types cross this boundary.
 *
 * Layouts are the reference's host layouts (rpy_mobility_solver.cu:65-71):
 *   x_segs [3*Ns]  x y z       per segment, body -> filament -> segment
 *   f_segs [6*Ns]  fx fy fz tx ty tz
 *   v_segs [6*Ns]  vx vy vz wx wy wz   (w written as 0 when seg_dof == 3)
 *   x_blobs, f_blobs, v_blobs [3*Nb]
 * where Ns = NSWIM*NFIL*NSEG and Nb = NSWIM*NBLOB.
 *
 * Every entry point returns 0 (TFS_OK) or a TFS_ERR_* code; tfs_last_error()
 * gives the message of the last failure on the calling thread.  Work is
 * enqueued on the handle's stream; results are visible after tfs_rpy_sync()
 * (the reference's wait_for_device, rpy_mobility_solver.cu:247-256).  A handle
 * is not thread-safe: use one handle per host thread.
 */
#ifndef TFS_RPY_H
#define TFS_RPY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFS_OK 0
#define TFS_ERR_INVALID_ARGUMENT 1
#define TFS_ERR_HIP 2
#define TFS_ERR_NON_FINITE 3
#define TFS_ERR_NOT_CONVERGED 4
#define TFS_ERR_NO_DEVICE 5
#define TFS_ERR_OUT_OF_MEMORY 6
#define TFS_ERR_IO 7     /* a data file cannot be opened, read or written (tfs_io.h) */
#define TFS_ERR_FORMAT 8 /* a data file record does not parse or has the wrong length */

/* Evaluation precision of the pair tensors. */
#define TFS_PRECISION_FP64 0       /* FP64 tiles, FP64 sums (reference FP64 build) */
#define TFS_PRECISION_FP32_ACC64 1 /* FP32 tiles, FP64 accumulation                */

/* Fluid domain of the pair tensor. */
#define TFS_DOMAIN_UNBOUNDED 0     /* cuda_functions.cu:207-354                     */
#define TFS_DOMAIN_HALF_SPACE 1    /* no-slip wall at z = 0, INFINITE_PLANE_WALL,
                                      cuda_functions.cu:12-203                      */

/* Write rule of the blob-segment block (cuda_functions.cu:739-751). */
#define TFS_BS_OVERWRITE 0         /* GMRES-for-linear-system, elastic filaments   */
#define TFS_BS_ACCUMULATE 1        /* PRESCRIBED_CILIA || USE_BROYDEN_FOR_EVERYTHING */

/* Mobility blocks (bit mask). */
#define TFS_BLOCK_SS 1u            /* V_s  = M_ss F_s   Mss_mult cuda_functions.cu:362-474 */
#define TFS_BLOCK_SB 2u            /* V_s += M_sb F_b   Msb_mult cuda_functions.cu:551-636 */
#define TFS_BLOCK_BB 4u            /* V_b  = M_bb F_b   Mbb_mult cuda_functions.cu:476-549 */
#define TFS_BLOCK_BS 8u            /* V_b (+)= M_bs F_s Mbs_mult cuda_functions.cu:638-757 */
#define TFS_BLOCK_ALL 15u          /* evaluate_full_mobility rpy_mobility_solver.cu:330-337 */

/* Where caller pointers live. */
#define TFS_MEMORY_HOST 0
#define TFS_MEMORY_DEVICE 1

/* Arrays moved by tfs_rpy_upload / tfs_rpy_download. */
#define TFS_SEG_POSITIONS 0
#define TFS_SEG_FORCES 1
#define TFS_BLOB_POSITIONS 2
#define TFS_BLOB_FORCES 3
#define TFS_SEG_VELOCITIES 4
#define TFS_BLOB_VELOCITIES 5

typedef struct tfs_rpy_desc {
    int32_t num_segs;      /* NSWIM*NFIL*NSEG (>= 0)                                   */
    int32_t num_blobs;     /* NSWIM*NBLOB (>= 0)                                       */
    int32_t seg_dof;       /* 6: <6,6>,<6,3>,<3,6> (!PRESCRIBED_CILIA); 3: <3,3>        */
    int32_t precision;     /* TFS_PRECISION_*                                          */
    int32_t bs_mode;       /* TFS_BS_*                                                  */
    int32_t memory;        /* TFS_MEMORY_*: where the pointers passed later live       */
    double seg_radius;     /* RSEG  (config.hpp:256)                                   */
    double blob_radius;    /* RBLOB (config.hpp:257)                                   */
    double mu;             /* MU    (config.hpp:254)                                   */
    void *stream;          /* hipStream_t to run on, or NULL for a library-owned one   */
} tfs_rpy_desc;

typedef struct tfs_rpy tfs_rpy;

/* Allocates the device buffers (allocate_device_memory, rpy_mobility_solver.cu:91-117). */
int tfs_rpy_create(const tfs_rpy_desc *desc, tfs_rpy **out);
/* free_device_memory (rpy_mobility_solver.cu:29-53). */
int tfs_rpy_destroy(tfs_rpy *h);

/* copy_{segment,blob}_{positions,forces}_to_device (rpy_mobility_solver.cu:119-161):
 * `what` is TFS_SEG_POSITIONS / TFS_SEG_FORCES / TFS_BLOB_POSITIONS / TFS_BLOB_FORCES,
 * `src` the full AoS array in the desc's memory space. Asynchronous.
 * TFS_SEG_VELOCITIES / TFS_BLOB_VELOCITIES are accepted too: they set the
 * velocities that an accumulating block (SB, or BS in TFS_BS_ACCUMULATE) adds to. */
int tfs_rpy_upload(tfs_rpy *h, int32_t what, const double *src);

/* copy_{segment,blob}_velocities_to_host (rpy_mobility_solver.cu:177-205):
 * `what` is TFS_SEG_VELOCITIES / TFS_BLOB_VELOCITIES. Asynchronous. */
int tfs_rpy_download(tfs_rpy *h, int32_t what, double *dst);

/* evaluate_{segment_segment,segment_blob,blob_blob,blob_segment,full}_mobility
 * (rpy_mobility_solver.cu:258-337) on the uploaded positions/forces, into the
 * handle's velocity buffers, with the reference's write rules: SS '=', SB '+=',
 * BB '=', BS '+=' or '=' per bs_mode; blocks given together are applied in the
 * reference's order (SS then SB; BB then BS). */
int tfs_rpy_evaluate(tfs_rpy *h, uint32_t blocks);

/* One call: V = M F for the requested blocks with caller arrays (any of them may
 * be NULL when its block is absent).  v_* are read first where a block
 * accumulates.  Positions must have been uploaded. */
int tfs_rpy_apply(tfs_rpy *h, uint32_t blocks, const double *f_segs, const double *f_blobs,
                  double *v_segs, double *v_blobs);

/* wait_for_device (rpy_mobility_solver.cu:247-256). */
int tfs_rpy_sync(tfs_rpy *h);

/* Switch the pair-tensor precision of later evaluations. */
int tfs_rpy_set_precision(tfs_rpy *h, int32_t precision);

/* Fluid domain of later evaluations (the reference's compile-time
 * INFINITE_PLANE_WALL, config.hpp): TFS_DOMAIN_UNBOUNDED (default) or
 * TFS_DOMAIN_HALF_SPACE, the Swan-Brady wall tensor of
 * cuda_functions.cu:12-203 for particles above z = 0, which like the
 * reference assumes radius 1 and viscosity 1 (other values are rejected) and
 * has no overlap branch.  FP32 tiles evaluate the wall tensor in float (as
 * the reference's Real = float build does) from exact fixed-point
 * separations, with FP64 sums across source batches.  The self term is
 * applied to the same particle (the reference tests the class-local indices,
 * which would also match segment i with blob i). */
int tfs_rpy_set_domain(tfs_rpy *h, int32_t domain);

/* Override the launch shape (threads per workgroup 64/128/256, targets per
 * thread 1/2, sources per work item: a multiple of block, or a divisor of it
 * of at least 32) for a precision; every lane holds its own targets (no
 * source split). */
int tfs_rpy_set_tile_shape(tfs_rpy *h, int32_t precision, int32_t block, int32_t per_thread,
                           int32_t chunk);

/* Device address of an internal array (TFS_SEG_POSITIONS ... TFS_BLOB_VELOCITIES,
 * AoS reference layout), for callers that keep data resident on the GPU. */
int tfs_rpy_device_buffer(tfs_rpy *h, int32_t what, double **out);

/* Positions were written in place through tfs_rpy_device_buffer: the
 * overlapping-pair lists must be rebuilt before the next evaluation. */
int tfs_rpy_positions_changed(tfs_rpy *h);

/* Independent bodies: the arrays hold num_bodies bodies one after another
 * (segs_per_body[i] segments and blobs_per_body[i] blobs each, summing to the
 * handle's sizes); evaluations couple particles of the same body only, i.e.
 * a block-diagonal grand mobility.  The reference couples all NSWIM bodies
 * (cuda_functions.cu:399, 504, 580, 675); one body is the default.  Unless a
 * tile shape was set or autotuned, work items are then sized so that the
 * largest body alone fills the device, and the bodies' target tiles are
 * interleaved through the launch so that they finish together. */
int tfs_rpy_set_bodies(tfs_rpy *h, int32_t num_bodies, const int32_t *segs_per_body,
                       const int32_t *blobs_per_body);

/* Persistent launches (0 = off): a fixed grid of workgroups_per_cu workgroups
 * per CU instead of one workgroup per item; each takes a contiguous range of
 * the work list and keeps a tile's sums in registers across its items,
 * writing one partial slot per tile it touches.  The source split has no
 * persistent form: a shape set with tfs_rpy_set_split_shape is refused, the
 * default shape gives way to the plain default (block 256, 2 targets per
 * thread). */
int tfs_rpy_set_persistent(tfs_rpy *h, int32_t precision, int32_t workgroups_per_cu);
/* Current launch shape of a precision: block, per_thread, chunk, persistent. */
int tfs_rpy_get_tile_shape(tfs_rpy *h, int32_t precision, int32_t shape[4]);
/* Launch shape with the source split (no counterpart in the reference, one
 * thread per target, cuda_functions.cu:362-757): the block/64 waves of a
 * workgroup (block 256 or 512) share one tile of 64 x per_thread targets and
 * split every LDS batch of the item's `chunk` sources between them, adding
 * their sums in LDS in wave order before the item's one store, so an item
 * writes block/64 times fewer partial bytes.  One workgroup per item (no
 * persistent form).  The default shape of both precisions; tfs_rpy_set_tile_shape
 * selects the form without it. */
int tfs_rpy_set_split_shape(tfs_rpy *h, int32_t precision, int32_t block, int32_t per_thread, int32_t chunk);
/* 1 if the precision's current shape uses the source split. */
int tfs_rpy_get_source_split(tfs_rpy *h, int32_t precision, int32_t *split);
/* Symmetric form (no counterpart in the reference, which evaluates every
 * ordered pair, cuda_functions.cu:362-757): the far-field tensor of a pair is
 * the same seen from either particle, so each unordered pair is evaluated once
 * and its term added to both particles' sums -- one 1/r, A and B for two
 * terms.  64-particle tiles in groups of four per workgroup of 256 threads;
 * results agree with the other shapes to FP32 rounding and are the same on
 * every run.  Independent bodies (tfs_rpy_set_bodies) pair within a body only,
 * and the items of bodies whose stop flag is set return at once.  Both
 * precisions; 3-DOF segments only (refused otherwise); evaluations whose
 * blocks lack SS, SB
 * or BS, and the half-space domain, take the default source-split shape
 * instead.  No persistent form.  The default shape (both precisions) of a
 * handle with 3-DOF segments whose (largest) body's groups of four tiles pair
 * into at least four items per CU (e.g. 12,800 segments + 9,000 blobs: 3,741
 * items); enable = 0 selects the source-split default instead. */
int tfs_rpy_set_symmetric(tfs_rpy *h, int32_t precision, int32_t enable);
/* 1 if the precision's current shape is the symmetric form. */
int tfs_rpy_get_symmetric(tfs_rpy *h, int32_t precision, int32_t *enabled);
/* Per-body completion times of the last launch (either form), in ms from its
 * first work item's start (enable before the launch; reading synchronises). */
int tfs_rpy_body_timing(tfs_rpy *h, int32_t enable);
int tfs_rpy_body_times(tfs_rpy *h, double *ms_per_body);

/* hipGraph of `evaluations` consecutive tfs_rpy_evaluate(h, blocks) calls,
 * replayed by tfs_graph_launch on the handle's stream (no counterpart in the
 * reference, which launches kernels one by one, rpy_mobility_solver.cu:258-337).
 * The work list, partial buffer and near-pair lists are fixed at capture:
 * re-create the graph after positions change. */
typedef struct tfs_graph tfs_graph;
int tfs_rpy_graph_create(tfs_rpy *h, uint32_t blocks, int32_t evaluations, tfs_graph **out);
int tfs_graph_launch(tfs_graph *g);
int tfs_graph_destroy(tfs_graph *g);

/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * THREADS_PER_BLOCK 64, config.hpp:294): times every launch shape (block 64/128/256,
 * 1 or 2 targets per thread, three chunk sizes, queued or one workgroup per
 * item, with and without the source split, and the symmetric form where it
 * applies) on the handle's current data and keeps the fastest for `precision`;
 * *best_ms is its mean time per evaluation. */
int tfs_rpy_autotune(tfs_rpy *h, int32_t precision, uint32_t blocks, int32_t reps, double *best_ms);

/* Live kernel timing (no counterpart in the reference, whose DISPLAYTIME
 * buckets bracket whole phases with cudaDeviceSynchronize, mobility_solver.cpp:2576-2657):
 * while enabled, HIP events are recorded on the handle's stream around every
 * pair-kernel launch.  profile_read synchronises and returns, indexed by
 * precision (TFS_PRECISION_FP64, TFS_PRECISION_FP32_ACC64), the summed kernel
 * milliseconds and launch counts since the last read, then resets them. */
int tfs_rpy_profile(tfs_rpy *h, int32_t enable);
int tfs_rpy_profile_read(tfs_rpy *h, double pair_ms[2], int32_t launches[2]);

/* Pinned host memory, as the reference's cudaHostAlloc(..., cudaHostAllocPortable)
 * host arrays (rpy_mobility_solver.cu:65-71; freed at :14-27). */
int tfs_host_alloc(size_t bytes, void **out);
int tfs_host_free(void *p);

/* Message of the last failure on this thread ("" if none). */
/* Barrier (short-range repulsion) forces of the elastic model,
 * apply_interparticle_forces (rpy_mobility_solver.cu:207-232) and its kernel
 * barrier_forces (cuda_functions.cu:1240-1436), at the handle's current
 * positions.  For each particle p and every other particle j closer than
 * 1.1 (a_p + a_j) -- segments skip their neighbours i-1, i+1 on the same
 * filament --
 *   f_p += strength * fac^4 * (x_p - x_j) / |x_p - x_j|,
 *   fac = min(1, 1 - 10 (|x_p - x_j| - a_p - a_j) / (a_p + a_j));
 * with `wall`, segments below z = dl/2 first get the normal force
 * strength * min(1, 1 - (z - a)/(dl/2 - a))^4 (INFINITE_PLANE_WALL).
 * f_segs [6*Ns]: forces added to slots 0-2 (the reference adds them to the
 * device forces the next matvec uses).  f_blobs_repulsion [3*Nb] (may be
 * NULL): overwritten (f_blobs_repulsion_device).  Pointers in the handle's
 * memory space; complete on return for host memory. */
typedef struct tfs_barrier_desc {
    double strength;      /* REPULSIVE_FORCE_FACTOR * END_FORCE_MAGNITUDE */
    int32_t seg_per_fil;  /* NSEG                                         */
    int32_t wall;         /* 1: plane wall at z = 0                       */
    double dl;            /* DL (wall term only)                          */
} tfs_barrier_desc;

int tfs_rpy_barrier_forces(tfs_rpy *h, const tfs_barrier_desc *desc, double *f_segs, double *f_blobs_repulsion);

/* ---- FP32 host arrays ------------------------------------------------------
 * The reference's host arrays are of its build's Real, float as shipped
 * (config.hpp:36,52; mobility_solver.hpp:57-63).  These take float arrays
 * (host or device memory, as the handle's desc says) with the same layouts
 * and meaning as the double entry points above: floats are widened exactly to
 * FP64 on the device, every evaluation stays FP64 (or FP32 tiles with FP64
 * sums, per the handle's precision), and results are rounded once to float.
 * Asynchronous as their double twins. */
/* copy_*_to_device of a Real = float build (rpy_mobility_solver.cu:119-161). */
int tfs_rpy_upload_f32(tfs_rpy *h, int32_t what, const float *src);
/* copy_*_velocities_to_host of a Real = float build (rpy_mobility_solver.cu:177-205). */
int tfs_rpy_download_f32(tfs_rpy *h, int32_t what, float *dst);
/* tfs_rpy_apply with float arrays. */
int tfs_rpy_apply_f32(tfs_rpy *h, uint32_t blocks, const float *f_segs, const float *f_blobs,
                      float *v_segs, float *v_blobs);
/* apply_interparticle_forces of a Real = float build (rpy_mobility_solver.cu:207-232). */
int tfs_rpy_barrier_forces_f32(tfs_rpy *h, const tfs_barrier_desc *desc, float *f_segs,
                               float *f_blobs_repulsion);

/* The matvec of tfs_rpy_apply without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No handle: `desc` gives the sizes, radii, precision, write rule and stream
 * (memory must be TFS_MEMORY_DEVICE: every pointer is a device pointer);
 * `blocks` and the write rules are those of tfs_rpy_apply (v_* are read
 * first where a block accumulates), for one body in the unbounded domain.
 * The work list, partial sums, overlapping-pair lists, their grid and the
 * FP32 tiles' fixed-point positions live in one static device array of about
 * 22 MiB that each GPU the library runs on holds until the process ends.
 * Limits: at most 32768 particles (segments + blobs) per call, at most 32
 * overlapping neighbours per particle (TFS_ERR_INVALID_ARGUMENT beyond
 * either).  It synchronises its stream once (the neighbour count is read
 * back) and is otherwise asynchronous.  Every call on one GPU shares that
 * array: calls that overlap in time on one GPU (two host threads, or two
 * streams) race and must be serialised by the caller. */
int tfs_rpy_apply_noalloc(const tfs_rpy_desc *desc, uint32_t blocks, const double *x_segs,
                          const double *x_blobs, const double *f_segs, const double *f_blobs,
                          double *v_segs, double *v_blobs);

const char *tfs_last_error(void);

/* Library version string. */
const char *tfs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TFS_RPY_H */
