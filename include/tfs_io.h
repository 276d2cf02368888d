/*
 * tfs_io.h -- C-ABI of the simulation data files: the reference's writers and
 * readers, and its file-driven flow-field run.
 *
 * The reference writes one record per line with std::ofstream, every value
 * followed by one space:
 *
 *   <name>.par                  "value %% NAME" lines                   cilia_sim_main.cu:168-210
 *   <name>_seg_states.dat       step, then per segment x y z (prescribed
 *                               cilia) or qw qx qy qz                    cilia_sim_main.cu:584-606,
 *                                                                        filament.cpp:1932-1950
 *   <name>_body_states.dat      step, then x y z qw qx qy qz per body    rigid_body.cpp:550-554
 *   <name>_seg_forces.dat       step, then f tau (6) per segment         mobility_solver.cpp:3076-3153
 *   <name>_seg_vels.dat         step, then v w (6) per segment           mobility_solver.cpp:3072-3149
 *   <name>_blob_forces.dat      step, then 3 per blob                    mobility_solver.cpp:3068-3136
 *   <name>_body_vels.dat        6 per body, no step                      mobility_solver.cpp:3064-3130
 *   <name>_fil_references.dat   3*NFIL, default format, no step          swimmer.cpp:1110-1125
 *   <name>_blob_references.dat  3*NBLOB, default format, no step         rigid_body.cpp:531-548
 *   <name>.backup               step, then per body its filament records
 *                               and its body record; overwritten         cilia_sim_main.cu:613-624,
 *                                                                        filament.cpp:1952-2020,
 *                                                                        rigid_body.cpp:556-578
 *   <name>_filament_phases.dat  step, then the phases                    cilia_sim_main.cu:626-640
 *   <name>_true_states.dat      step, PERIOD, phases (and angles)        cilia_sim_main.cu:632-668
 *
 * Data values are written std::scientific with precision 10 (as "%.10e"),
 * steps and integer parameters as integers, and the reference-position files
 * and .par reals in the stream's default format (as "%g").
 *
 * All of this is host code: only tfs_io_flow_run uses the GPU (through
 * tfs_flow.h).  Entry points return TFS_OK or a TFS_ERR_* code of tfs_rpy.h
 * (TFS_ERR_IO, TFS_ERR_FORMAT for file problems), with the message in
 * tfs_last_error().
 */
#ifndef TFS_IO_H
#define TFS_IO_H

#include <stddef.h>
#include <stdint.h>

#include "tfs_rpy.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status of tfs_io_read_record at the end of a file (not an error). */
#define TFS_IO_END 100

/* tfs_io_write_record flags. */
#define TFS_IO_APPEND 1        /* append to the file (std::ios::app); else overwrite     */
#define TFS_IO_WITH_STEP 2     /* the record starts with the integer step and a space    */
#define TFS_IO_DEFAULT_FLOAT 4 /* values in the default format ("%g"), else "%.10e"      */

/* One "value %% NAME" line of a .par file. */
typedef struct tfs_io_param {
    char name[64];
    double value;
    int32_t is_integer; /* written (and read back) as an integer token */
} tfs_io_param;

/* Writes a .par file: the parameters in order, one per line
 * (cilia_sim_main.cu:168-210).  Overwrites the file. */
int tfs_io_write_par(const char *path, int32_t count, const tfs_io_param *params);
/* Reads every "value %% NAME" line of a .par file (up to capacity of them);
 * *count is the number of lines in the file. */
int tfs_io_read_par(const char *path, int32_t capacity, tfs_io_param *params, int32_t *count);

/* Writes one record: "<step> " if TFS_IO_WITH_STEP, then each value and a
 * space, then a newline. */
int tfs_io_write_record(const char *path, int32_t flags, int64_t step, int64_t count, const double *values);

/* Sequential reader of a record file. */
typedef struct tfs_io_reader tfs_io_reader;
int tfs_io_open(const char *path, tfs_io_reader **out);
int tfs_io_close(tfs_io_reader *r);
/* Reads the next record: its integer step if with_step (into *step, which may
 * be NULL) and then exactly count values.  TFS_ERR_FORMAT if the record holds
 * another number of values or a token that is not a number; TFS_IO_END when no
 * record is left (blank lines are skipped). */
int tfs_io_read_record(tfs_io_reader *r, int32_t with_step, int64_t count, int64_t *step, double *values);
/* Number of numbers (step included) in the next record, without consuming
 * it; TFS_IO_END when no record is left. */
int tfs_io_peek_count(tfs_io_reader *r, int64_t *count);
/* Back to the first record (the flow-field run's periodic cycles,
 * flow_field_main.cpp:190-206). */
int tfs_io_rewind(tfs_io_reader *r);
/* Number of records (non-blank lines) in a file. */
int tfs_io_count_records(const char *path, int64_t *count);

/* Source positions of one save from the state records
 * (flow_field_evaluator.cu:401-498): per body b with state X q
 * (body_states[7b..7b+6] = x y z qw qx qy qz) and R = R(q) the unit-quaternion
 * rotation (quaternion.cpp:327-370),
 *   x_bodies[b] = X,  x_blobs[b, k] = X + R blob_refs[k],
 *   filament f base = X + R fil_refs[f];
 * with quaternion_based, segment 0 sits at its base and each next segment at
 *   x[s] = x[s-1] + dl/2 (t[s-1] + t[s]),  t = first column of R(q_seg)
 * (quaternion.cpp:285-291); otherwise seg_states already holds positions and
 * is copied.  The reference's reader uses dl = 2.2 RSEG.  x_blobs and blob_refs
 * may be NULL when num_blobs is 0; x_bodies may be NULL. */
int tfs_io_sources_from_states(int32_t num_bodies, int32_t num_fil, int32_t num_seg, int32_t num_blobs,
                               double dl, int32_t quaternion_based, const double *body_states,
                               const double *seg_states, const double *fil_refs, const double *blob_refs,
                               double *x_segs, double *x_blobs, double *x_bodies);

/* The reference's flow-field program (flow_field_main.cpp:8-220) on the GPU.
 * input_path names its parameter file ("flow_field_simulation_parameters.input":
 * SIMULATION_NAME, NFLOW, QUATERNION_BASED_DATA, TREAT_DATA_AS_PERIODIC,
 * NUMBER_OF_CYCLES_FOR_PERIODIC_DATA, REMOVE_NET_FORCE_STOKESLET,
 * ADVANCE_TRACER_POSITIONS as "KEY value" pairs); SIMULATION_NAME is taken
 * relative to that file's folder.  For each save it reads the state and force
 * records, evaluates the Stokeslet/rotlet flow (tfs_flow.h) at the tracers in
 * the given precision, moves the tracers by DT*PLOT_FREQUENCY_IN_STEPS times
 * the last velocities if asked (flow_field_evaluator.cu:511-522), and appends
 * the tracer positions and velocities to <name>_flow_locations.dat and
 * <name>_flow_velocities.dat.  x_points (3*NFLOW) gives the first tracer
 * positions; NULL takes the reference's seeding, a 551-wide grid of spacing 8
 * in the plane x = 0 (flow_field_evaluator.cu:338-398).  A cycle covers the
 * saves present (at most 1 + (TOTAL_TIME_STEPS-1)/PLOT_FREQUENCY_IN_STEPS,
 * ending early at the end of the shortest file); *saves (may be NULL) is the
 * number evaluated, cycles included. */
int tfs_io_flow_run(const char *input_path, int32_t precision, const double *x_points, int64_t *saves);

#ifdef __cplusplus
}
#endif

#endif /* TFS_IO_H */
