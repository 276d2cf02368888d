/*
 * tfs_real.hpp -- C++ overloads of the C-ABI's array entry points on the
 * caller's element type.
 *
 * The reference fixes Real at compile time (config.hpp:36-67: float as
 * shipped, double with FIL_USE_DOUBLE_PRECISION).  A back-end class written
 * against these overloads (tfs_real::upload(h, what, x_segs_host), ...)
 * compiles for either: double arrays go to the FP64 entry points, float
 * arrays to their *_f32 twins.  Header-only forwarding; the library's exported
 * surface is the C headers it includes.
 */
#ifndef TFS_REAL_HPP
#define TFS_REAL_HPP

#include "tfs_flow.h"
#include "tfs_rpy.h"
#include "tfs_solver.h"

namespace tfs_real {

// copy_*_to_device / copy_*_velocities_to_host (rpy_mobility_solver.cu:119-205)
inline int upload(tfs_rpy *h, int32_t what, const double *src) { return tfs_rpy_upload(h, what, src); }
inline int upload(tfs_rpy *h, int32_t what, const float *src) { return tfs_rpy_upload_f32(h, what, src); }
inline int download(tfs_rpy *h, int32_t what, double *dst) { return tfs_rpy_download(h, what, dst); }
inline int download(tfs_rpy *h, int32_t what, float *dst) { return tfs_rpy_download_f32(h, what, dst); }
inline int apply(tfs_rpy *h, uint32_t blocks, const double *fs, const double *fb, double *vs, double *vb) {
    return tfs_rpy_apply(h, blocks, fs, fb, vs, vb);
}
inline int apply(tfs_rpy *h, uint32_t blocks, const float *fs, const float *fb, float *vs, float *vb) {
    return tfs_rpy_apply_f32(h, blocks, fs, fb, vs, vb);
}
// apply_interparticle_forces (rpy_mobility_solver.cu:207-232)
inline int barrier_forces(tfs_rpy *h, const tfs_barrier_desc *d, double *fs, double *fb_rep) {
    return tfs_rpy_barrier_forces(h, d, fs, fb_rep);
}
inline int barrier_forces(tfs_rpy *h, const tfs_barrier_desc *d, float *fs, float *fb_rep) {
    return tfs_rpy_barrier_forces_f32(h, d, fs, fb_rep);
}

// the solver layer of mobility_solver (mobility_solver.cpp:1806-2570)
inline int set_geometry(tfs_solver *s, const double *xs, const double *xb, const double *kp, const double *kt,
                        int32_t memory) {
    return tfs_solver_set_geometry(s, xs, xb, kp, kt, memory);
}
inline int set_geometry(tfs_solver *s, const float *xs, const float *xb, const float *kp, const float *kt,
                        int32_t memory) {
    return tfs_solver_set_geometry_f32(s, xs, xb, kp, kt, memory);
}
inline int set_bodies(tfs_solver *s, const double *x, int32_t memory) { return tfs_solver_set_bodies(s, x, memory); }
inline int set_bodies(tfs_solver *s, const float *x, int32_t memory) { return tfs_solver_set_bodies_f32(s, x, memory); }
inline int system_mult(tfs_solver *s, const double *in, double *out, int32_t precision, int32_t memory) {
    return tfs_solver_system_mult(s, in, out, precision, memory);
}
inline int system_mult(tfs_solver *s, const float *in, float *out, int32_t precision, int32_t memory) {
    return tfs_solver_system_mult_f32(s, in, out, precision, memory);
}
inline int precondition(tfs_solver *s, const double *in, double *out, int32_t memory) {
    return tfs_solver_precondition(s, in, out, memory);
}
inline int precondition(tfs_solver *s, const float *in, float *out, int32_t memory) {
    return tfs_solver_precondition_f32(s, in, out, memory);
}
inline int solve(tfs_solver *s, const double *rhs, double *x, int32_t memory, tfs_solve_stats *st) {
    return tfs_solver_solve(s, rhs, x, memory, st);
}
inline int solve(tfs_solver *s, const float *rhs, float *x, int32_t memory, tfs_solve_stats *st) {
    return tfs_solver_solve_f32(s, rhs, x, memory, st);
}
inline int residual(tfs_solver *s, const double *rhs, const double *x, int32_t precision, int32_t memory,
                    double *relres) {
    return tfs_solver_residual(s, rhs, x, precision, memory, relres);
}
inline int residual(tfs_solver *s, const float *rhs, const float *x, int32_t precision, int32_t memory,
                    double *relres) {
    return tfs_solver_residual_f32(s, rhs, x, precision, memory, relres);
}

// compute_velocities of elastic filaments (mobility_solver.cpp:2598-2655)
inline int compute(tfs_velocities *v, const double *xs, const double *fs, const double *xb, const double *xbod,
                   const double *vbod, double *vs, double *fb, tfs_solve_stats *st) {
    return tfs_velocities_compute(v, xs, fs, xb, xbod, vbod, vs, fb, st);
}
inline int compute(tfs_velocities *v, const float *xs, const float *fs, const float *xb, const float *xbod,
                   const float *vbod, float *vs, float *fb, tfs_solve_stats *st) {
    return tfs_velocities_compute_f32(v, xs, fs, xb, xbod, vbod, vs, fb, st);
}
inline int compute_free(tfs_velocities *v, const double *xs, const double *fs, const double *xb, const double *xbod,
                        const double *fbod, const double *qbod, double *vs, double *fb, double *vbod,
                        tfs_solve_stats *st) {
    return tfs_velocities_compute_free(v, xs, fs, xb, xbod, fbod, qbod, vs, fb, vbod, st);
}
inline int compute_free(tfs_velocities *v, const float *xs, const float *fs, const float *xb, const float *xbod,
                        const float *fbod, const float *qbod, float *vs, float *fb, float *vbod, tfs_solve_stats *st) {
    return tfs_velocities_compute_free_f32(v, xs, fs, xb, xbod, fbod, qbod, vs, fb, vbod, st);
}

// flow_field_evaluator (flow_field_evaluator.hpp:8-63)
inline int set_sources(tfs_flow *h, const double *xs, const double *fs, const double *xb, const double *fb,
                       const double *xbod) {
    return tfs_flow_set_sources(h, xs, fs, xb, fb, xbod);
}
inline int set_sources(tfs_flow *h, const float *xs, const float *fs, const float *xb, const float *fb,
                       const float *xbod) {
    return tfs_flow_set_sources_f32(h, xs, fs, xb, fb, xbod);
}
inline int evaluate(tfs_flow *h, int64_t n, const double *x, double *v) { return tfs_flow_evaluate(h, n, x, v); }
inline int evaluate(tfs_flow *h, int64_t n, const float *x, float *v) { return tfs_flow_evaluate_f32(h, n, x, v); }

}  // namespace tfs_real

#endif /* TFS_REAL_HPP */
