// tfs_api_internal.h -- the handle behind the C-ABI (not part of the C-ABI).
#pragma once
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tfs_rpy.h"
#include "../../include/tfs_solver.h"
#include "../../include/tfs_broyden.h"
#include "tfs_internal.h"

namespace tfs {

// Device staging of FP32 caller arrays (the reference's Real = float build):
// floats cross PCIe and are widened to FP64 on the device, or narrowed from
// it, so every computation behind the boundary stays FP64.
struct F32Stage {
    float* p = nullptr;
    size_t cap = 0;
};
void free_stage(F32Stage& st);
// dst[0:n] (device, FP64) = src[0:n] (FP32 in host or device memory), exact.
int widen_in(F32Stage& st, const float* src, double* dst, int64_t n, bool host, hipStream_t s, const char* where);
// dst[0:n] (FP32 in host or device memory) = src[0:n] (device, FP64) rounded
// to nearest.  Asynchronous on s like the FP64 copies.
int narrow_out(F32Stage& st, const double* src, float* dst, int64_t n, bool host, hipStream_t s, const char* where);

// A cached work list for one block mask and launch shape.
struct Plan {
    std::vector<MatvecTile> tiles;
    std::vector<MatvecItem> items;
    MatvecTile* d_tiles = nullptr;
    MatvecItem* d_items = nullptr;
    uint32_t blocks = 0;  // blocks actually evaluated (after the reference's overwrite rule)
    bool symmetric = false;  // items are pairs of tile groups (the symmetric form)
};

}  // namespace tfs


struct tfs_rpy {
    tfs_rpy_desc desc{};
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double* arena = nullptr;
    // Reference-layout (AoS) device arrays, carved out of `arena`.
    double* x_segs = nullptr;
    double* f_segs = nullptr;
    double* x_blobs = nullptr;
    double* f_blobs = nullptr;
    double* v_segs = nullptr;
    double* v_blobs = nullptr;
    // Partial sums of the work items (grown on demand).
    double* partial = nullptr;
    size_t partial_len = 0;
    tfs::TileShape shape[2];
    // Independent bodies (one by default) and, per precision, the overlapping
    // pairs of the current positions.
    std::vector<tfs::BodyRange> bodies;
    tfs::NearList near[2];
    bool near_valid[2] = {false, false};
    // Rebuilt near lists whose capacity check is left to the caller: the
    // largest count is copied to host_int[8 + precision] without a wait, and
    // tfs::rpy_near_check reads it after the caller's own synchronisation.
    bool defer_near_check = false;
    bool body_table_valid = false;  // d_body_range holds `bodies`
    bool near_pending[2] = {false, false};
    tfs::NearGrid grid;
    int domain = 0;  // TFS_DOMAIN_UNBOUNDED or TFS_DOMAIN_HALF_SPACE
    // Fixed-point positions for FP32 tiles (valid with the near lists):
    // per-body centres c_b (bounding-box centres) and one lambda = 2^k.
    int4* x_fx = nullptr;
    double fx_lambda = 1.0;
    bool fx_valid = false;
    double* d_body_box = nullptr;    // 6 per body: min xyz, max xyz
    int32_t* d_body_range = nullptr; // 4 per body (BodyRange), for the box and conversion kernels
    size_t body_cap = 0;
    std::vector<double> fx_centre;   // 3 per body
    int32_t* host_int = nullptr;  // pinned scratch
    double* blob_rep = nullptr;   // device blob repulsion staging (tfs_rpy_barrier_forces, host memory)
    // Persistent launches with timing: per-item stamps of the last launch and
    // the plan they belong to.
    int64_t* d_item_stamps = nullptr;
    size_t item_stamps_len = 0;
    const tfs::Plan* stamped_plan = nullptr;
    bool body_timing = false;
    // Tile shape set by tfs_rpy_set_tile_shape or tfs_rpy_autotune (else the
    // default follows the body list).
    bool shape_chosen[2] = {false, false};
    // Stop flag passed to the kernels of later evaluations (set by a solver
    // around its speculative Krylov iterations; null otherwise).
    const int32_t* skip_flag = nullptr;
    // Per-body stop flags passed to the kernels (MatvecArgs::body_skip), set
    // by a solver around the lock-step iterations of independent bodies.
    const int32_t* body_skip = nullptr;
    // Live kernel timing: event pairs recorded around each pair-kernel launch.
    bool profiling = false;
    struct EvPair {
        hipEvent_t begin, end;
        int precision;
    };
    std::vector<EvPair> ev_used, ev_free;
    std::map<std::tuple<uint32_t, int, int, int, int, int>, tfs::Plan> plans;
    tfs::F32Stage f32;  // FP32 entry points (tfs_rpy_*_f32)
};

struct tfs_velocities;
struct tfs_broyden;

namespace tfs {

// The blob-motion mode of a velocities handle (tfs_velocities_set_blob_motion).
int32_t velocities_blob_motion(const tfs_velocities* v);
// tfs_velocities_compute(_free) whose device-memory results are ordered on
// the handle's stream instead of complete on return.
int velocities_compute_ordered(tfs_velocities* v, const double* x_segs, const double* f_segs, const double* x_blobs,
                               const double* x_bodies, const double* bodies, const double* q_bodies, double* v_segs,
                               double* f_blobs, double* v_bodies_io, tfs_solve_stats* stats, bool free_bodies);
hipStream_t velocities_stream(const tfs_velocities* v);
tfs_rpy* velocities_rpy(const tfs_velocities* v);
// A device-memory handle whose next computes enqueue fixed launches only (the
// dense direct blob solve formed, the blobs' motion declared): capturable.
bool velocities_graph_ready(const tfs_velocities* v);
// Deferred near-list checks (tfs_rpy::defer_near_check): after `waited` has
// been synchronised, reads the largest neighbour count of every list rebuilt
// since the last check; a list that overflowed its capacity is dropped with
// the capacity grown, and *overflow is set: the evaluations since the rebuild
// must be repeated.
int rpy_near_check(tfs_rpy* h, hipStream_t waited, bool* overflow);
// Norm of the latest residual given to a Broyden engine: the new_error of its
// last end_of_iter (slot 0 after an update, slot 1 otherwise), else the
// error it started from.  Synchronises.
int broyden_latest_norm(tfs_broyden* b, double* out);
hipStream_t broyden_stream(const tfs_broyden* b);
// Device-counted Broyden iterations, for hipGraph capture (device memory, at
// most 65,536 unknowns): begin copies the iteration count to the device;
// find_update and end_of_iter (residual in slot 1, its too-large flag read on
// the device) then take it from there, with the update committed by copies
// instead of slot swaps; finish hands the count and the latest slot back.
bool broyden_dev_supported(const tfs_broyden* b);
int broyden_dev_begin(tfs_broyden* b);
int broyden_dev_find_update(tfs_broyden* b, double* update);
// pending (device, may be null): the end of iteration runs only while it is
// nonzero, and clears it.
int broyden_dev_end_of_iter(tfs_broyden* b, int32_t* pending);
void broyden_dev_finish(tfs_broyden* b, int32_t iter, int32_t latest);
// tfs_broyden_compute_errors in two halves: the launches and (copy_flags)
// the copy of the flags to the host (no wait), and the flags once the stream
// has been waited for; broyden_copy_error_flags enqueues that copy alone.
int broyden_errors_enqueue(tfs_broyden* b, const tfs_broyden_state* st, int32_t slot, bool copy_flags = true);
int broyden_copy_error_flags(tfs_broyden* b);
int broyden_errors_result(tfs_broyden* b, int32_t* too_large);
// Device address of the residual's two flags: its too-large measure (> 1:
// too large) and nonzero if it was not finite.
const double* broyden_error_flags(const tfs_broyden* b);
// While flags is set, every residual's last kernel also counts one more
// iteration in flags[0] unless flags[2] (stopped) is set, marks its end of
// iteration pending in flags[1], and sets flags[2] to 1 once the residual is
// small enough or max_iter iterations have run, 2 if it is not finite.
void broyden_set_bookkeeping(tfs_broyden* b, int32_t* flags, int max_iter);
// The too-large measure of the residual whose flags were last read (> 1: too large).
double broyden_last_measure(const tfs_broyden* b);
// Grows the near-pair lists of `precision` to at least `cap` entries per
// particle (rebuilt at the next evaluation).
int rpy_near_reserve(tfs_rpy* h, int precision, int cap);

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* where);
// TFS_OK for hipSuccess, else hip_fail(e, where).
int hip_fail_if(hipError_t e, const char* where);
TileShape default_shape(int precision, int ns, int nb, int num_cus, int seg_dof, int num_bodies);
Plan* get_plan(tfs_rpy* h, uint32_t blocks, const TileShape& sh, int* err);
// Whether the symmetric form can evaluate `blocks` of this handle in
// `precision` (3-DOF segments, unbounded domain, the SS, SB and BS blocks all
// active).
bool symmetric_applies(const tfs_rpy* h, int precision, uint32_t blocks);
// The launch shape of one evaluation: the handle's shape, or its default
// source-split shape where a symmetric shape does not apply.
TileShape effective_shape(const tfs_rpy* h, int precision, uint32_t blocks);
// Positions changed: the near-pair lists must be rebuilt before the next matvec.
void invalidate_near(tfs_rpy* h);
// Builds the near-pair list of `precision` (and, for FP32, the fixed-point
// positions) if stale; allocates and synchronises, so never under capture.
int ensure_near(tfs_rpy* h, int precision);

// Matvec with explicit views: forces fs (seg_dof comps) / fb (3 comps) and
// outputs vs (seg_out_slots comps) / vb (3 comps), with the write rules of
// `blocks` and the handle's positions.  `sad` (may be null) folds the
// saddle-point couplings into the reduce kernel.
int run_matvec(tfs_rpy* h, uint32_t blocks, int precision, CView fs, CView fb, View vs, View vb,
               int seg_out_slots, const SaddleEpilogue* sad = nullptr);

}  // namespace tfs
