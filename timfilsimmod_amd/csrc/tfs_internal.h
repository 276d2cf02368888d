// tfs_internal.h -- shared declarations of the HIP library (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace tfs {

// Strided view of an n x c array: element (particle i, component k) lives at
// p[i*ld_p + k*ld_c].  AoS with 6 slots: ld_p = 6, ld_c = 1.  SoA: ld_p = 1,
// ld_c = n.
struct CView {
    const double* p;
    int64_t ld_p;
    int64_t ld_c;
};
struct View {
    double* p;
    int64_t ld_p;
    int64_t ld_c;
};

// One independent body: its segments [seg_begin, seg_end) and blobs
// [blob_begin, blob_end) in the concatenated arrays.  Bodies do not interact.
struct BodyRange {
    int32_t seg_begin, seg_end, blob_begin, blob_end;
};

// One unit of matvec work: the pairs between the targets of `tile` and the
// sources [s_begin, s_end) of class src_cls (0 segments, 1 blobs).  Its
// partial sums go to partial slot = the item's index.
struct MatvecItem {
    int32_t tile;
    int32_t src_cls;
    int32_t s_begin;
    int32_t s_end;
};

// A tile of targets [t_begin, t_end) of class cls; its partial sums live in
// slots [first_item, first_item + n_items).
struct MatvecTile {
    int32_t cls;
    int32_t t_begin;
    int32_t t_end;
    int32_t first_item;
    int32_t n_items;
    int32_t write_mode;  // 0 overwrite, 1 accumulate into the output
    int32_t body;        // index into the body list
    int32_t group0;      // symmetric form: first tile group of the body (slot of group q:
                         // first_item + q - group0)
};

// Overlapping pairs (r^2 <= (ai+aj)^2 as evaluated by the pair kernel of one
// precision): excluded from the branch-free pair kernel and added exactly, in
// FP64 with the reference's overlap formulas, by the reduce kernel.  Target t
// (segments first, then blobs) owns idx[t*cap .. t*cap + cnt[t]); an entry is
// the source index, segments first then blobs.
struct NearList {
    int32_t* cnt = nullptr;
    int32_t* idx = nullptr;
    int32_t* max_cnt = nullptr;
    int cap = 0;
};

// Hashed uniform grid over every particle (segments, then blobs) that
// build_near_list sorts the particles into: buckets of cells of side
// >= max(ai + aj), a power-of-two table.  Device arrays, owned by the handle.
struct NearGrid {
    int32_t* start = nullptr;      // table + 1: first particle of each bucket in `order`
    int32_t* fill = nullptr;       // table: per-bucket counters
    int32_t* bucket_of = nullptr;  // n
    int32_t* order = nullptr;      // n: particles by bucket
    int32_t* body_of = nullptr;    // n: body index of each particle
    int table = 0;
    int n = 0;
    // Small problems (build_near_list's all-pairs form): overlap bit masks,
    // 64 sources per word, and a body index per particle of its own.
    uint64_t* mask = nullptr;
    int64_t mask_len = 0;
    int32_t* pair_body = nullptr;
    int pair_body_len = 0;
    bool may_alloc = false;        // the all-pairs form may allocate its buffers
};

// Barrier forces (cuda_functions.cu:1240-1436): strength
// REPULSIVE_FORCE_FACTOR * END_FORCE_MAGNITUDE, NSEG for the neighbour
// exclusion, and the plane-wall term below z = DL/2.
struct BarrierParams {
    double strength = 0.0;
    int nseg = 1;
    int wall = 0;
    double dl = 0.0;
};

// Launch shape of the pair kernel; chosen by the autotuner.
struct TileShape {
    int block;       // threads per workgroup (64, 128, 256; 512 with split)
    int per_thread;  // targets per thread (1 or 2)
    int chunk;       // sources per work item (a multiple of block)
    int persistent;  // 0: one workgroup per item; >0: that many workgroups per CU, each
                     // taking a contiguous range of the work list
    int split = 0;   // 1: the block/64 waves share one tile of 64 x per_thread targets and
                     // split every LDS batch of sources between them
    int symmetric = 0;  // 1: each pair once for both particles (FP32 tiles, 3-DOF, one
                        // body; 64-particle tiles in groups of four, 256 threads)
};

// Up to this many particles the near-pair lists test every pair (bit masks of
// n^2 / 8 bytes) instead of sorting the particles into a hashed grid.
constexpr int kNearPairsMax = 6144;

// Tiles of 64 particles per group (and waves per workgroup) of the symmetric
// form; values per particle in its partial slots (3 x 64 per slot); the
// largest partial buffer its default shape may take (the slots grow as
// tiles x groups).
constexpr int kSymGroupTiles = 4;
constexpr int kSymSlotDof = 3;
constexpr double kSymMaxPartialBytes = 4.0 * (1 << 30);

// Targets per tile (and per partial slot) of a launch shape.
inline int tile_targets(const TileShape& s) { return s.symmetric ? 64 : (s.split ? 64 : s.block) * s.per_thread; }

// State of a persistent launch: optional per-item start/finish stamps
// (wall_clock64 ticks), two per item.
struct PersistState {
    int64_t* item_stamps = nullptr;
    int num_cus = 256;
};

// Everything the pair kernel needs besides the work list.
struct MatvecArgs {
    CView xs, xb;      // positions (3 components)
    CView fs, fb;      // source strengths (seg_dof or 3 components)
    int ns, nb;
    double a_seg, a_blob, mu;
    // Device stop flag (may be null): when nonzero the pair and reduce kernels
    // return at once (iterations a Krylov solver enqueued past convergence).
    const int32_t* skip = nullptr;
    // FP32 tiles: fixed-point positions X = rint(lambda (x - c_body)), segments
    // then blobs (w unused), with lambda = 2^k and |X| < 2^30.
    const int4* xfx = nullptr;
    double lam = 1.0;
    // Per-body stop flags (may be null): items and tiles of a body whose flag
    // is nonzero do nothing (independent bodies solved in lock-step, whose
    // converged members drop out).
    const int32_t* body_skip = nullptr;
};

// Saddle-point couplings of the prescribed-cilia operator
// (mobility_solver.cpp:1806-2090), applied by the reduce kernel when `in` is
// set: segment rows 3t..3t+2 (filament m = t / nseg) get
// -(phi_dot_m K_phi + theta_dot_m K_theta), and extra workgroups (one wave per
// filament) write out[off_phase + m] = -K_phi^T lambda_m and
// out[off_angle + m] = -K_theta^T lambda_m.  Layout of in/out:
// [lambda_s (3 Ns) | lambda_b (3 Nb) | phi_dot (Nfil) | theta_dot (Nfil)].
struct SaddleEpilogue {
    const double* in = nullptr;
    double* out = nullptr;
    const double* kphi = nullptr;
    const double* kth = nullptr;
    int nfil = 0, nseg = 0, dyn_phase = 0, dyn_angle = 0;
    int64_t off_phase = 0, off_angle = 0;
    // free bodies: particle rows also get -(V + Omega x (x - x_body)) of
    // their swimmer, V at in[off_body + 6 s]
    const double* x_bodies = nullptr;
    int free_body = 0, segs_per_swim = 0, blobs_per_swim = 0;
    int64_t off_body = 0;
    // independent bodies: the body of each filament, for the per-body stop
    // flags of MatvecArgs::body_skip
    const int32_t* fil_body = nullptr;
};

struct OutArgs {
    View vs, vb;       // outputs; vs has seg_out_slots components
    int seg_out_slots; // 6 in the reference layout, 3 for a packed 3-dof vector
    uint32_t blocks;   // active blocks (bit 0 SS, 1 SB, 2 BB, 3 BS) for the near-pair terms
    SaddleEpilogue sad;
};

// Builds the work list for one evaluation of independent bodies.
// seg_write / blob_write: -1 no write, 0 overwrite, 1 accumulate.
void build_worklist(const std::vector<BodyRange>& bodies, int tile_targets, int chunk, bool ss, bool sb,
                    bool bb, bool bs, int seg_write, int blob_write, std::vector<MatvecTile>& tiles,
                    std::vector<MatvecItem>& items);
// The symmetric form's work list: 64-particle tiles (each body's segments,
// then its blobs, body after body) in groups of four, one item (tile = ga,
// src_cls = gb) per pair of groups ga <= gb that share a body, the diagonal
// ones last; every tile one partial slot per group its body overlaps.
// Returns the number of partial slots.
size_t build_sym_worklist(const std::vector<BodyRange>& bodies, int seg_write, int blob_write,
                          std::vector<MatvecTile>& tiles, std::vector<MatvecItem>& items);

// Rebuilds `near` for the given precision: every pair of particles of one
// body whose squared separation, in the pair kernel's own arithmetic for that
// precision, is <= (ai+aj)^2, found through a hashed uniform grid (O(n)).
// body_ranges: nbody BodyRange records on the device.  Grows `grid` as
// needed; returns the largest per-target count in *max_count (synchronises).
hipError_t build_near_list(int precision, const MatvecArgs& a, NearGrid& grid, const int32_t* body_ranges, int nbody,
                           NearList& near, int n_targets, int* max_count, int* host_scratch, hipStream_t stream);
void free_grid(NearGrid& g);
// Barrier forces of the positions in `a` (xs, xb, ns, nb, radii): fs += on
// segment slots 0-2 (stride 6), fb = (stride 3, may be null).  Uses `grid`
// as scratch.
hipError_t launch_barrier(const MatvecArgs& a, NearGrid& grid, const BarrierParams& bp, double* fs, double* fb,
                          hipStream_t stream);

// Launches pair kernel + reduce on `stream`.  d_tiles/d_items/partial are device
// buffers sized for the work list; precision 0 = fp64, 1 = fp32 tiles with fp64
// accumulation.  ev_begin/ev_end (may be null) are recorded around the pair
// kernel only.
hipError_t launch_matvec(int precision, int seg_dof, const TileShape& shape, const MatvecArgs& a,
                         const OutArgs& o, const MatvecTile* d_tiles, int n_tiles,
                         const MatvecItem* d_items, int n_items, double* partial, const NearList& near,
                         const PersistState& ps, hipStream_t stream, hipEvent_t ev_begin = nullptr,
                         hipEvent_t ev_end = nullptr);

// Half-space domain (no-slip wall at z = 0): the wall pair kernel (FP64, or
// FP32 tiles from fixed-point separations) over a work list of
// kWallBlock-target tiles (one target per lane), then the reduce kernel
// without near-pair terms.
constexpr int kWallBlock = 256;
hipError_t launch_matvec_wall(int precision, int seg_dof, const MatvecArgs& a, const OutArgs& o,
                              const MatvecTile* d_tiles, int n_tiles, const MatvecItem* d_items, int n_items,
                              double* partial, hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end);

// Flow-field evaluation (Stokeslet + rotlet mode and RPY mode).
struct FlowArgs {
    CView xt;          // targets (3 components)
    View vt;           // velocities out (3 components)
    int64_t nt;
    CView xs, fs;      // segment sources: positions, force+torque (6 components)
    CView xb, fb;      // blob sources: positions, force (3 components)
    int ns_total, nb_total;
    int ns_per_body, nb_per_body, num_bodies;
    const double* x_bodies;  // 3 per body, for net-force removal
    int remove_net_force;
    int mode;          // 0 Stokeslet + rotlet (+ Blakelet without blobs), 1 RPY point tracer
    double a_seg, a_blob, mu;
};
// net[3*b..] = total force on body b (segments + blobs).
hipError_t launch_flow_net(const FlowArgs& a, double* net, int nbod, hipStream_t stream);
hipError_t launch_flow_eval(int precision, const FlowArgs& a, const double* net, int block, hipStream_t stream);

}  // namespace tfs

#define TFS_HIP_CHECK(expr)                                 \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return _e;                    \
    } while (0)
