// rpy_matvec.hip -- fused, tiled RPY mobility matvec for gfx950.
//
// Replaces the four reference kernels Mss_mult / Msb_mult / Mbb_mult /
// Mbs_mult (src/cilia/cuda_functions.cu:362-757) with one launch over a work
// list.  A work item pairs a tile of targets (segments or blobs of one body)
// with a chunk of sources (segments or blobs of the same body).  A workgroup
// stages its chunk through LDS in batches of BLOCK sources (one coalesced
// load per lane, converted to the evaluation type as it is stored); every
// lane then walks the batch with broadcast LDS reads and accumulates
// PER_THREAD targets in registers.  Each item writes its partial sums to its
// own slot; a second kernel adds the slots of a tile in a fixed order and
// applies the reference's write rule (= or +=), so results are deterministic
// run to run, unlike atomics.
//
// The pair loop is branch-free: it evaluates the far-field RPY forms for
// every pair, with r^2 clamped below at (ai+aj)^2 (the self pair included).
// The pairs that clamp touches -- r^2 <= (ai+aj)^2, a handful per particle --
// are listed once per geometry by build_near_list (a hashed uniform grid),
// using the pair loop's own arithmetic for r^2, and the reduce kernel takes out their clamped far-field
// value and adds the reference's overlap and inner-branch formulas in FP64.
// Every pair is counted exactly once.
//
// Splitting the source range into chunks is what fills the chip: 21,800
// targets are only 341 wave64s, far fewer than the 1,024 SIMDs of an MI355X.
// The default shape also splits each item's sources between the waves of its
// workgroup (the waves share one tile of 64 x per_thread targets and add
// their sums in LDS before the item's one store): 12,800 segments + 9,000
// blobs in FP32 tiles are ~10k items of 64 targets x 768 sources, each
// writing 64 x 3 FP32 sums, a quarter of what four waves with their own
// targets would write.
//
// Two evaluation types:
//   double : the FP64 path (parity with the FP64 oracle at <= 1e-12).
//   float  : pair tensors in FP32 from fixed-point positions (per body,
//            X = rint(2^k (x - centre)) with |X| < 2^30): the separation is an
//            exact integer difference rounded once to FP32, so near pairs --
//            the ones that dominate -- lose nothing to cancellation, in two
//            full-rate instructions per component; per-batch FP32 partial sums
//            are flushed into FP64 accumulators after every LDS batch.
#include <algorithm>
#include <type_traits>

#include "rpy_pair.h"
#include "tfs_internal.h"

namespace tfs {
namespace {

constexpr int kMaxDof = 6;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X

// LDS record length (in T) of one staged source with FS strength components
// for targets with VT velocity components.  FP64: x y z f0..f(FS-1) [pad].
// FP32: X Y Z (fixed-point integers, stored as their bits) f0..f(FS-1), then
// the negated torque (FS = 6) and the negated force (VT = 6), padded to 16 B:
// with the negations staged, every cross-product term of the pair loop is an
// accumulate (the double-rate VOP2 form on gfx950) instead of a negated fma.
template <typename T, int FS, int VT = 3>
constexpr int rec_len() {
    if constexpr (sizeof(T) == 8) {
        return FS == 6 ? 10 : 6;
    } else {
        return (3 + FS + (FS == 6 ? 3 : 0) + (VT == 6 ? 3 : 0) + 3) / 4 * 4;
    }
}

// Largest record over the class combinations a kernel can stage.
template <typename T>
constexpr int rec_len_max() {
    return rec_len<T, 6, 6>();
}

// Separation and squared distance exactly as the pair loop evaluates them;
// build_near_list calls the same functions so the two agree bit for bit.
// FP64: d = x_i - x_j.  FP32: d = lambda (x_i - x_j) from the fixed-point
// coordinates -- an exact integer difference (|X| < 2^30) rounded once to
// FP32, two full-rate instructions per component.
__device__ __forceinline__ double separation(const double* xi, const double* rec, double& dx, double& dy,
                                             double& dz) {
    dx = xi[0] - rec[0];
    dy = xi[1] - rec[1];
    dz = xi[2] - rec[2];
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}
__device__ __forceinline__ float separation(const int32_t* Xi, const float* rec, float& dx, float& dy, float& dz) {
    dx = float(Xi[0] - __float_as_int(rec[0]));
    dy = float(Xi[1] - __float_as_int(rec[1]));
    dz = float(Xi[2] - __float_as_int(rec[2]));
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Target coordinates in the evaluation type's form.
__device__ __forceinline__ void load_target(const MatvecArgs& a, int cls, int t, double* xi) {
    const CView x = cls == 0 ? a.xs : a.xb;
#pragma unroll
    for (int k = 0; k < 3; ++k) xi[k] = x.p[(int64_t)t * x.ld_p + k * x.ld_c];
}
__device__ __forceinline__ void load_target(const MatvecArgs& a, int cls, int t, int32_t* Xi) {
    const int4 q = a.xfx[(cls == 0 ? 0 : a.ns) + t];
    Xi[0] = q.x;
    Xi[1] = q.y;
    Xi[2] = q.z;
}

// One source of class src_cls into its LDS record: its position and, if
// with_force, its strengths -- forces multiplied by fscale, torques by
// tscale.
template <typename T, int FS, int VT = 3>
__device__ __forceinline__ void stage_source(T* rec, const MatvecArgs& a, int src_cls, int s, bool with_force,
                                             double fscale = 1.0, double tscale = 1.0) {
    constexpr int F0 = 3;
    if constexpr (sizeof(T) == 4) {
        const int4 q = a.xfx[(src_cls == 0 ? 0 : a.ns) + s];
        rec[0] = __int_as_float(q.x);
        rec[1] = __int_as_float(q.y);
        rec[2] = __int_as_float(q.z);
    } else {
        const CView x = src_cls == 0 ? a.xs : a.xb;
#pragma unroll
        for (int k = 0; k < 3; ++k) rec[k] = x.p[(int64_t)s * x.ld_p + k * x.ld_c];
    }
    if (with_force) {
        const CView f = src_cls == 0 ? a.fs : a.fb;
#pragma unroll
        for (int k = 0; k < FS; ++k) rec[F0 + k] = T((k < 3 ? fscale : tscale) * f.p[(int64_t)s * f.ld_p + k * f.ld_c]);
        if constexpr (sizeof(T) == 4) {
            int o = F0 + FS;
            if constexpr (FS == 6) {
#pragma unroll
                for (int k = 0; k < 3; ++k) rec[o + k] = -rec[F0 + 3 + k];
                o += 3;
            }
            if constexpr (VT == 6) {
#pragma unroll
                for (int k = 0; k < 3; ++k) rec[o + k] = -rec[F0 + k];
            }
        }
    }
}

// Target coordinates of a tile for the lanes of this thread (clamped into
// the tile; surplus lanes compute a duplicate and their slot entries are
// never read).  A tile has LANES x R targets: LANES = the workgroup size, or
// 64 when the waves split the sources (every wave then holds the same
// targets).
template <typename T>
using TargetT = typename std::conditional<sizeof(T) == 4, int32_t, double>::type;

template <typename T, int LANES, int R>
__device__ __forceinline__ void load_targets(const MatvecTile& tl, const MatvecArgs& a, TargetT<T> (&xi)[R][3]) {
    const int lane = (int)threadIdx.x & (LANES - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int t = tl.t_begin + lane + r * LANES;
        t = t < tl.t_end ? t : tl.t_end - 1;
        load_target(a, tl.cls, t, xi[r]);
    }
}

// The pairs of one work item added into acc (the first VT entries of each
// target's AV).  FP32 tiles work in the fixed-point frame d' = lambda d
// (lambda = 2^k, exact): with forces staged as lambda c8 F and torques as
// lambda^2 c8 tau, the translational sums come out in physical units and the
// rotational ones divided by lambda, which store_slot multiplies back.  FP32
// sums one LDS batch in `part` and flushes it into the FP64 accumulators.
// With SPLIT, wave w evaluates records [64 w, 64 w + 64) of every batch.
template <int VT, int FS, typename T, int BLOCK, int R, int AV, bool SPLIT = false>
__device__ __forceinline__ void accumulate_item(const MatvecItem it, const MatvecTile tl, const MatvecArgs& a,
                                                const TargetT<T> (&xi)[R][3], double (&acc)[R][AV],
                                                T* __restrict__ lds) {
    constexpr int REC = rec_len<T, FS, VT>();
    constexpr int F0 = 3;
    constexpr bool F32 = sizeof(T) == 4;
    const int tid = threadIdx.x;

    const double ai = tl.cls == 0 ? a.a_seg : a.a_blob;
    const double aj = it.src_cls == 0 ? a.a_seg : a.a_blob;
    const double PI = 3.14159265358979323846;
    const double lam = F32 ? a.lam : 1.0;
    const T sum_a2 = T(lam * lam * ((ai + aj) * (ai + aj)));
    const T s3 = T(lam * lam * ((ai * ai + aj * aj) / 3.0));
    // -3 s3 as its own value (a negated operand would force the VOP3 form of the fma)
    const T m3s3 = T(-3) * s3;
    const double c8 = 1.0 / (8.0 * PI * a.mu);  // folded into the staged strengths

    T part[R][F32 ? VT : 1];
    // this wave's records of a batch start at j0 (uniform over the wave)
    const int j0 = SPLIT ? __builtin_amdgcn_readfirstlane(tid >> 6) * 64 : 0;
    for (int s0 = it.s_begin; s0 < it.s_end; s0 += BLOCK) {
        const int n = (it.s_end - s0) < BLOCK ? (it.s_end - s0) : BLOCK;
        const int j1 = SPLIT ? (n < j0 + 64 ? n : j0 + 64) : n;
        __syncthreads();  // the previous batch has been consumed
        if (tid < n) stage_source<T, FS, VT>(lds + tid * REC, a, it.src_cls, s0 + tid, true, lam * c8, lam * lam * c8);
        __syncthreads();
        if constexpr (F32) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < VT; ++k) part[r][k] = T(0);
        }

#pragma unroll 2
        for (int j = j0; j < j1; ++j) {
            const T* rec = lds + j * REC;
            T f[FS];
#pragma unroll
            for (int k = 0; k < FS; ++k) f[k] = rec[F0 + k];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                T dx, dy, dz;
                const T r2 = separation(xi[r], rec, dx, dy, dz);
                // Overlapping pairs (and the self pair) are evaluated here with r
                // clamped to ai + aj -- one max instead of a compare and select --
                // and the reduce kernel swaps that value for the exact one.
                const T rm1 = rsqrt_refined(fmax_t(r2, sum_a2));
                if constexpr (F32) pair_far_scaled<T, VT, FS>(part[r], dx, dy, dz, rm1, f, rec + F0 + FS, s3, m3s3);
                else pair_far_scaled<T, VT, FS>(acc[r], dx, dy, dz, rm1, f, nullptr, s3, m3s3);
            }
        }
        if constexpr (F32) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < VT; ++k) acc[r][k] += double(part[r][k]);
        }
    }
}

// A tile's sums (VT per target) into its partial slot (LANES x R targets).
// FP32 tiles store them in FP32: the reduce kernel reads half the bytes.
template <int VT, typename T, int LANES, int R, int AV>
__device__ __forceinline__ void store_slot(T* __restrict__ slot, const double (&acc)[R][AV], double lam) {
    constexpr int TILE = LANES * R;
    const int lane = (int)threadIdx.x & (LANES - 1);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < VT; ++k)
            slot[k * TILE + lane + r * LANES] = T(k < 3 ? acc[r][k] : lam * acc[r][k]);
}

// Source split: waves 1 .. S-1 hand their sums to wave 0 through LDS, which
// adds them in wave order (a fixed order: results are reproducible).
template <int VT, int BLOCK, int R, int AV>
__device__ __forceinline__ void combine_waves(double (&acc)[R][AV], double* __restrict__ buf) {
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    __syncthreads();  // every wave is done with the last batch's records
    if (w > 0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k = 0; k < VT; ++k) buf[(((w - 1) * R + r) * VT + k) * 64 + lane] = acc[r][k];
    }
    __syncthreads();
    if (w == 0) {
        for (int v = 1; v < BLOCK / 64; ++v)
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < VT; ++k) acc[r][k] += buf[(((v - 1) * R + r) * VT + k) * 64 + lane];
    }
}

// LDS of the one-workgroup-per-item kernel, in T: a batch of the largest
// records, or (source split) the sums of BLOCK/64 - 1 waves, if larger.
template <typename T, int BLOCK, int R, bool SPLIT>
constexpr int item_lds_len() {
    constexpr int batch = BLOCK * rec_len_max<T>();
    constexpr int sums = SPLIT ? (BLOCK / 64 - 1) * 64 * R * kMaxDof * (int)(sizeof(double) / sizeof(T)) : 0;
    return batch > sums ? batch : sums;
}

// The pairs of one work item into its own slot.
template <int VT, int FS, typename T, int BLOCK, int R, bool SPLIT>
__device__ __forceinline__ void run_item(const MatvecItem it, const MatvecTile tl, const MatvecArgs& a,
                                         T* __restrict__ slot, T* __restrict__ lds) {
    constexpr int LANES = SPLIT ? 64 : BLOCK;
    TargetT<T> xi[R][3];
    load_targets<T, LANES, R>(tl, a, xi);
    double acc[R][VT];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < VT; ++k) acc[r][k] = 0.0;
    accumulate_item<VT, FS, T, BLOCK, R, VT, SPLIT>(it, tl, a, xi, acc, lds);
    if constexpr (SPLIT) {
        combine_waves<VT, BLOCK, R, VT>(acc, reinterpret_cast<double*>(lds));
        if (threadIdx.x >= 64) return;
    }
    store_slot<VT, T, LANES, R, VT>(slot, acc, sizeof(T) == 4 ? a.lam : 1.0);
}

// One workgroup per item.  With `stamps`, thread 0 stores the item's start
// and finish (wall_clock64 ticks; equal for an item of a skipped body).
// SPLIT: the BLOCK/64 waves share the tile's 64 x R targets and split every
// batch of sources (BLOCK/64 times fewer partial bytes per item).
template <typename T, int SEGDOF, int BLOCK, int R, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void rpy_pairs_kernel(const MatvecItem* __restrict__ items,
                                                          const MatvecTile* __restrict__ tiles,
                                                          MatvecArgs a, T* __restrict__ partial,
                                                          int64_t* __restrict__ stamps) {
    constexpr int TILE = (SPLIT ? 64 : BLOCK) * R;
    __shared__ __align__(16) T lds[item_lds_len<T, BLOCK, R, SPLIT>()];
    if (a.skip && *a.skip) return;
    const int64_t t_start = stamps ? (int64_t)wall_clock64() : 0;
    const MatvecItem it = items[blockIdx.x];
    const MatvecTile tl = tiles[it.tile];
    if (!(a.body_skip && a.body_skip[tl.body])) {
        T* slot = partial + (size_t)blockIdx.x * (kMaxDof * TILE);
        if (tl.cls == 0) {
            if (it.src_cls == 0) run_item<SEGDOF, SEGDOF, T, BLOCK, R, SPLIT>(it, tl, a, slot, lds);
            else run_item<SEGDOF, 3, T, BLOCK, R, SPLIT>(it, tl, a, slot, lds);
        } else {
            if (it.src_cls == 0) run_item<3, SEGDOF, T, BLOCK, R, SPLIT>(it, tl, a, slot, lds);
            else run_item<3, 3, T, BLOCK, R, SPLIT>(it, tl, a, slot, lds);
        }
    }
    if (stamps && threadIdx.x == 0) {
        const int64_t t_end = (a.body_skip && a.body_skip[tl.body]) ? t_start : (int64_t)wall_clock64();
        stamps[2 * (int64_t)blockIdx.x] = t_start;
        stamps[2 * (int64_t)blockIdx.x + 1] = t_end;
    }
}

// ---- symmetric form (3-DOF particles) -----------------------------------------
//
// The far-field tensor of a pair is the same seen from either particle: the
// separation enters twice (its sign cancels) and the radii symmetrically, so
// M_ji = M_ij and one evaluation of 1/r, A and B serves both.  Tiles of 64
// particles are grouped kSymTiles to a workgroup; item (ga, gb), ga <= gb,
// evaluates the tile pairs of groups ga and gb, wave w owning tile w of group
// ga.  In stage t the wave takes source tile (w + t) mod kSymTiles of group gb
// from LDS, lane l pairing its particle with source (l + kSymRot k) mod 64 at
// step k: the lane's own sum gets the source's term and the source's sum gets
// the lane's term.  The source sums live in one register per component that
// moves one lane per step (a DPP wave rotation ahead of the accumulates), so
// after 64 steps lane l holds the sum of source (l + 63 kSymRot) mod 64.  Each stage
// adds those sums into the source tile's FP64 LDS sums; the stages are fenced
// by barriers and the writer of a tile's sums in stage t is wave
// (tile - t) mod kSymTiles, a fixed order.  A pair of distinct groups is
// evaluated once for both tiles; in a diagonal item (ga == gb) the pairs with
// tile a < tile b are evaluated once, a == b one-sided (the self pair and the
// overlapping pairs take the clamped far field, which the reduce kernel swaps
// for the exact value), and a > b not at all.
//
// Partial slots: tile t of group g has one slot per group q its body overlaps
// (groups run body after body; tile pairs of different bodies are skipped),
// holding the sums it received from group q (its own sums when q > g, the
// source sums when q < g, both added for q == g), each written once per
// launch, so the reduce kernel sums them in slot order.
constexpr int kSymTiles = kSymGroupTiles;
constexpr int kSymRot = 1;  // DPP wave_rol:1 -- lane l reads lane l + 1 (mod 64)
__device__ __forceinline__ float rotate_lanes(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), kSymRot > 0 ? 0x134 : 0x13C, 0xF, 0xF, false));
}
__device__ __forceinline__ double rotate_lanes(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), kSymRot > 0 ? 0x134 : 0x13C, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), kSymRot > 0 ? 0x134 : 0x13C, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// LDS records of one source tile, twice over (index j and j + 64 hold
// particle j, so that lane l reads record l + kSymRot k + (kSymRot < 0 ? 64 :
// 0) without a wrap), split into arrays of 16 or 8 bytes per record so that
// consecutive lanes, reading consecutive records, share no banks within a
// read phase.  FP32: X Y Z (fixed-point integers as their bits) and the first
// strength, then the other two strengths.  FP64: x y, z and the first
// strength, the other two.
template <typename T>
struct SymTileLds;
template <>
struct SymTileLds<float> {
    float4 xf[128];
    float2 yz[128];
};
template <>
struct SymTileLds<double> {
    double2 xy[128];
    double2 zf[128];
    double2 gh[128];
};

__device__ __forceinline__ void sym_record(const SymTileLds<float>& L, int j, float (&rec)[3], float (&f)[3]) {
    const float4 q = L.xf[j];
    const float2 g = L.yz[j];
    rec[0] = q.x;
    rec[1] = q.y;
    rec[2] = q.z;
    f[0] = q.w;
    f[1] = g.x;
    f[2] = g.y;
}
__device__ __forceinline__ void sym_record(const SymTileLds<double>& L, int j, double (&rec)[3], double (&f)[3]) {
    const double2 a = L.xy[j], b = L.zf[j], c = L.gh[j];
    rec[0] = a.x;
    rec[1] = a.y;
    rec[2] = b.x;
    f[0] = b.y;
    f[1] = c.x;
    f[2] = c.y;
}
__device__ __forceinline__ void sym_put(SymTileLds<float>& L, int j, const int32_t (&X)[3], const float (&f)[3]) {
    L.xf[j] = make_float4(__int_as_float(X[0]), __int_as_float(X[1]), __int_as_float(X[2]), f[0]);
    L.yz[j] = make_float2(f[1], f[2]);
}
__device__ __forceinline__ void sym_put(SymTileLds<double>& L, int j, const double (&X)[3], const double (&f)[3]) {
    L.xy[j] = make_double2(X[0], X[1]);
    L.zf[j] = make_double2(X[2], f[0]);
    L.gh[j] = make_double2(f[1], f[2]);
}

// The 64 steps of one stage.  BOTH: add the lane's terms into the rotating
// source sums pj as well (an off-diagonal tile pair); otherwise only the
// lane's own sum pi (the tile against itself).  FP32 tiles work in the
// fixed-point frame (separation from integer coordinates, s3 and (ai+aj)^2
// scaled by lambda^2).
template <bool BOTH, typename T>
__device__ __forceinline__ void sym_stage(const SymTileLds<T>& recs, int lane, const TargetT<T> (&Xi)[3],
                                          const T (&fi)[3], T sum_a2, T s3, T m3s3, T (&pi)[3], T (&pj)[3]) {
    const int j0 = lane + (kSymRot > 0 ? 0 : 64);
    // Not unrolled: FP32 48 VGPRs, eight waves per SIMD to hide the LDS and
    // dependency latency (measured on MI355X, production body: 0.212 ms per
    // launch, against 0.219 unrolled twice (59 VGPRs) and 0.239 eight times
    // (119 VGPRs, four waves per SIMD)).
#pragma unroll 1
    for (int k = 0; k < 64; ++k) {
        T rec[3], fj[3];
        sym_record(recs, j0 + kSymRot * k, rec, fj);
        T dx, dy, dz;
        const T r2 = separation(Xi, rec, dx, dy, dz);
        const T rm1 = rsqrt_refined(fmax_t(r2, sum_a2));
        const T rm2 = rm1 * rm1;
        const T rm3 = rm1 * rm2;
        const T A = fma_t(s3, rm3, rm1);
        const T u = m3s3 * rm2;
        const T B = fma_t(u, rm3, rm3);
        const T Bd = B * fma_t(dz, fj[2], fma_t(dy, fj[1], dx * fj[0]));
        pi[0] = fma_t(Bd, dx, fma_t(A, fj[0], pi[0]));
        pi[1] = fma_t(Bd, dy, fma_t(A, fj[1], pi[1]));
        pi[2] = fma_t(Bd, dz, fma_t(A, fj[2], pi[2]));
        if constexpr (BOTH) {
            // the source's term: the same A, B and separation (d_ji = -d
            // enters twice), the lane's strengths
            const T Be = B * fma_t(dz, fi[2], fma_t(dy, fi[1], dx * fi[0]));
            pj[0] = fma_t(Be, dx, fma_t(A, fi[0], rotate_lanes(pj[0])));
            pj[1] = fma_t(Be, dy, fma_t(A, fi[1], rotate_lanes(pj[1])));
            pj[2] = fma_t(Be, dz, fma_t(A, fi[2], rotate_lanes(pj[2])));
        }
    }
}

// One particle of a tile for the symmetric form: its position (FP32 tiles:
// fixed-point) and strengths scaled by fsc (zero, with the tile's first
// position, for lanes past the tile's end).
__device__ __forceinline__ void sym_particle(const MatvecTile& tl, const MatvecArgs& a, int lane, double fsc,
                                             int32_t (&X)[3], float (&f)[3]) {
    const int s = tl.t_begin + lane;
    const bool ok = s < tl.t_end;
    const int sc = ok ? s : tl.t_begin;
    const int4 q = a.xfx[(tl.cls == 0 ? 0 : a.ns) + sc];
    X[0] = q.x;
    X[1] = q.y;
    X[2] = q.z;
    const CView fv = tl.cls == 0 ? a.fs : a.fb;
#pragma unroll
    for (int k = 0; k < 3; ++k) f[k] = ok ? float(fsc * fv.p[(int64_t)sc * fv.ld_p + k * fv.ld_c]) : 0.0f;
}
__device__ __forceinline__ void sym_particle(const MatvecTile& tl, const MatvecArgs& a, int lane, double fsc,
                                             double (&X)[3], double (&f)[3]) {
    const int s = tl.t_begin + lane;
    const bool ok = s < tl.t_end;
    const int sc = ok ? s : tl.t_begin;
    const CView xv = tl.cls == 0 ? a.xs : a.xb;
    const CView fv = tl.cls == 0 ? a.fs : a.fb;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        X[k] = xv.p[(int64_t)sc * xv.ld_p + k * xv.ld_c];
        f[k] = ok ? fsc * fv.p[(int64_t)sc * fv.ld_p + k * fv.ld_c] : 0.0;
    }
}

// One workgroup of kSymTiles waves per item (tile = ga, src_cls = gb).
// `bb`: blob-blob pairs are evaluated (off when an overwriting BS block
// replaces BB).  Tile pairs of different bodies are skipped, and so is an item
// whose bodies all have their stop flag set; with `stamps`, thread 0 stores
// the item's start and finish (wall_clock64 ticks).
template <typename T>
__global__ __launch_bounds__(64 * kSymTiles) void rpy_pairs_sym_kernel(const MatvecItem* __restrict__ items,
                                                                        const MatvecTile* __restrict__ tiles,
                                                                        int n_tiles, int bb, MatvecArgs a,
                                                                        T* __restrict__ partial,
                                                                        int64_t* __restrict__ stamps) {
    constexpr bool F32 = sizeof(T) == 4;
    __shared__ SymTileLds<T> recs[kSymTiles];
    __shared__ double jsum[kSymTiles][3][64];
    if (a.skip && *a.skip) return;
    const int64_t t_start = stamps ? (int64_t)wall_clock64() : 0;
    const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const MatvecItem it = items[blockIdx.x];
    const int ga = it.tile, gb = it.src_cls;
    if (a.body_skip) {
        // the bodies of the item's tiles are a contiguous range
        const int b0 = tiles[kSymTiles * ga].body;
        const int tl_last = kSymTiles * gb + kSymTiles < n_tiles ? kSymTiles * gb + kSymTiles : n_tiles;
        const int b1 = tiles[tl_last - 1].body;
        bool any = false;
        for (int b = b0; b <= b1 && !any; ++b) any = a.body_skip[b] == 0;
        if (!any) {
            if (stamps && threadIdx.x == 0) {
                stamps[2 * (int64_t)blockIdx.x] = t_start;
                stamps[2 * (int64_t)blockIdx.x + 1] = t_start;
            }
            return;
        }
    }
    const double PI = 3.14159265358979323846;
    const double lam = F32 ? a.lam : 1.0;
    const double fsc = lam / (8.0 * PI * a.mu);
    // wave w stages source tile w of group gb (twice: records lane and lane + 64)
    {
        const int tb = kSymTiles * gb + w;
        if (tb < n_tiles) {
            TargetT<T> X[3];
            T f[3];
            sym_particle(tiles[tb], a, lane, fsc, X, f);
            sym_put(recs[w], lane, X, f);
            sym_put(recs[w], lane + 64, X, f);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) jsum[w][k][lane] = 0.0;
    }
    const int ta = kSymTiles * ga + w;
    const bool has_a = ta < n_tiles;
    TargetT<T> Xi[3] = {0, 0, 0};
    T fi[3] = {T(0), T(0), T(0)};
    MatvecTile tla{};
    bool run_a = false;
    if (has_a) {
        tla = tiles[ta];
        run_a = !(a.body_skip && a.body_skip[tla.body]);
        sym_particle(tla, a, lane, fsc, Xi, fi);
    }
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < kSymTiles; ++t) {
        const int bq = (w + t) & (kSymTiles - 1);
        const int tb = kSymTiles * gb + bq;
        int mode = 0;  // 0 nothing, 1 one-sided, 2 both particles
        int cls_b = 0;
        if (run_a && tb < n_tiles) {
            const MatvecTile tlb = tiles[tb];
            cls_b = tlb.cls;
            mode = (ga < gb || ta < tb) ? 2 : (ta == tb ? 1 : 0);
            if ((tla.cls == 1 && cls_b == 1 && !bb) || tlb.body != tla.body) mode = 0;
        }
        if (mode) {
            const double ai = tla.cls == 0 ? a.a_seg : a.a_blob;
            const double aj = cls_b == 0 ? a.a_seg : a.a_blob;
            const T sum_a2 = T(lam * lam * ((ai + aj) * (ai + aj)));
            const T s3 = T(lam * lam * ((ai * ai + aj * aj) / 3.0));
            const T m3s3 = T(-3) * s3;
            T pi[3] = {T(0), T(0), T(0)}, pj[3] = {T(0), T(0), T(0)};
            if (mode == 2) sym_stage<true, T>(recs[bq], lane, Xi, fi, sum_a2, s3, m3s3, pi, pj);
            else sym_stage<false, T>(recs[bq], lane, Xi, fi, sum_a2, s3, m3s3, pi, pj);
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += double(pi[k]);
            if (mode == 2) {
                const int j = (lane + 63 * kSymRot) & 63;
#pragma unroll
                for (int k = 0; k < 3; ++k) jsum[bq][k][j] += double(pj[k]);
            }
        }
        __syncthreads();  // this stage's source sums are in before the next writer's
    }
    // slot of (tile, partner group q) if q is one of the groups of the tile's
    // body: 3 x 64 values
    auto slot = [&](const MatvecTile& tl, int q) -> T* {
        return (q >= tl.group0 && q < tl.group0 + tl.n_items)
                   ? partial + ((size_t)tl.first_item + (q - tl.group0)) * (kSymSlotDof * 64)
                   : nullptr;
    };
    if (ga < gb) {
        if (has_a) {
            if (T* s = slot(tla, gb)) {
#pragma unroll
                for (int k = 0; k < 3; ++k) s[k * 64 + lane] = T(acc[k]);
            }
        }
        const int tb = kSymTiles * gb + w;
        if (tb < n_tiles) {
            if (T* s = slot(tiles[tb], ga)) {
#pragma unroll
                for (int k = 0; k < 3; ++k) s[k * 64 + lane] = T(jsum[w][k][lane]);
            }
        }
    } else if (has_a) {
        T* s = slot(tla, ga);
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k * 64 + lane] = T(acc[k] + jsum[w][k][lane]);
    }
    if (stamps && threadIdx.x == 0) {
        stamps[2 * (int64_t)blockIdx.x] = t_start;
        stamps[2 * (int64_t)blockIdx.x + 1] = (int64_t)wall_clock64();
    }
}

// Half-space domain (no-slip wall at z = 0): every pair of an item through
// pair_wall, one target per lane; the self pair (same global index) takes the
// wall self-mobility.  FP64 tiles stage positions as they are.  FP32 tiles
// take separations from the fixed-point positions (exact integer
// differences, scaled by 1/lambda) and heights as floats of the FP64
// positions, sum one LDS batch in FP32 and flush it into FP64 accumulators.
// Partial slots as run_item, in the evaluation type.
template <int VT, int FS, typename T>
__device__ void run_item_wall(const MatvecItem it, const MatvecTile tl, const MatvecArgs& a,
                              T* __restrict__ slot, T* __restrict__ lds) {
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int F0 = F32 ? 4 : 3;          // FP32 records: X bits (3), height, strengths
    constexpr int REC = F0 + FS;
    constexpr int TILE = kWallBlock;
    const int tid = threadIdx.x;
    int t = tl.t_begin + tid;
    t = t < tl.t_end ? t : tl.t_end - 1;
    const CView xt = tl.cls == 0 ? a.xs : a.xb;
    const double zi_d = xt.p[(int64_t)t * xt.ld_p + 2 * xt.ld_c];
    double xi[3];
    int32_t Xi[3];
    if constexpr (F32) load_target(a, tl.cls, t, Xi);
    else load_target(a, tl.cls, t, xi);
    const T zi = T(zi_d);
    const T inv_lam = T(F32 ? 1.0 / a.lam : 1.0);
    const int tg = (tl.cls == 0 ? 0 : a.ns) + t;
    const int src_base = it.src_cls == 0 ? 0 : a.ns;
    const CView xsrc = it.src_cls == 0 ? a.xs : a.xb;
    const CView fsrc = it.src_cls == 0 ? a.fs : a.fb;
    double acc[VT];
    T part[VT];
#pragma unroll
    for (int k = 0; k < VT; ++k) acc[k] = 0.0;
    for (int s0 = it.s_begin; s0 < it.s_end; s0 += kWallBlock) {
        const int n = (it.s_end - s0) < kWallBlock ? (it.s_end - s0) : kWallBlock;
        __syncthreads();
        if (tid < n) {
            T* rec = lds + tid * REC;
            const int s = s0 + tid;
            if constexpr (F32) {
                const int4 q = a.xfx[src_base + s];
                rec[0] = __int_as_float(q.x);
                rec[1] = __int_as_float(q.y);
                rec[2] = __int_as_float(q.z);
                rec[3] = float(xsrc.p[(int64_t)s * xsrc.ld_p + 2 * xsrc.ld_c]);
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) rec[k] = xsrc.p[(int64_t)s * xsrc.ld_p + k * xsrc.ld_c];
            }
#pragma unroll
            for (int k = 0; k < FS; ++k) rec[F0 + k] = T(fsrc.p[(int64_t)s * fsrc.ld_p + k * fsrc.ld_c]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < VT; ++k) part[k] = T(0);
        for (int j = 0; j < n; ++j) {
            const T* rec = lds + j * REC;
            T dx, dy, dz, zj;
            if constexpr (F32) {
                dx = float(Xi[0] - __float_as_int(rec[0])) * inv_lam;
                dy = float(Xi[1] - __float_as_int(rec[1])) * inv_lam;
                dz = float(Xi[2] - __float_as_int(rec[2])) * inv_lam;
                zj = rec[3];
            } else {
                dx = xi[0] - rec[0];
                dy = xi[1] - rec[1];
                dz = xi[2] - rec[2];
                zj = rec[2];
            }
            pair_wall<T, VT, FS>(part, src_base + s0 + j == tg, dx, dy, dz, zi, zj, rec + F0);
        }
#pragma unroll
        for (int k = 0; k < VT; ++k) acc[k] += double(part[k]);
    }
#pragma unroll
    for (int k = 0; k < VT; ++k) slot[k * TILE + tid] = T(acc[k]);
}

template <int SEGDOF, typename T>
__global__ __launch_bounds__(kWallBlock) void rpy_pairs_wall_kernel(const MatvecItem* __restrict__ items,
                                                                    const MatvecTile* __restrict__ tiles,
                                                                    MatvecArgs a, T* __restrict__ partial) {
    __shared__ __align__(16) T lds[kWallBlock * (4 + 6)];
    if (a.skip && *a.skip) return;
    const MatvecItem it = items[blockIdx.x];
    const MatvecTile tl = tiles[it.tile];
    if (a.body_skip && a.body_skip[tl.body]) return;
    T* slot = partial + (size_t)blockIdx.x * (kMaxDof * kWallBlock);
    if (tl.cls == 0) {
        if (it.src_cls == 0) run_item_wall<SEGDOF, SEGDOF, T>(it, tl, a, slot, lds);
        else run_item_wall<SEGDOF, 3, T>(it, tl, a, slot, lds);
    } else {
        if (it.src_cls == 0) run_item_wall<3, SEGDOF, T>(it, tl, a, slot, lds);
        else run_item_wall<3, 3, T>(it, tl, a, slot, lds);
    }
}

// Persistent form: workgroup w takes the contiguous items [w k, (w+1) k) of
// the work list (items of one tile are contiguous, of similar size) and keeps
// one tile's sums in registers across its items, storing them once per tile
// segment, into the slot of the segment's first item: the reduce kernel reads
// only those slots (segment starts are the tiles' first items and the
// multiples of k), about k times fewer partial bytes than one slot per item.
// Loop counts and exits are uniform over the workgroup.  With timing on,
// every lane of a workgroup stores the same per-item start/finish stamps
// (wall_clock64 ticks; equal for an item of a skipped body).
template <typename T, int SEGDOF, int BLOCK, int R>
__global__ __launch_bounds__(BLOCK) void rpy_pairs_persistent_kernel(const MatvecItem* __restrict__ items,
                                                                     const MatvecTile* __restrict__ tiles,
                                                                     MatvecArgs a, T* __restrict__ partial,
                                                                     int n_items, int k, PersistState ps) {
    constexpr int REC = rec_len_max<T>();
    __shared__ __align__(16) T lds[BLOCK * REC];
    if (a.skip && *a.skip) return;
    const int i0 = blockIdx.x * k;
    const int i1 = i0 + k < n_items ? i0 + k : n_items;
    const double lam = sizeof(T) == 4 ? a.lam : 1.0;
    TargetT<T> xi[R][3];
    double acc[R][SEGDOF];
    int cur = -1, cur_cls = 0, seg = 0;
    auto flush = [&]() {
        T* slot = partial + (size_t)seg * (kMaxDof * BLOCK * R);
        if (cur_cls == 0) store_slot<SEGDOF, T, BLOCK, R, SEGDOF>(slot, acc, lam);
        else store_slot<3, T, BLOCK, R, SEGDOF>(slot, acc, lam);
    };
    for (int item = i0; item < i1; ++item) {
        if (ps.item_stamps) ps.item_stamps[2 * item] = (int64_t)wall_clock64();
        const MatvecItem it = items[item];
        const MatvecTile tl = tiles[it.tile];
        if (a.body_skip && a.body_skip[tl.body]) {  // uniform over the workgroup
            if (ps.item_stamps) ps.item_stamps[2 * item + 1] = ps.item_stamps[2 * item];
            continue;
        }
        if (it.tile != cur) {
            if (cur >= 0) flush();
            cur = it.tile;
            cur_cls = tl.cls;
            seg = item;
            load_targets<T, BLOCK, R>(tl, a, xi);
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int c = 0; c < SEGDOF; ++c) acc[r][c] = 0.0;
        }
        if (tl.cls == 0) {
            if (it.src_cls == 0) accumulate_item<SEGDOF, SEGDOF, T, BLOCK, R, SEGDOF>(it, tl, a, xi, acc, lds);
            else accumulate_item<SEGDOF, 3, T, BLOCK, R, SEGDOF>(it, tl, a, xi, acc, lds);
        } else {
            if (it.src_cls == 0) accumulate_item<3, SEGDOF, T, BLOCK, R, SEGDOF>(it, tl, a, xi, acc, lds);
            else accumulate_item<3, 3, T, BLOCK, R, SEGDOF>(it, tl, a, xi, acc, lds);
        }
        if (ps.item_stamps) ps.item_stamps[2 * item + 1] = (int64_t)wall_clock64();
    }
    if (cur >= 0) flush();
}

// ---------------------------------------------------------------------------
// Near-pair listing on a hashed uniform grid.  Cells have side h >= (1 + 1e-5)
// max(ai + aj), so every pair that either arithmetic (FP64, or FP32 fixed
// point, whose r^2 errs by well under 1e-5 relative) could classify as
// overlapping lies in neighbouring cells; each candidate is then classified
// with the pair kernel's own staging and separation functions for the
// requested precision, so the lists and the pair loop's clamp agree bit for
// bit.  One thread per target, entries appended in visiting order (the
// reduce kernel's order comes from near_sort_kernel).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void particle_pos(const MatvecArgs& a, int p, double* x) {
    const CView v = p < a.ns ? a.xs : a.xb;
    const int i = p < a.ns ? p : p - a.ns;
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = v.p[(int64_t)i * v.ld_p + k * v.ld_c];
}

__device__ __forceinline__ int cell_coord(double x, double inv_h) {
    const double c = __builtin_floor(x * inv_h);
    return (int)(c < -1073741824.0 ? -1073741824.0 : (c > 1073741824.0 ? 1073741824.0 : c));
}

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz, uint32_t mask) {
    return ((uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u ^ (uint32_t)cz * 83492791u) & mask;
}

__global__ __launch_bounds__(256) void grid_body_kernel(const int4* __restrict__ ranges, int ns,
                                                       int32_t* __restrict__ body_of) {
    const int4 r = ranges[blockIdx.x];
    for (int i = r.x + threadIdx.x; i < r.y; i += 256) body_of[i] = blockIdx.x;
    for (int i = r.z + threadIdx.x; i < r.w; i += 256) body_of[ns + i] = blockIdx.x;
}

__global__ __launch_bounds__(256) void grid_count_kernel(MatvecArgs a, int n, double inv_h, uint32_t mask,
                                                        int32_t* __restrict__ bucket_of,
                                                        int32_t* __restrict__ fill) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    double x[3];
    particle_pos(a, p, x);
    const uint32_t b = cell_hash(cell_coord(x[0], inv_h), cell_coord(x[1], inv_h), cell_coord(x[2], inv_h), mask);
    bucket_of[p] = (int32_t)b;
    atomicAdd(&fill[b], 1);
}

// start = exclusive prefix sum of fill (table entries, a power of two >= 1024),
// start[table] = total; fill is cleared for the scatter.  One workgroup, in
// chunks of up to 16384 entries (1024 x per).  The chunk is loaded and stored
// coalesced, every load (and store) of a thread issued together, and staged in
// LDS with one pad word per 16 entries, so that each thread's `per`
// consecutive entries (a stride of 17 words across the lanes) share no bank.
// Every thread sums its entries, the sums are scanned across each wave by
// shuffles and across the sixteen waves from their totals, and a running
// carry joins the chunks.
__global__ __launch_bounds__(1024) void grid_scan_kernel(int32_t* __restrict__ fill, int32_t* __restrict__ start,
                                                        int table) {
    constexpr int kChunk = 16384, kMaxPer = kChunk / 1024, kWaves = 1024 / 64;
    __shared__ int32_t buf[kChunk + kChunk / 16];
    __shared__ int32_t wsum[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const auto pad = [](int i) { return i + (i >> 4); };
    int32_t carry = 0;
    for (int c0 = 0; c0 < table; c0 += kChunk) {
        const int len = table - c0 < kChunk ? table - c0 : kChunk;
        const int per = len / 1024;
        int32_t v[kMaxPer];
#pragma unroll
        for (int k = 0; k < kMaxPer; ++k) v[k] = k < per ? fill[c0 + k * 1024 + tid] : 0;
#pragma unroll
        for (int k = 0; k < kMaxPer; ++k)
            if (k < per) buf[pad(k * 1024 + tid)] = v[k];
        __syncthreads();
        const int b = tid * per;  // this thread's first entry of the chunk
        int32_t sum = 0;
#pragma unroll
        for (int i = 0; i < kMaxPer; ++i)
            if (i < per) sum += buf[pad(b + i)];
        int32_t inc = sum;  // inclusive scan of the threads' sums across the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t t = __shfl_up(inc, off, 64);
            if (lane >= off) inc += t;
        }
        if (lane == 63) wsum[wid] = inc;
        __syncthreads();
        int32_t before = 0, total = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            const int32_t t = wsum[k];
            before += k < wid ? t : 0;
            total += t;
        }
        int32_t run = carry + before + inc - sum;  // exclusive prefix of this thread's entries
#pragma unroll
        for (int i = 0; i < kMaxPer; ++i)
            if (i < per) {
                const int32_t cnt = buf[pad(b + i)];
                buf[pad(b + i)] = run;
                run += cnt;
            }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kMaxPer; ++k) v[k] = k < per ? buf[pad(k * 1024 + tid)] : 0;
#pragma unroll
        for (int k = 0; k < kMaxPer; ++k)
            if (k < per) {
                start[c0 + k * 1024 + tid] = v[k];
                fill[c0 + k * 1024 + tid] = 0;
            }
        carry += total;
        __syncthreads();  // buf and wsum are written again by the next chunk
    }
    if (tid == 0) start[table] = carry;
}

__global__ __launch_bounds__(256) void grid_scatter_kernel(int n, const int32_t* __restrict__ bucket_of,
                                                          const int32_t* __restrict__ start,
                                                          int32_t* __restrict__ fill, int32_t* __restrict__ order) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int32_t b = bucket_of[p];
    order[start[b] + atomicAdd(&fill[b], 1)] = p;
}

// Puts each bucket's particles in increasing index order (the scatter's
// atomics leave them in arrival order), so that sums over a cell's particles
// are reproducible.  Insertion sort per bucket; buckets hold a few particles
// (cells of side ~2a hold at most a handful), larger ones are left as they are.
__global__ __launch_bounds__(256) void grid_sort_buckets_kernel(int table, const int32_t* __restrict__ start,
                                                               int32_t* __restrict__ order) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= table) return;
    const int lo = start[b], hi = start[b + 1];
    if (hi - lo < 2 || hi - lo > 256) return;
    for (int i = lo + 1; i < hi; ++i) {
        const int32_t v = order[i];
        int k = i - 1;
        while (k >= lo && order[k] > v) {
            order[k + 1] = order[k];
            --k;
        }
        order[k + 1] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void near_grid_kernel(MatvecArgs a, int n, double inv_h, uint32_t mask,
                                                       const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ body_of,
                                                       int32_t* __restrict__ cnt, int32_t* __restrict__ idx,
                                                       int cap) {
    // 32 lanes per target, lane l < 27 visiting neighbour cell l; near sources
    // are appended in any order (near_rank_sort_kernel orders them).  cnt must
    // start at zero and ends as the full count, entries past cap dropped.
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int REC = rec_len<T, 3>();
    const int tg = blockIdx.x * 8 + (int)(threadIdx.x >> 5);
    const int lane = (int)(threadIdx.x & 31);
    if (tg >= n || lane >= 27) return;
    const int cls = tg < a.ns ? 0 : 1;
    const int t = cls == 0 ? tg : tg - a.ns;
    using XT = typename std::conditional<F32, int32_t, double>::type;
    XT xi[3];
    load_target(a, cls, t, xi);
    double xd[3];
    particle_pos(a, tg, xd);
    const int nx = cell_coord(xd[0], inv_h) + lane % 3 - 1;
    const int ny = cell_coord(xd[1], inv_h) + (lane / 3) % 3 - 1;
    const int nz = cell_coord(xd[2], inv_h) + lane / 9 - 1;
    const int body = body_of[tg];
    const double ai = cls == 0 ? a.a_seg : a.a_blob;
    const double lam = F32 ? a.lam : 1.0;
    const uint32_t b = cell_hash(nx, ny, nz, mask);
    for (int q = start[b]; q < start[b + 1]; ++q) {
        const int j = order[q];
        if (body_of[j] != body) continue;
        double xj[3];
        particle_pos(a, j, xj);
        // the bucket may also hold other cells: visit j from its own cell only
        if (cell_coord(xj[0], inv_h) != nx || cell_coord(xj[1], inv_h) != ny || cell_coord(xj[2], inv_h) != nz)
            continue;
        const int jcls = j < a.ns ? 0 : 1;
        T rec[REC];
        stage_source<T, 3>(rec, a, jcls, jcls == 0 ? j : j - a.ns, false);
        const double aj = jcls == 0 ? a.a_seg : a.a_blob;
        const T sum_a2 = T(lam * lam * ((ai + aj) * (ai + aj)));
        T ddx, ddy, ddz;
        const T r2 = separation(xi, rec, ddx, ddy, ddz);
        if (!(r2 > sum_a2)) {
            const int k = atomicAdd(cnt + tg, 1);
            if (k < cap) idx[(int64_t)tg * cap + k] = j;
        }
    }
}

// All pairs, for problems of a few thousand particles: one wave per 64
// targets x 64 sources sets bit k of mask[c n + t] when source 64 c + k
// overlaps target t, classified exactly as near_grid_kernel does (the pair
// kernel's staging and separation, the same r^2 bound); body_of null: one
// body.  Workgroup (0, 0) clears max_cnt for near_emit_kernel.
template <typename T>
__global__ __launch_bounds__(64) void near_pairs_kernel(MatvecArgs a, int n, const int32_t* __restrict__ body_of,
                                                       uint64_t* __restrict__ mask, int32_t* __restrict__ max_cnt) {
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int REC = rec_len<T, 3>();
    __shared__ __align__(16) T recs[64 * REC];
    __shared__ int32_t sbody[64];
    __shared__ int32_t scls[64];
    const int lane = (int)threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0 && lane == 0) *max_cnt = 0;
    const int j0 = (int)blockIdx.y * 64;
    const int jl = j0 + lane;
    if (jl < n) {
        const int jcls = jl < a.ns ? 0 : 1;
        stage_source<T, 3>(recs + lane * REC, a, jcls, jcls == 0 ? jl : jl - a.ns, false);
        sbody[lane] = body_of ? body_of[jl] : 0;
        scls[lane] = jcls;
    }
    __syncthreads();
    const int tg = (int)blockIdx.x * 64 + lane;
    if (tg >= n) return;
    const int cls = tg < a.ns ? 0 : 1;
    using XT = typename std::conditional<F32, int32_t, double>::type;
    XT xi[3];
    load_target(a, cls, cls == 0 ? tg : tg - a.ns, xi);
    const int body = body_of ? body_of[tg] : 0;
    const double ai = cls == 0 ? a.a_seg : a.a_blob;
    const double lam = F32 ? a.lam : 1.0;
    const T sum_seg = T(lam * lam * ((ai + a.a_seg) * (ai + a.a_seg)));
    const T sum_blob = T(lam * lam * ((ai + a.a_blob) * (ai + a.a_blob)));
    const int m = n - j0 < 64 ? n - j0 : 64;
    uint64_t bits = 0;
#pragma unroll 8
    for (int k = 0; k < m; ++k) {
        T ddx, ddy, ddz;
        const T r2 = separation(xi, recs + k * REC, ddx, ddy, ddz);
        const bool near = !(r2 > (scls[k] == 0 ? sum_seg : sum_blob)) && sbody[k] == body;
        bits |= near ? uint64_t(1) << k : uint64_t(0);
    }
    mask[(int64_t)blockIdx.y * n + tg] = bits;
}

// One wave per target: lane q takes the masks of chunks q, q + 64, ...; an
// exclusive scan of the lanes' bit counts places each lane's sources, so the
// list comes out in increasing source order (entries past cap dropped, cnt
// the full count, max_cnt the largest count past cap).
__global__ __launch_bounds__(256) void near_emit_kernel(const uint64_t* __restrict__ mask, int n, int nchunk,
                                                       int32_t* __restrict__ cnt, int32_t* __restrict__ idx, int cap,
                                                       int32_t* __restrict__ max_cnt) {
    const int tg = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    if (tg >= n) return;  // uniform over the wave
    int32_t* e = idx + (int64_t)tg * cap;
    int base = 0;
    for (int q0 = 0; q0 < nchunk; q0 += 64) {
        const int q = q0 + lane;
        uint64_t b = q < nchunk ? mask[(int64_t)q * n + tg] : uint64_t(0);
        const int pc = __popcll(b);
        int incl = pc;  // inclusive scan over the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            incl += lane >= off ? t : 0;
        }
        int c = base + incl - pc;
        while (b) {
            const int k = __builtin_ctzll(b);
            b &= b - 1;
            if (c < cap) e[c] = q * 64 + k;
            ++c;
        }
        base += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
        cnt[tg] = base;
        if (base > cap) atomicMax(max_cnt, base);
    }
}

__global__ void fill_int_kernel(int32_t* p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Barrier (short-range repulsion) forces of the elastic model
// (cuda_functions.cu:1240-1436).  Per particle p, over every other particle j
// within 1.1 (ai + aj) -- except, for segments, j itself and its neighbours
// on the same filament --
//   f_p += S fac^4 (x_p - x_j)/|x_p - x_j|,  fac = min(1, 1 - 10 (r - ai - aj)/(ai + aj)),
// and for segments below z = DL/2 with the plane wall a normal force
// S min(1, 1 - (z - a)/(DL/2 - a))^4 first.  Segment forces are added to
// slots 0-2 of fs (stride 6); blob forces overwrite fb (stride 3, may be
// null).  Candidates come from the 27 neighbouring cells of side
// >= 1.1 * 2 max(a), in a fixed order (cells, then indices within a bucket),
// so results are reproducible; each is visited from its own cell only.
__global__ __launch_bounds__(256) void barrier_kernel(MatvecArgs a, int n, double inv_h, uint32_t mask,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ order, BarrierParams bp,
                                                     double* __restrict__ fs, double* __restrict__ fb) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    double xi[3];
    particle_pos(a, p, xi);
    const bool seg = p < a.ns;
    const double ai = seg ? a.a_seg : a.a_blob;
    const int fili = seg ? p / bp.nseg : -1;
    double f[3] = {0.0, 0.0, 0.0};
    if (seg && bp.wall && xi[2] < 0.5 * bp.dl) {
        double fz = fmin(1.0, 1.0 - (xi[2] - a.a_seg) / (0.5 * bp.dl - a.a_seg));
        fz *= bp.strength * fz * fz * fz;
        f[2] = fz;
    }
    const int cx = cell_coord(xi[0], inv_h), cy = cell_coord(xi[1], inv_h), cz = cell_coord(xi[2], inv_h);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
                const uint32_t b = cell_hash(nx, ny, nz, mask);
                for (int q = start[b]; q < start[b + 1]; ++q) {
                    const int j = order[q];
                    if (j == p) continue;
                    const bool jseg = j < a.ns;
                    if (seg && jseg && j / bp.nseg == fili && (j - p <= 1 && p - j <= 1)) continue;
                    double xj[3];
                    particle_pos(a, j, xj);
                    if (cell_coord(xj[0], inv_h) != nx || cell_coord(xj[1], inv_h) != ny ||
                        cell_coord(xj[2], inv_h) != nz)
                        continue;
                    const double asum = ai + (jseg ? a.a_seg : a.a_blob);
                    const double ddx = xi[0] - xj[0], ddy = xi[1] - xj[1], ddz = xi[2] - xj[2];
                    const double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    if (dist < 1.1 * asum) {
                        double fac = fmin(1.0, 1.0 - (10.0 / asum) * (dist - asum));
                        fac *= bp.strength * fac * fac * fac;
                        const double dm1 = 1.0 / dist;
                        f[0] += fac * ddx * dm1;
                        f[1] += fac * ddy * dm1;
                        f[2] += fac * ddz * dm1;
                    }
                }
            }
    if (seg) {
#pragma unroll
        for (int k = 0; k < 3; ++k) fs[(int64_t)p * 6 + k] += f[k];
    } else if (fb) {
#pragma unroll
        for (int k = 0; k < 3; ++k) fb[(int64_t)(p - a.ns) * 3 + k] = f[k];
    }
}

template <int VT>
__device__ void add_near(double* acc, int src, const double* xi, double ai, const MatvecArgs& a, int seg_dof,
                         uint32_t cls_mask) {
    const bool seg = src < a.ns;
    if (!(cls_mask & (seg ? 1u : 2u))) return;  // that block is not active
    const int s = seg ? src : src - a.ns;
    const CView xv = seg ? a.xs : a.xb;
    const CView fv = seg ? a.fs : a.fb;
    const double d0 = xi[0] - xv.p[(int64_t)s * xv.ld_p];
    const double d1 = xi[1] - xv.p[(int64_t)s * xv.ld_p + xv.ld_c];
    const double d2 = xi[2] - xv.p[(int64_t)s * xv.ld_p + 2 * xv.ld_c];
    double f[6];
    const int fs = seg ? seg_dof : 3;
    for (int k = 0; k < 6; ++k) f[k] = k < fs ? fv.p[(int64_t)s * fv.ld_p + k * fv.ld_c] : 0.0;
    const double aj = seg ? a.a_seg : a.a_blob;
    // Take out what the pair loop added for this pair -- the far-field form at
    // r clamped to ai + aj, with the strengths scaled by 1/(8 pi mu) -- then add
    // the reference's overlap / inner-branch value.
    const double PI = 3.14159265358979323846;
    const double c8 = 1.0 / (8.0 * PI * a.mu);
    const double s3 = (ai * ai + aj * aj) / 3.0;
    double fsc[6], clamp[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 6; ++k) fsc[k] = c8 * f[k];
    const double rm1c = rsqrt_refined((ai + aj) * (ai + aj));
    if (fs == 6) pair_far_scaled<double, VT, 6>(clamp, d0, d1, d2, rm1c, fsc, nullptr, s3, -3.0 * s3);
    else pair_far_scaled<double, VT, 3>(clamp, d0, d1, d2, rm1c, fsc, nullptr, s3, -3.0 * s3);
    for (int k = 0; k < VT; ++k) acc[k] -= clamp[k];
    if (fs == 6) pair_near<VT, 6>(acc, d0, d1, d2, f, ai, aj, a.mu);
    else pair_near<VT, 3>(acc, d0, d1, d2, f, ai, aj, a.mu);
}

// Filament rows of the saddle-point operator: -K_phi^T lambda and
// -K_theta^T lambda over the filament's 3*nseg entries, by one wave.
__device__ void saddle_filament_rows(const SaddleEpilogue& e, int m, int lane) {
    const int len = 3 * e.nseg;
    const int64_t seg0 = (int64_t)m * len;
    double sp = 0.0, st = 0.0;
    for (int k = lane; k < len; k += 64) {
        const double lam = e.in[seg0 + k];
        sp += lam * e.kphi[seg0 + k];
        st += lam * e.kth[seg0 + k];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_down(sp, off, 64);
        st += __shfl_down(st, off, 64);
    }
    if (lane == 0) {
        if (e.dyn_phase) e.out[e.off_phase + m] = -sp;
        if (e.dyn_angle) e.out[e.off_angle + m] = -st;
    }
}

// Adds the partial slots of each tile and the overlapping pairs, and applies
// the write rule.  A slot holds slot_dof x tile_targets values (component k
// of target l at k * tile_targets + l).  A workgroup covers 64 targets of one tile with four groups
// of 64 threads: group g sums the items g, g+4, g+8, ... (loads of four items
// issued together) and then the near-list entries g, g+4, ... of its target
// (the lists are sorted when built); the four group sums are added in group
// order, so the result is the same on every run.  Workgroups past
// `tile_blocks` write the filament rows of the saddle-point operator.
template <int SEGDOF, typename P>
__global__ __launch_bounds__(256) void rpy_reduce_kernel(const MatvecTile* __restrict__ tiles,
                                                         const P* __restrict__ partial, int tile_targets,
                                                         int tile_blocks, int merge_k, int slot_dof, OutArgs o,
                                                         MatvecArgs a, NearList near) {
    __shared__ double grp_acc[3][6][64];
    if (a.skip && *a.skip) return;
    const int grp = threadIdx.x >> 6;
    if ((int)blockIdx.x >= tile_blocks) {
        const int m = ((int)blockIdx.x - tile_blocks) * 4 + grp;
        if (m < o.sad.nfil && !(a.body_skip && o.sad.fil_body && a.body_skip[o.sad.fil_body[m]]))
            saddle_filament_rows(o.sad, m, threadIdx.x & 63);
        return;
    }
    const int per_tile = (tile_targets + 63) / 64;
    const MatvecTile tl = tiles[blockIdx.x / per_tile];
    if (a.body_skip && a.body_skip[tl.body]) return;  // the whole workgroup has this tile
    const int l = (blockIdx.x % per_tile) * 64 + (threadIdx.x & 63);
    const int t = tl.t_begin + l;
    const bool active = l < tile_targets && t < tl.t_end;
    const bool seg_t = tl.cls == 0;
    const int vt = seg_t ? SEGDOF : 3;
    const size_t slot_len = (size_t)slot_dof * tile_targets;
    const int lc = l < tile_targets ? l : tile_targets - 1;
    const P* p = partial + (size_t)tl.first_item * slot_len + lc;
    // Filled slots: every item's, or (persistent form, merge_k items per
    // workgroup) the tile's first item and the multiples of merge_k after it.
    const int q0 = merge_k > 0 ? tl.first_item / merge_k : 0;
    const int nslots = merge_k > 0 && tl.n_items > 0 ? (tl.first_item + tl.n_items - 1) / merge_k - q0 + 1
                                                     : tl.n_items;
    auto slot_of = [&](int q) { return merge_k > 0 ? (q == 0 ? 0 : (q0 + q) * merge_k - tl.first_item) : q; };
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int m0 = grp; m0 < nslots; m0 += 16) {
        double v[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int m = m0 + 4 * u;
            const int mm = slot_of(m < nslots ? m : grp);
#pragma unroll
            for (int k = 0; k < 6; ++k)
                v[u][k] = k < vt ? double(p[(size_t)mm * slot_len + (size_t)k * tile_targets]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = m0 + 4 * u < nslots;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] += in ? v[u][k] : 0.0;
        }
    }
    // Active source classes for this target class: bit 0 segments, bit 1 blobs.
    const uint32_t cls_mask = seg_t ? ((o.blocks & 1u) ? 1u : 0u) | ((o.blocks & 2u) ? 2u : 0u)
                                    : ((o.blocks & 8u) ? 1u : 0u) | ((o.blocks & 4u) ? 2u : 0u);
    if (near.cnt && cls_mask && active) {
        const int tg = (seg_t ? 0 : a.ns) + t;
        const int c = near.cnt[tg] < near.cap ? near.cnt[tg] : near.cap;
        const int32_t* e = near.idx + (int64_t)tg * near.cap;
        const CView xv = seg_t ? a.xs : a.xb;
        const double xi[3] = {xv.p[(int64_t)t * xv.ld_p], xv.p[(int64_t)t * xv.ld_p + xv.ld_c],
                              xv.p[(int64_t)t * xv.ld_p + 2 * xv.ld_c]};
        const double ai = seg_t ? a.a_seg : a.a_blob;
        for (int q = grp; q < c; q += 4) {
            if (seg_t && SEGDOF == 6) add_near<6>(acc, e[q], xi, ai, a, SEGDOF, cls_mask);
            else add_near<3>(acc, e[q], xi, ai, a, SEGDOF, cls_mask);
        }
    }
    if (grp > 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) grp_acc[grp - 1][k][threadIdx.x & 63] = acc[k];
    }
    __syncthreads();
    if (grp > 0 || !active) return;
#pragma unroll
    for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += grp_acc[g][k][threadIdx.x];
    if (seg_t && o.sad.in && (o.sad.dyn_phase || o.sad.dyn_angle)) {
        const int m = t / o.sad.nseg;
        const double pd = o.sad.dyn_phase ? o.sad.in[o.sad.off_phase + m] : 0.0;
        const double td = o.sad.dyn_angle ? o.sad.in[o.sad.off_angle + m] : 0.0;
        for (int k = 0; k < 3; ++k) acc[k] -= pd * o.sad.kphi[3 * (int64_t)t + k] + td * o.sad.kth[3 * (int64_t)t + k];
    }
    if (o.sad.free_body) {
        // -(V + Omega x r) of the target's swimmer (mobility_solver.cpp:1856-1870, 1885-1894)
        const int sw = seg_t ? t / o.sad.segs_per_swim : t / o.sad.blobs_per_swim;
        const double* V = o.sad.in + o.sad.off_body + 6 * (int64_t)sw;
        const CView xv = seg_t ? a.xs : a.xb;
        const double r0 = xv.p[(int64_t)t * xv.ld_p] - o.sad.x_bodies[3 * sw];
        const double r1 = xv.p[(int64_t)t * xv.ld_p + xv.ld_c] - o.sad.x_bodies[3 * sw + 1];
        const double r2 = xv.p[(int64_t)t * xv.ld_p + 2 * xv.ld_c] - o.sad.x_bodies[3 * sw + 2];
        acc[0] -= V[0] + (V[4] * r2 - V[5] * r1);
        acc[1] -= V[1] + (V[5] * r0 - V[3] * r2);
        acc[2] -= V[2] + (V[3] * r1 - V[4] * r0);
    }
    const int slots = seg_t ? o.seg_out_slots : 3;
    const View out = seg_t ? o.vs : o.vb;
    for (int k = 0; k < slots; ++k) {
        double* dst = out.p + (int64_t)t * out.ld_p + k * out.ld_c;
        if (tl.write_mode == 0) *dst = k < vt ? acc[k] : 0.0;
        else if (k < vt) *dst += acc[k];
    }
}

// Sorts every target's near list into increasing source order, 32 lanes per
// target: each lane ranks up to kRankPer entries against the whole list (the
// entries are distinct source indices), then all write after a barrier.
// Lists up to 32 kRankPer long; *max_cnt gets any count above cap.
constexpr int kRankPer = 8;
__global__ __launch_bounds__(256) void near_rank_sort_kernel(const int32_t* __restrict__ cnt,
                                                             int32_t* __restrict__ idx, int cap, int64_t n,
                                                             int32_t* __restrict__ max_cnt) {
    const int64_t tg = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);
    const int lane = (int)(threadIdx.x & 31);
    int c = 0;
    int32_t* e = idx;
    if (tg < n) {
        const int k = cnt[tg];
        c = k < cap ? k : cap;
        e = idx + tg * cap;
        if (lane == 0 && k > cap) atomicMax(max_cnt, k);
    }
    int32_t val[kRankPer];
    int rank[kRankPer];
#pragma unroll
    for (int s = 0; s < kRankPer; ++s) {
        const int i = lane + 32 * s;
        val[s] = 0;
        rank[s] = -1;
        if (i < c) {
            const int32_t v = e[i];
            int r = 0;
            for (int j = 0; j < c; ++j) r += e[j] < v ? 1 : 0;
            val[s] = v;
            rank[s] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kRankPer; ++s)
        if (rank[s] >= 0) e[rank[s]] = val[s];
}

// The same for longer lists (insertion sort, one thread per target).
__global__ __launch_bounds__(256) void near_sort_kernel(const int32_t* __restrict__ cnt, int32_t* __restrict__ idx,
                                                        int cap, int64_t n, int32_t* __restrict__ max_cnt) {
    const int64_t tg = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (tg >= n) return;
    if (cnt[tg] > cap) atomicMax(max_cnt, cnt[tg]);
    const int c = cnt[tg] < cap ? cnt[tg] : cap;
    int32_t* e = idx + tg * cap;
    for (int i = 1; i < c; ++i) {
        const int32_t v = e[i];
        int j = i - 1;
        while (j >= 0 && e[j] > v) {
            e[j + 1] = e[j];
            --j;
        }
        e[j + 1] = v;
    }
}

// Items per workgroup of the persistent form: the work list over
// persistent x CUs workgroups, as contiguous ranges.
int persistent_items(const TileShape& sh, int n_items, int num_cus) {
    const long g = (long)sh.persistent * (num_cus > 0 ? num_cus : 256);
    return (int)std::max<long>(1, (n_items + g - 1) / g);
}

template <typename T, int SEGDOF, int BLOCK, int R, bool SPLIT>
hipError_t launch_pairs(const TileShape& sh, const MatvecArgs& a, const MatvecItem* d_items, int n_items,
                        const MatvecTile* d_tiles, double* partial, const PersistState& ps, hipStream_t stream) {
    if (SPLIT && sh.persistent > 0) return hipErrorInvalidValue;  // no persistent form of the split
    if (!SPLIT && sh.persistent > 0) {
        const int k = persistent_items(sh, n_items, ps.num_cus);
        hipLaunchKernelGGL((rpy_pairs_persistent_kernel<T, SEGDOF, BLOCK, R>), dim3((unsigned)((n_items + k - 1) / k)),
                           dim3(BLOCK), 0, stream, d_items, d_tiles, a, reinterpret_cast<T*>(partial), n_items, k,
                           ps);
    } else {
        hipLaunchKernelGGL((rpy_pairs_kernel<T, SEGDOF, BLOCK, R, SPLIT>), dim3(n_items), dim3(BLOCK), 0, stream,
                           d_items, d_tiles, a, reinterpret_cast<T*>(partial), ps.item_stamps);
    }
    return hipGetLastError();
}

template <typename T, int SEGDOF>
hipError_t dispatch_shape(const TileShape& sh, const MatvecArgs& a, const MatvecItem* d_items, int n_items,
                          const MatvecTile* d_tiles, double* partial, const PersistState& ps,
                          hipStream_t stream) {
#define TFS_SHAPE(B, R, S)                                                                    \
    if (sh.block == B && sh.per_thread == R && sh.split == S)                                \
        return launch_pairs<T, SEGDOF, B, R, S>(sh, a, d_items, n_items, d_tiles, partial, ps, stream);
    TFS_SHAPE(64, 1, 0) TFS_SHAPE(64, 2, 0) TFS_SHAPE(128, 1, 0) TFS_SHAPE(128, 2, 0) TFS_SHAPE(256, 1, 0)
    TFS_SHAPE(256, 2, 0) TFS_SHAPE(256, 1, 1) TFS_SHAPE(256, 2, 1) TFS_SHAPE(512, 1, 1) TFS_SHAPE(512, 2, 1)
#undef TFS_SHAPE
    return hipErrorInvalidValue;
}

}  // namespace

void build_worklist(const std::vector<BodyRange>& bodies, int tile_targets, int chunk, bool ss, bool sb,
                    bool bb, bool bs, int seg_write, int blob_write, std::vector<MatvecTile>& tiles,
                    std::vector<MatvecItem>& items) {
    tiles.clear();
    items.clear();
    // Target tiles of every body; with several bodies the tiles are then
    // interleaved in proportion to each body's tile count (tile k of T_b at
    // position (k + 1/2) / T_b), so that every body's work is spread over the
    // whole launch and the bodies finish together.
    struct Pending {
        MatvecTile tl;
        double pos;
    };
    std::vector<Pending> pend;
    for (const BodyRange& body : bodies) {
        const size_t first = pend.size();
        auto add_tiles = [&](int cls, int t_lo, int t_hi, int write) {
            if (write < 0) return;
            for (int t0 = t_lo; t0 < t_hi; t0 += tile_targets) {
                MatvecTile tl{};
                tl.cls = cls;
                tl.t_begin = t0;
                tl.t_end = std::min(t0 + tile_targets, t_hi);
                tl.write_mode = write;
                tl.body = (int32_t)(&body - bodies.data());
                pend.push_back({tl, 0.0});
            }
        };
        add_tiles(0, body.seg_begin, body.seg_end, seg_write);
        add_tiles(1, body.blob_begin, body.blob_end, blob_write);
        const size_t nt = pend.size() - first;
        for (size_t k = 0; k < nt; ++k) pend[first + k].pos = (double(k) + 0.5) / double(nt);
    }
    if (bodies.size() > 1)
        std::stable_sort(pend.begin(), pend.end(), [](const Pending& x, const Pending& y) { return x.pos < y.pos; });
    for (Pending& pt : pend) {
        MatvecTile tl = pt.tl;
        const BodyRange& body = bodies[tl.body];
        const bool from_segs = tl.cls == 0 ? ss : bs, from_blobs = tl.cls == 0 ? sb : bb;
        tl.first_item = (int)items.size();
        const int tile_id = (int)tiles.size();
        std::vector<MatvecItem> chunks;  // the tile's items in source order
        auto add_src = [&](int src_cls, int lo, int hi) {
            for (int s0 = lo; s0 < hi; s0 += chunk) chunks.push_back({tile_id, src_cls, s0, std::min(s0 + chunk, hi)});
        };
        if (from_segs) add_src(0, body.seg_begin, body.seg_end);
        if (from_blobs) add_src(1, body.blob_begin, body.blob_end);
        // Workgroup i of a launch runs on XCD i mod 8.  Chunk c of every tile
        // goes to a position p with p = c (mod 8) where one is free, so each
        // XCD's L2 holds about an eighth of the sources instead of all of them.
        const int n = (int)chunks.size();
        std::vector<int> next(kXcds, 0), order(n, -1);
        std::vector<char> used(n, 0);
        for (int i = 0; i < n; ++i) {
            const int want = (tl.first_item + i) % kXcds;
            int& c = next[want];
            while (c * kXcds + want < n && used[c * kXcds + want]) ++c;
            if (c * kXcds + want < n) {
                order[i] = c * kXcds + want;
                used[order[i]] = 1;
            }
        }
        for (int i = 0, c = 0; i < n; ++i) {
            if (order[i] >= 0) continue;
            while (used[c]) ++c;
            order[i] = c;
            used[c] = 1;
        }
        for (int i = 0; i < n; ++i) items.push_back(chunks[order[i]]);
        tl.n_items = n;
        tiles.push_back(tl);
    }
}

size_t build_sym_worklist(const std::vector<BodyRange>& bodies, int seg_write, int blob_write,
                          std::vector<MatvecTile>& tiles, std::vector<MatvecItem>& items) {
    tiles.clear();
    items.clear();
    for (size_t b = 0; b < bodies.size(); ++b) {
        auto add_tiles = [&](int cls, int lo, int hi, int write) {
            for (int t0 = lo; t0 < hi; t0 += 64) {
                MatvecTile tl{};
                tl.cls = cls;
                tl.t_begin = t0;
                tl.t_end = std::min(t0 + 64, hi);
                tl.write_mode = write;
                tl.body = (int32_t)b;
                tiles.push_back(tl);
            }
        };
        add_tiles(0, bodies[b].seg_begin, bodies[b].seg_end, seg_write);
        add_tiles(1, bodies[b].blob_begin, bodies[b].blob_end, blob_write);
    }
    const int n_tiles = (int)tiles.size();
    const int n_groups = (n_tiles + kSymTiles - 1) / kSymTiles;
    // each body's groups [gfirst, glast]; each group's bodies [blo, bhi]
    std::vector<int> gfirst(bodies.size(), -1), glast(bodies.size(), -1), blo(n_groups), bhi(n_groups);
    for (int t = 0; t < n_tiles; ++t) {
        const int b = tiles[t].body, g = t / kSymTiles;
        if (gfirst[b] < 0) gfirst[b] = g;
        glast[b] = g;
        if (t % kSymTiles == 0) blo[g] = b;
        bhi[g] = b;
    }
    size_t slots = 0;
    for (MatvecTile& tl : tiles) {
        tl.group0 = gfirst[tl.body];
        tl.first_item = (int32_t)slots;
        tl.n_items = glast[tl.body] - gfirst[tl.body] + 1;
        slots += (size_t)tl.n_items;
    }
    // off-diagonal items (both tiles' sums, the costlier ones) first; groups
    // are in body order, so ga < gb share a body when gb starts no later than
    // the last body of ga
    for (int ga = 0; ga < n_groups; ++ga)
        for (int gb = ga + 1; gb < n_groups && blo[gb] <= bhi[ga]; ++gb) items.push_back({ga, gb, 0, 0});
    for (int g = 0; g < n_groups; ++g) items.push_back({g, g, 0, 0});
    return slots;
}

void free_grid(NearGrid& g) {
    for (int32_t* p : {g.start, g.fill, g.bucket_of, g.order, g.body_of, g.pair_body})
        if (p) (void)hipFree(p);
    if (g.mask) (void)hipFree(g.mask);
    const bool may_alloc = g.may_alloc;
    g = NearGrid{};
    g.may_alloc = may_alloc;
}

// Sorts the n particles (segments, then blobs) into the grid's buckets for
// cells of side 1 / inv_h: start (exclusive bucket offsets) and order.
static hipError_t sort_into_grid(const MatvecArgs& a, NearGrid& grid, int n, double inv_h, hipStream_t stream) {
    int table = 1024;
    while (table < 2 * n) table <<= 1;
    if (grid.n != n || grid.table != table) {
        free_grid(grid);
        TFS_HIP_CHECK(hipMalloc(&grid.start, sizeof(int32_t) * (table + 1)));
        TFS_HIP_CHECK(hipMalloc(&grid.fill, sizeof(int32_t) * table));
        TFS_HIP_CHECK(hipMalloc(&grid.bucket_of, sizeof(int32_t) * n));
        TFS_HIP_CHECK(hipMalloc(&grid.order, sizeof(int32_t) * n));
        TFS_HIP_CHECK(hipMalloc(&grid.body_of, sizeof(int32_t) * n));
        grid.n = n;
        grid.table = table;
    }
    const uint32_t mask = (uint32_t)table - 1;
    const unsigned gn = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((table + 255) / 256)), dim3(256), 0, stream, grid.fill, 0,
                       (int64_t)table);
    hipLaunchKernelGGL(grid_count_kernel, dim3(gn), dim3(256), 0, stream, a, n, inv_h, mask, grid.bucket_of, grid.fill);
    hipLaunchKernelGGL(grid_scan_kernel, dim3(1), dim3(1024), 0, stream, grid.fill, grid.start, table);
    hipLaunchKernelGGL(grid_scatter_kernel, dim3(gn), dim3(256), 0, stream, n, grid.bucket_of, grid.start, grid.fill,
                       grid.order);
    hipLaunchKernelGGL(grid_sort_buckets_kernel, dim3((unsigned)((table + 255) / 256)), dim3(256), 0, stream, table,
                       grid.start, grid.order);
    return hipGetLastError();
}

hipError_t build_near_list(int precision, const MatvecArgs& a, NearGrid& grid, const int32_t* body_ranges, int nbody,
                           NearList& near, int n_targets, int* max_count, int* host_scratch, hipStream_t stream) {
    const int n = n_targets;
    if (max_count) *max_count = 0;
    if (n <= 0) return hipSuccess;
    const int nchunk = (n + 63) / 64;
    if (n <= kNearPairsMax && (grid.may_alloc || grid.mask_len >= (int64_t)nchunk * n)) {
        if (grid.mask_len < (int64_t)nchunk * n) {
            TFS_HIP_CHECK(hipStreamSynchronize(stream));
            if (grid.mask) (void)hipFree(grid.mask);
            grid.mask = nullptr;
            grid.mask_len = 0;
            TFS_HIP_CHECK(hipMalloc(&grid.mask, sizeof(uint64_t) * (size_t)nchunk * n));
            grid.mask_len = (int64_t)nchunk * n;
        }
        const int32_t* body_of = nullptr;
        if (nbody > 1) {
            if (grid.pair_body_len < n) {
                TFS_HIP_CHECK(hipStreamSynchronize(stream));
                if (grid.pair_body) (void)hipFree(grid.pair_body);
                grid.pair_body = nullptr;
                grid.pair_body_len = 0;
                TFS_HIP_CHECK(hipMalloc(&grid.pair_body, sizeof(int32_t) * n));
                grid.pair_body_len = n;
            }
            hipLaunchKernelGGL(grid_body_kernel, dim3((unsigned)nbody), dim3(256), 0, stream,
                               reinterpret_cast<const int4*>(body_ranges), a.ns, grid.pair_body);
            body_of = grid.pair_body;
        }
        const dim3 g2((unsigned)nchunk, (unsigned)nchunk);
        if (precision == 0)
            hipLaunchKernelGGL(near_pairs_kernel<double>, g2, dim3(64), 0, stream, a, n, body_of, grid.mask,
                               near.max_cnt);
        else
            hipLaunchKernelGGL(near_pairs_kernel<float>, g2, dim3(64), 0, stream, a, n, body_of, grid.mask,
                               near.max_cnt);
        hipLaunchKernelGGL(near_emit_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, stream,
                           (const uint64_t*)grid.mask, n, nchunk, near.cnt, near.idx, near.cap, near.max_cnt);
        TFS_HIP_CHECK(hipGetLastError());
        if (!host_scratch) return hipSuccess;  // no capacity check wanted
        TFS_HIP_CHECK(hipMemcpyAsync(host_scratch, near.max_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        if (!max_count) return hipSuccess;
        TFS_HIP_CHECK(hipStreamSynchronize(stream));
        *max_count = *host_scratch;
        return hipSuccess;
    }
    const double h = 2.0 * std::max(a.a_seg, a.a_blob) * (1.0 + 1e-5);
    const double inv_h = 1.0 / h;
    TFS_HIP_CHECK(sort_into_grid(a, grid, n, inv_h, stream));
    const uint32_t mask = (uint32_t)grid.table - 1;
    const unsigned gn = (unsigned)((n + 255) / 256);
    const unsigned gl = (unsigned)((n + 7) / 8);   // 8 targets of 32 lanes per workgroup
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, stream, near.max_cnt, 0, (int64_t)1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(gn), dim3(256), 0, stream, near.cnt, 0, (int64_t)n);
    hipLaunchKernelGGL(grid_body_kernel, dim3((unsigned)nbody), dim3(256), 0, stream,
                       reinterpret_cast<const int4*>(body_ranges), a.ns, grid.body_of);
    if (precision == 0)
        hipLaunchKernelGGL(near_grid_kernel<double>, dim3(gl), dim3(256), 0, stream, a, n, inv_h, mask, grid.start,
                           grid.order, grid.body_of, near.cnt, near.idx, near.cap);
    else
        hipLaunchKernelGGL(near_grid_kernel<float>, dim3(gl), dim3(256), 0, stream, a, n, inv_h, mask, grid.start,
                           grid.order, grid.body_of, near.cnt, near.idx, near.cap);
    TFS_HIP_CHECK(hipGetLastError());
    if (near.cap <= 32 * kRankPer)
        hipLaunchKernelGGL(near_rank_sort_kernel, dim3(gl), dim3(256), 0, stream, near.cnt, near.idx, near.cap,
                           (int64_t)n, near.max_cnt);
    else
        hipLaunchKernelGGL(near_sort_kernel, dim3(gn), dim3(256), 0, stream, near.cnt, near.idx, near.cap,
                           (int64_t)n, near.max_cnt);
    TFS_HIP_CHECK(hipGetLastError());
    if (!host_scratch) return hipSuccess;  // no capacity check wanted
    TFS_HIP_CHECK(hipMemcpyAsync(host_scratch, near.max_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    if (!max_count) return hipSuccess;  // the caller reads host_scratch after its own wait
    TFS_HIP_CHECK(hipStreamSynchronize(stream));
    *max_count = *host_scratch;
    return hipSuccess;
}

hipError_t launch_barrier(const MatvecArgs& a, NearGrid& grid, const BarrierParams& bp, double* fs, double* fb,
                          hipStream_t stream) {
    const int n = a.ns + a.nb;
    if (n <= 0) return hipSuccess;
    const double h = 2.2 * std::max(a.a_seg, a.a_blob) * (1.0 + 1e-6);
    const double inv_h = 1.0 / h;
    TFS_HIP_CHECK(sort_into_grid(a, grid, n, inv_h, stream));
    hipLaunchKernelGGL(barrier_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a, n, inv_h,
                       (uint32_t)grid.table - 1, grid.start, grid.order, bp, fs, fb);
    return hipGetLastError();
}

hipError_t launch_matvec_wall(int precision, int seg_dof, const MatvecArgs& a, const OutArgs& o,
                              const MatvecTile* d_tiles, int n_tiles, const MatvecItem* d_items, int n_items,
                              double* partial, hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end) {
    if (n_tiles == 0) return hipSuccess;
    float* pf = reinterpret_cast<float*>(partial);
    if (n_items > 0) {
        if (ev_begin) TFS_HIP_CHECK(hipEventRecord(ev_begin, stream));
        const dim3 grid((unsigned)n_items), block(kWallBlock);
        if (precision == 0) {
            if (seg_dof == 6)
                hipLaunchKernelGGL((rpy_pairs_wall_kernel<6, double>), grid, block, 0, stream, d_items, d_tiles, a,
                                   partial);
            else
                hipLaunchKernelGGL((rpy_pairs_wall_kernel<3, double>), grid, block, 0, stream, d_items, d_tiles, a,
                                   partial);
        } else {
            if (seg_dof == 6)
                hipLaunchKernelGGL((rpy_pairs_wall_kernel<6, float>), grid, block, 0, stream, d_items, d_tiles, a, pf);
            else
                hipLaunchKernelGGL((rpy_pairs_wall_kernel<3, float>), grid, block, 0, stream, d_items, d_tiles, a, pf);
        }
        TFS_HIP_CHECK(hipGetLastError());
        if (ev_end) TFS_HIP_CHECK(hipEventRecord(ev_end, stream));
    }
    const int tile_blocks = n_tiles * ((kWallBlock + 63) / 64);
    const int fil_blocks = o.sad.in ? (o.sad.nfil + 3) / 4 : 0;
    const NearList none{};
    const dim3 grid((unsigned)(tile_blocks + fil_blocks));
    if (precision == 0) {
        if (seg_dof == 6)
            hipLaunchKernelGGL((rpy_reduce_kernel<6, double>), grid, dim3(256), 0, stream, d_tiles,
                               (const double*)partial, kWallBlock, tile_blocks, 0, kMaxDof, o, a, none);
        else
            hipLaunchKernelGGL((rpy_reduce_kernel<3, double>), grid, dim3(256), 0, stream, d_tiles,
                               (const double*)partial, kWallBlock, tile_blocks, 0, kMaxDof, o, a, none);
    } else {
        if (seg_dof == 6)
            hipLaunchKernelGGL((rpy_reduce_kernel<6, float>), grid, dim3(256), 0, stream, d_tiles, (const float*)pf,
                               kWallBlock, tile_blocks, 0, kMaxDof, o, a, none);
        else
            hipLaunchKernelGGL((rpy_reduce_kernel<3, float>), grid, dim3(256), 0, stream, d_tiles, (const float*)pf,
                               kWallBlock, tile_blocks, 0, kMaxDof, o, a, none);
    }
    return hipGetLastError();
}

hipError_t launch_matvec(int precision, int seg_dof, const TileShape& shape, const MatvecArgs& a,
                         const OutArgs& o, const MatvecTile* d_tiles, int n_tiles,
                         const MatvecItem* d_items, int n_items, double* partial, const NearList& near,
                         const PersistState& ps, hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end) {
    if (n_tiles == 0) return hipSuccess;
    if (shape.symmetric && (seg_dof != 3 || (precision == 1 && !a.xfx))) return hipErrorInvalidValue;
    if (n_items > 0) {
        hipError_t e;
        if (ev_begin) TFS_HIP_CHECK(hipEventRecord(ev_begin, stream));
        if (shape.symmetric) {
            const int bb = (o.blocks & 4u) ? 1 : 0;
            if (precision == 1)
                hipLaunchKernelGGL(rpy_pairs_sym_kernel<float>, dim3((unsigned)n_items), dim3(64 * kSymTiles), 0, stream,
                                   d_items, d_tiles, n_tiles, bb, a, reinterpret_cast<float*>(partial), ps.item_stamps);
            else
                hipLaunchKernelGGL(rpy_pairs_sym_kernel<double>, dim3((unsigned)n_items), dim3(64 * kSymTiles), 0,
                                   stream, d_items, d_tiles, n_tiles, bb, a, partial, ps.item_stamps);
            e = hipGetLastError();
        } else if (precision == 0) {
            e = seg_dof == 6 ? dispatch_shape<double, 6>(shape, a, d_items, n_items, d_tiles, partial, ps, stream)
                             : dispatch_shape<double, 3>(shape, a, d_items, n_items, d_tiles, partial, ps, stream);
        } else {
            e = seg_dof == 6 ? dispatch_shape<float, 6>(shape, a, d_items, n_items, d_tiles, partial, ps, stream)
                             : dispatch_shape<float, 3>(shape, a, d_items, n_items, d_tiles, partial, ps, stream);
        }
        if (e != hipSuccess) return e;
        if (ev_end) TFS_HIP_CHECK(hipEventRecord(ev_end, stream));
    }
    const int tt = tile_targets(shape);
    const int tile_blocks = n_tiles * ((tt + 63) / 64);
    const int fil_blocks = o.sad.in ? (o.sad.nfil + 3) / 4 : 0;
    const dim3 grid((unsigned)(tile_blocks + fil_blocks));
    // The partial slots hold the evaluation type (FP32 sums for FP32 tiles);
    // the persistent form fills only its tile segments' first slots.
    const int merge_k = shape.persistent > 0 && n_items > 0 ? persistent_items(shape, n_items, ps.num_cus) : 0;
    const int slot_dof = shape.symmetric ? kSymSlotDof : kMaxDof;
    const float* pf = reinterpret_cast<const float*>(partial);
    if (precision == 0) {
        if (seg_dof == 6)
            hipLaunchKernelGGL((rpy_reduce_kernel<6, double>), grid, dim3(256), 0, stream, d_tiles,
                               (const double*)partial, tt, tile_blocks, merge_k, slot_dof, o, a, near);
        else
            hipLaunchKernelGGL((rpy_reduce_kernel<3, double>), grid, dim3(256), 0, stream, d_tiles,
                               (const double*)partial, tt, tile_blocks, merge_k, slot_dof, o, a, near);
    } else {
        if (seg_dof == 6)
            hipLaunchKernelGGL((rpy_reduce_kernel<6, float>), grid, dim3(256), 0, stream, d_tiles, pf, tt,
                               tile_blocks, merge_k, slot_dof, o, a, near);
        else
            hipLaunchKernelGGL((rpy_reduce_kernel<3, float>), grid, dim3(256), 0, stream, d_tiles, pf, tt,
                               tile_blocks, merge_k, slot_dof, o, a, near);
    }
    return hipGetLastError();
}

}  // namespace tfs
