// tfs_rpy_api.hip -- the C-ABI of include/tfs_rpy.h over the gfx950 kernels.
//
// Mirrors the data ownership of rpy_mobility_solver
// (src/cilia/mobility/rpy_mobility_solver.cu:54-117): the handle owns one set
// of device arrays in the reference's AoS layouts (positions 3, forces 6/3,
// velocities 6/3 per particle).  Unlike the reference there is no per-GPU
// target split and no host staging between kernels: the four blocks run as one
// fused launch (plus a reduce) on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "tfs_api_internal.h"

namespace tfs {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(TFS_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

int hip_fail_if(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }

// s.chunk for ~per_cu items per CU over a body of ns segments and nb blobs
// (a multiple of the block).
static void size_chunk(TileShape& s, int ns, int nb, int num_cus, long per_cu) {
    const int tt = tile_targets(s);
    const long n_tiles = (ns + tt - 1) / tt + (nb + tt - 1) / tt;
    const long n_src = (long)ns + nb;
    const long want_items = per_cu * (num_cus > 0 ? num_cus : 256);
    long chunk = n_tiles > 0 ? (n_src * n_tiles + want_items - 1) / want_items : n_src;
    chunk = ((chunk + s.block - 1) / s.block) * s.block;
    if (chunk < s.block) chunk = s.block;
    s.chunk = (int)chunk;
}

TileShape default_shape(int precision, int ns, int nb, int num_cus, int seg_dof, int num_bodies) {
    // Measured on MI355X (pair kernel, ms per launch): one body of 12,800
    // segments + 9,000 blobs, 3-DOF segments, is fastest with the source
    // split and 64-target tiles in FP32 (0.266 against 0.279 for the best
    // plain shape) and 128-target tiles in FP64 (0.430 against 0.444), with
    // items sized for ~48 and ~32 per CU; the 6-DOF matvec of that body
    // likewise (FP32 0.391 with 64-target tiles, 0.400 with 128; FP64 0.716,
    // against 0.418 and 0.792 without the split).  64 uneven bodies in one
    // launch have work to spare: there the split with 128-target tiles and
    // ~512-source items is within 2% of the best plain shape in both
    // precisions, and 64-target tiles lose 5-10%.
    // A body too small to fill the device at that shape (under ~2 items per
    // CU, e.g. 320 segments + 2,000 blobs: 190 items) takes the finest shape,
    // one wave per item of 64 targets x 32 sources (the LDS batch then holds
    // half a wave's worth): its launches are bound by the longest item's
    // latency, not by throughput.
    (void)seg_dof;
    TileShape s;
    s.block = 256;
    s.persistent = 0;
    s.split = 1;
    s.per_thread = num_bodies == 1 && precision == 1 ? 1 : 2;
    const long per_cu = num_bodies > 1 ? 64 : s.per_thread == 1 ? 48 : 32;
    size_chunk(s, ns, nb, num_cus, per_cu);
    const int tt = tile_targets(s);
    const long items = ((long)(ns + tt - 1) / tt + (nb + tt - 1) / tt) *
                       (((long)ns + s.chunk - 1) / s.chunk + ((long)nb + s.chunk - 1) / s.chunk);
    if (items < 2L * (num_cus > 0 ? num_cus : 256)) s = TileShape{64, 1, 32, 0, 0};
    return s;
}

// The default launch shape of a handle: the symmetric form for 3-DOF
// segments when the (largest) body alone gives every CU ~4 items (measured on
// MI355X, 12,800 segments + 9,000 blobs: 0.212 ms per FP32 launch against
// 0.266 for the best source-split shape, 0.358 ms per FP64 launch against
// 0.432; 64 uneven bodies in lock-step, a mixed timestep in 831 ms against
// 1,027), else default_shape.
// `sym_slots`: the symmetric form's partial slots over all bodies (its buffer
// takes 3 x 64 doubles each), which must stay below kSymMaxPartialBytes.
static TileShape preferred_shape(int precision, int ns, int nb, int num_cus, int seg_dof, int num_bodies,
                                 double sym_slots) {
    if (seg_dof == 3 && sym_slots * kSymSlotDof * 64 * sizeof(double) <= kSymMaxPartialBytes) {
        const long tiles = (ns + 63L) / 64 + (nb + 63L) / 64;
        const long groups = (tiles + 3) / 4;
        if (groups * (groups + 1) / 2 >= 4L * (num_cus > 0 ? num_cus : 256)) {
            TileShape s{256, 1, 0, 0, 0};
            s.symmetric = 1;
            return s;
        }
    }
    return default_shape(precision, ns, nb, num_cus, seg_dof, num_bodies);
}

// Without the split: 512-target tiles, ~16 items per CU (short items even
// out the tail).
static TileShape plain_shape(int ns, int nb, int num_cus) {
    TileShape s{256, 2, 256, 0, 0};
    size_chunk(s, ns, nb, num_cus, 16);
    return s;
}

// The largest body's segment and blob counts: with independent bodies, work
// items are sized so that this body alone fills the device (a body whose
// neighbours have converged still runs at full width).
static void largest_body(const tfs_rpy* h, int* ns, int* nb) {
    *ns = h->desc.num_segs;
    *nb = h->desc.num_blobs;
    if (h->bodies.size() < 2) return;
    long best = -1;
    for (const BodyRange& b : h->bodies) {
        const long n = (long)(b.seg_end - b.seg_begin) + (b.blob_end - b.blob_begin);
        if (n > best) {
            best = n;
            *ns = b.seg_end - b.seg_begin;
            *nb = b.blob_end - b.blob_begin;
        }
    }
}

// The symmetric form's partial slots for these bodies (each tile one per
// group of four tiles its body overlaps).
static double sym_slot_count(const std::vector<BodyRange>& bodies) {
    double slots = 0.0;
    for (const BodyRange& b : bodies) {
        const double t = double((b.seg_end - b.seg_begin + 63) / 64 + (b.blob_end - b.blob_begin + 63) / 64);
        slots += t * (std::ceil(t / kSymGroupTiles) + 1.0);
    }
    return slots;
}

Plan* get_plan(tfs_rpy* h, uint32_t blocks, const TileShape& sh, int* err) {
    bool ss = blocks & TFS_BLOCK_SS, sb = blocks & TFS_BLOCK_SB;
    bool bb = blocks & TFS_BLOCK_BB, bs = blocks & TFS_BLOCK_BS;
    int seg_write = -1, blob_write = -1;
    // Sequential semantics: with an overwriting BS, an earlier BB is overwritten.
    if (bs && bb && h->desc.bs_mode == TFS_BS_OVERWRITE) bb = false;
    seg_write = ss ? 0 : (sb ? 1 : -1);
    if (bb) blob_write = 0;
    else if (bs) blob_write = h->desc.bs_mode == TFS_BS_OVERWRITE ? 0 : 1;
    const uint32_t eff = (ss ? 1u : 0u) | (sb ? 2u : 0u) | (bb ? 4u : 0u) | (bs ? 8u : 0u);
    const auto key = std::make_tuple(eff, sh.block, sh.per_thread, sh.chunk, sh.split, sh.symmetric);
    auto it = h->plans.find(key);
    if (it != h->plans.end()) return &it->second;
    Plan p;
    p.blocks = eff;
    size_t slots = 0;
    if (sh.symmetric) {
        slots = build_sym_worklist(h->bodies, seg_write, blob_write, p.tiles, p.items);
        p.symmetric = true;
    } else {
        build_worklist(h->bodies, tile_targets(sh), sh.chunk, ss, sb, bb, bs, seg_write, blob_write, p.tiles, p.items);
        slots = p.items.size();
    }
    if (!p.tiles.empty()) {
        hipError_t e = hipMalloc(&p.d_tiles, sizeof(MatvecTile) * p.tiles.size());
        if (e == hipSuccess && !p.items.empty())
            e = hipMalloc(&p.d_items, sizeof(MatvecItem) * p.items.size());
        if (e == hipSuccess)
            e = hipMemcpy(p.d_tiles, p.tiles.data(), sizeof(MatvecTile) * p.tiles.size(),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess && !p.items.empty())
            e = hipMemcpy(p.d_items, p.items.data(), sizeof(MatvecItem) * p.items.size(),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(p.d_tiles);
            hipFree(p.d_items);
            *err = hip_fail(e, "work list upload");
            return nullptr;
        }
        const size_t need = slots * (size_t)(sh.symmetric ? kSymSlotDof : 6) * (size_t)tile_targets(sh);
        if (need > h->partial_len) {
            if (h->partial) {
                hipStreamSynchronize(h->stream);
                hipFree(h->partial);
            }
            h->partial = nullptr;
            h->partial_len = 0;
            e = hipMalloc(&h->partial, need * sizeof(double));
            if (e != hipSuccess) {
                *err = hip_fail(e, "partial-sum buffer");
                return nullptr;
            }
            h->partial_len = need;
        }
    }
    auto res = h->plans.emplace(key, std::move(p));
    return &res.first->second;
}

bool symmetric_applies(const tfs_rpy* h, int precision, uint32_t blocks) {
    (void)precision;
    return h->desc.seg_dof == 3 && h->domain == TFS_DOMAIN_UNBOUNDED && (blocks & TFS_BLOCK_SS) &&
           (blocks & TFS_BLOCK_SB) && (blocks & TFS_BLOCK_BS);
}

TileShape effective_shape(const tfs_rpy* h, int precision, uint32_t blocks) {
    const TileShape& sh = h->shape[precision];
    if (!sh.symmetric || symmetric_applies(h, precision, blocks)) return sh;
    int ns = 0, nb = 0;
    largest_body(h, &ns, &nb);
    return default_shape(precision, ns, nb, h->num_cus, h->desc.seg_dof, (int)h->bodies.size());
}

void invalidate_near(tfs_rpy* h) {
    h->near_valid[0] = h->near_valid[1] = false;
    h->fx_valid = false;
}

namespace {

// Bounding box of each body's segments and blobs (min xyz, max xyz); a
// non-finite coordinate makes the box NaN.  One workgroup per body.
__global__ __launch_bounds__(256) void body_box_kernel(const double* __restrict__ xs, const double* __restrict__ xb,
                                                       const int4* __restrict__ ranges, double* __restrict__ box) {
    __shared__ double red[6][256];
    __shared__ int bad_any;
    const int4 r = ranges[blockIdx.x];
    const int tid = threadIdx.x;
    if (tid == 0) bad_any = 0;
    double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    bool bad = false;
    for (int i = r.x + tid; i < r.y; i += 256)
        for (int k = 0; k < 3; ++k) {
            const double v = xs[3 * (int64_t)i + k];
            bad |= !__builtin_isfinite(v);
            lo[k] = __builtin_fmin(lo[k], v);
            hi[k] = __builtin_fmax(hi[k], v);
        }
    for (int i = r.z + tid; i < r.w; i += 256)
        for (int k = 0; k < 3; ++k) {
            const double v = xb[3 * (int64_t)i + k];
            bad |= !__builtin_isfinite(v);
            lo[k] = __builtin_fmin(lo[k], v);
            hi[k] = __builtin_fmax(hi[k], v);
        }
    for (int k = 0; k < 3; ++k) {
        red[k][tid] = lo[k];
        red[3 + k][tid] = hi[k];
    }
    __syncthreads();
    if (bad) bad_any = 1;
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < 3; ++k) {
                red[k][tid] = __builtin_fmin(red[k][tid], red[k][tid + w]);
                red[3 + k][tid] = __builtin_fmax(red[3 + k][tid], red[3 + k][tid + w]);
            }
        __syncthreads();
    }
    if (tid < 6) box[6 * blockIdx.x + tid] = bad_any ? __builtin_nan("") : red[tid][0];
}

// X = rint(lambda (x - c_body)) for every segment and blob of body blockIdx.x.
__global__ __launch_bounds__(256) void to_fixed_kernel(const double* __restrict__ xs, const double* __restrict__ xb,
                                                       const int4* __restrict__ ranges,
                                                       const double* __restrict__ centre, double lam, int ns,
                                                       int4* __restrict__ out) {
    const int4 r = ranges[blockIdx.x];
    const double c0 = centre[3 * blockIdx.x], c1 = centre[3 * blockIdx.x + 1], c2 = centre[3 * blockIdx.x + 2];
    for (int i = r.x + threadIdx.x; i < r.y; i += 256) {
        const double* x = xs + 3 * (int64_t)i;
        out[i] = make_int4((int)__builtin_rint((x[0] - c0) * lam), (int)__builtin_rint((x[1] - c1) * lam),
                           (int)__builtin_rint((x[2] - c2) * lam), 0);
    }
    for (int i = r.z + threadIdx.x; i < r.w; i += 256) {
        const double* x = xb + 3 * (int64_t)i;
        out[ns + i] = make_int4((int)__builtin_rint((x[0] - c0) * lam), (int)__builtin_rint((x[1] - c1) * lam),
                                (int)__builtin_rint((x[2] - c2) * lam), 0);
    }
}

}  // namespace

// Fixed-point coordinates for FP32 tiles.  Each body is centred on its
// bounding box; lambda = 2^k is the largest power of two with
// lambda * (largest half-extent) <= 2^30 - 2, so every coordinate and every
// difference of two fits an int32.  The quantisation step 1/lambda must stay
// below 2^-19 of the smallest radius (a body within ~1024 radii of its
// centre); larger bodies are refused in FP32.  Synchronises the stream.
// The body table on the device (BodyRange records, plus room for 6 doubles
// per body), uploaded on the handle's stream from h->bodies.
int ensure_body_table(tfs_rpy* h) {
    if (h->body_table_valid) return TFS_OK;
    const size_t nbody = h->bodies.size();
    hipError_t e = hipSuccess;
    if (h->body_cap < nbody) {
        if (h->d_body_box) (void)hipFree(h->d_body_box);
        if (h->d_body_range) (void)hipFree(h->d_body_range);
        h->d_body_box = nullptr;
        h->d_body_range = nullptr;
        e = hipMalloc(&h->d_body_box, sizeof(double) * 6 * nbody);
        if (e == hipSuccess) e = hipMalloc(&h->d_body_range, sizeof(int4) * nbody);
        h->body_cap = e == hipSuccess ? nbody : 0;
    }
    if (e != hipSuccess) return fail(TFS_ERR_OUT_OF_MEMORY, "body table");
    e = hipMemcpyAsync(h->d_body_range, h->bodies.data(), sizeof(int4) * nbody, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);  // the host table may change after return
    if (e != hipSuccess) return hip_fail(e, "body table");
    h->body_table_valid = true;
    return TFS_OK;
}

// The fixed-point frame of the bodies' bounding boxes (6 per body: min xyz,
// max xyz): each body's centre and one lambda = 2^k for all of them.
static int fixed_point_frame(const double* box, size_t nbody, double amin, double* centre, double* lam) {
    double ext = 0.0;
    for (size_t b = 0; b < nbody; ++b) {
        for (int k = 0; k < 3; ++k) {
            const double lo = box[6 * b + k], hi = box[6 * b + 3 + k];
            if (std::isnan(lo) || std::isnan(hi)) return fail(TFS_ERR_NON_FINITE, "positions are not finite");
            centre[3 * b + k] = 0.0;
            if (!(lo <= hi)) continue;  // body without particles
            const double c = 0.5 * (lo + hi);
            centre[3 * b + k] = c;
            ext = std::max(ext, std::max(hi - c, c - lo));
        }
    }
    const double lim = 1073741822.0;  // 2^30 - 2
    int k = 60;
    if (ext > 0.0) {
        k = (int)std::floor(std::log2(lim / ext));
        while (k > -1000 && std::ldexp(ext, k) > lim) --k;
        while (k < 60 && std::ldexp(ext, k + 1) <= lim) ++k;
        k = std::min(k, 60);
    }
    if (std::ldexp(1.0, -k) > std::ldexp(amin, -19))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "FP32 tiles: a body extends more than about 1024 particle radii from its centre (use FP64)");
    *lam = std::ldexp(1.0, k);
    return TFS_OK;
}

int ensure_fixed(tfs_rpy* h) {
    if (h->fx_valid) return TFS_OK;
    const int ns = h->desc.num_segs, nb = h->desc.num_blobs;
    const size_t nbody = h->bodies.size();
    hipError_t e = hipSuccess;
    if (!h->x_fx) e = hipMalloc(&h->x_fx, sizeof(int4) * (size_t)std::max(ns + nb, 1));
    if (e != hipSuccess) return fail(TFS_ERR_OUT_OF_MEMORY, "fixed-point positions");
    int rc = ensure_body_table(h);
    if (rc) return rc;
    std::vector<double> box(6 * nbody), centre(3 * nbody, 0.0);
    {
        hipLaunchKernelGGL(body_box_kernel, dim3((unsigned)nbody), dim3(256), 0, h->stream, h->x_segs, h->x_blobs,
                           (const int4*)h->d_body_range, h->d_body_box);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(box.data(), h->d_body_box, sizeof(double) * 6 * nbody, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "fixed-point positions: bounding boxes");
    int rc2 = fixed_point_frame(box.data(), nbody, std::min(h->desc.seg_radius, h->desc.blob_radius), centre.data(),
                                &h->fx_lambda);
    if (rc2) return rc2;
    h->fx_centre = centre;
    e = hipMemcpyAsync(h->d_body_box, h->fx_centre.data(), sizeof(double) * 3 * nbody, hipMemcpyHostToDevice,
                       h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(to_fixed_kernel, dim3((unsigned)nbody), dim3(256), 0, h->stream, h->x_segs, h->x_blobs,
                           (const int4*)h->d_body_range, (const double*)h->d_body_box, h->fx_lambda, ns, h->x_fx);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "fixed-point positions");
    h->fx_valid = true;
    return TFS_OK;
}

void free_near(NearList& nl) {
    hipFree(nl.cnt);
    hipFree(nl.idx);
    hipFree(nl.max_cnt);
    nl = NearList{};
}

MatvecArgs matvec_args(tfs_rpy* h, CView fs, CView fb) {
    MatvecArgs a;
    a.xs = {h->x_segs, 3, 1};
    a.xb = {h->x_blobs, 3, 1};
    a.fs = fs;
    a.fb = fb;
    a.ns = h->desc.num_segs;
    a.nb = h->desc.num_blobs;
    a.a_seg = h->desc.seg_radius;
    a.a_blob = h->desc.blob_radius;
    a.mu = h->desc.mu;
    a.skip = h->skip_flag;
    a.body_skip = h->body_skip;
    a.xfx = h->x_fx;
    a.lam = h->fx_lambda;
    return a;
}

// Lists the overlapping pairs of the current positions for `precision`,
// growing the per-target capacity until every list fits.
int ensure_near(tfs_rpy* h, int precision) {
    if (h->near_valid[precision] && (precision == 0 || h->fx_valid)) return TFS_OK;
    const int nt = h->desc.num_segs + h->desc.num_blobs;
    if (nt == 0) {
        h->near_valid[precision] = true;
        return TFS_OK;
    }
    int err = precision == 1 ? ensure_fixed(h) : ensure_body_table(h);
    if (err) return err;
    NearList& nl = h->near[precision];
    MatvecArgs a = matvec_args(h, CView{h->f_segs, 6, 1}, CView{h->f_blobs, 3, 1});
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (!nl.cnt) {
            if (nl.cap == 0) nl.cap = 32;
            hipError_t e = hipMalloc(&nl.cnt, sizeof(int32_t) * nt);
            if (e == hipSuccess) e = hipMalloc(&nl.idx, sizeof(int32_t) * (size_t)nt * nl.cap);
            if (e == hipSuccess) e = hipMalloc(&nl.max_cnt, sizeof(int32_t));
            if (e != hipSuccess) {
                free_near(nl);
                return fail(TFS_ERR_OUT_OF_MEMORY, "near-pair lists");
            }
        }
        if (h->defer_near_check && nl.cnt) {
            // a capacity of every particle cannot overflow: nothing to check
            const bool roomy = nl.cap >= nt;
            hipError_t e = build_near_list(precision, a, h->grid, h->d_body_range, (int)h->bodies.size(), nl, nt,
                                           nullptr, roomy ? nullptr : h->host_int + 8 + precision, h->stream);
            if (e != hipSuccess) return hip_fail(e, "near-pair lists");
            h->near_valid[precision] = true;
            h->near_pending[precision] = !roomy;
            return TFS_OK;
        }
        int max_count = 0;
        hipError_t e = build_near_list(precision, a, h->grid, h->d_body_range, (int)h->bodies.size(), nl, nt,
                                       &max_count, h->host_int, h->stream);
        if (e != hipSuccess) return hip_fail(e, "near-pair lists");
        if (max_count <= nl.cap) {
            h->near_valid[precision] = true;
            return TFS_OK;
        }
        int cap = nl.cap;
        while (cap < max_count) cap *= 2;
        hipStreamSynchronize(h->stream);
        free_near(nl);
        nl.cap = cap;
    }
    return fail(TFS_ERR_INVALID_ARGUMENT, "near-pair lists: too many overlapping neighbours");
}

int rpy_near_reserve(tfs_rpy* h, int precision, int cap) {
    NearList& nl = h->near[precision];
    if (nl.cap >= cap) return TFS_OK;
    if (nl.cnt && hipStreamSynchronize(h->stream) != hipSuccess) return fail(TFS_ERR_HIP, "near-pair lists: synchronise");
    free_near(nl);
    nl.cap = cap;
    h->near_valid[precision] = false;
    return TFS_OK;
}

int rpy_near_check(tfs_rpy* h, hipStream_t waited, bool* overflow) {
    *overflow = false;
    for (int p = 0; p < 2; ++p) {
        if (!h->near_pending[p]) continue;
        if (h->stream != waited && hipStreamSynchronize(h->stream) != hipSuccess)
            return fail(TFS_ERR_HIP, "near-pair lists: synchronise");
        h->near_pending[p] = false;
        NearList& nl = h->near[p];
        const int m = h->host_int[8 + p];
        if (m <= nl.cap) continue;
        int cap = nl.cap;
        while (cap < m) cap *= 2;
        hipStreamSynchronize(h->stream);
        free_near(nl);
        nl.cap = cap;
        h->near_valid[p] = false;
        *overflow = true;
    }
    return TFS_OK;
}

int ensure_stamps(tfs_rpy* h, size_t n_items) {
    if (h->item_stamps_len >= 2 * n_items) return TFS_OK;
    if (h->d_item_stamps) {
        hipStreamSynchronize(h->stream);
        hipFree(h->d_item_stamps);
    }
    h->d_item_stamps = nullptr;
    h->item_stamps_len = 0;
    hipError_t e = hipMalloc(&h->d_item_stamps, sizeof(int64_t) * 2 * n_items);
    if (e != hipSuccess) return hip_fail(e, "persistent-launch stamps");
    h->item_stamps_len = 2 * n_items;
    return TFS_OK;
}

void free_persist(tfs_rpy* h) {
    hipFree(h->d_item_stamps);
    h->d_item_stamps = nullptr;
    h->item_stamps_len = 0;
    h->stamped_plan = nullptr;
}

void free_plans(tfs_rpy* h) {
    for (auto& kv : h->plans) {
        hipFree(kv.second.d_tiles);
        hipFree(kv.second.d_items);
    }
    h->plans.clear();
}

int run_matvec(tfs_rpy* h, uint32_t blocks, int precision, CView fs, CView fb, View vs, View vb,
               int seg_out_slots, const SaddleEpilogue* sad) {
    const bool wall = h->domain == TFS_DOMAIN_HALF_SPACE;
    TileShape sh = effective_shape(h, precision, blocks);
    if (wall) {
        // one target per lane, ~16 workgroups per CU over the work list
        const int ns = h->desc.num_segs, nb = h->desc.num_blobs;
        const long n_tiles = (ns + kWallBlock - 1) / kWallBlock + (nb + kWallBlock - 1) / kWallBlock;
        const long want = 16L * (h->num_cus > 0 ? h->num_cus : 256);
        long chunk = n_tiles > 0 ? ((long)(ns + nb) * n_tiles + want - 1) / want : (long)(ns + nb);
        chunk = std::max<long>(kWallBlock, (chunk + kWallBlock - 1) / kWallBlock * kWallBlock);
        sh = TileShape{kWallBlock, 1, (int)chunk, 0};
    }
    int err = TFS_OK;
    Plan* p = get_plan(h, blocks, sh, &err);
    if (!p) return err;
    if (p->tiles.empty()) return TFS_OK;
    if (!wall) err = ensure_near(h, precision);
    else if (precision == TFS_PRECISION_FP32_ACC64) err = ensure_fixed(h);  // separations in fixed point
    if (err) return err;
    MatvecArgs a = matvec_args(h, fs, fb);
    OutArgs o;
    o.vs = vs;
    o.vb = vb;
    o.seg_out_slots = seg_out_slots;
    o.blocks = p->blocks;
    if (sad) o.sad = *sad;
    hipEvent_t eb = nullptr, ee = nullptr;
    if (h->profiling && !p->items.empty()) {
        if (h->ev_free.empty()) {
            hipEvent_t a0, a1;
            hipError_t ce = hipEventCreate(&a0);
            if (ce == hipSuccess) ce = hipEventCreate(&a1);
            if (ce != hipSuccess) return hip_fail(ce, "profiling events");
            h->ev_free.push_back({a0, a1, 0});
        }
        tfs_rpy::EvPair ep = h->ev_free.back();
        h->ev_free.pop_back();
        ep.precision = precision;
        eb = ep.begin;
        ee = ep.end;
        h->ev_used.push_back(ep);
    }
    PersistState ps;
    ps.num_cus = h->num_cus;
    if (h->body_timing) {
        int rc = ensure_stamps(h, p->items.size());
        if (rc) return rc;
        ps.item_stamps = h->d_item_stamps;
        h->stamped_plan = p;
    }
    hipError_t e = wall ? launch_matvec_wall(precision, h->desc.seg_dof, a, o, p->d_tiles, (int)p->tiles.size(), p->d_items,
                                             (int)p->items.size(), h->partial, h->stream, eb, ee)
                        : launch_matvec(precision, h->desc.seg_dof, sh, a, o, p->d_tiles, (int)p->tiles.size(),
                                        p->d_items, (int)p->items.size(), h->partial, h->near[precision], ps,
                                        h->stream, eb, ee);
    if (e != hipSuccess) return hip_fail(e, "matvec launch");
    return TFS_OK;
}

}  // namespace tfs

using namespace tfs;

extern "C" {

const char* tfs_last_error(void) { return g_last_error.c_str(); }

const char* tfs_version(void) { return "tfs-rpy-mi355x 0.1 (gfx950)"; }

int tfs_rpy_create(const tfs_rpy_desc* desc, tfs_rpy** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: null argument");
    *out = nullptr;
    if (desc->num_segs < 0 || desc->num_blobs < 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: negative particle count");
    if (desc->seg_dof != 3 && desc->seg_dof != 6)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: seg_dof must be 3 or 6");
    if (desc->precision != TFS_PRECISION_FP64 && desc->precision != TFS_PRECISION_FP32_ACC64)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: unknown precision");
    if (desc->bs_mode != TFS_BS_OVERWRITE && desc->bs_mode != TFS_BS_ACCUMULATE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: unknown bs_mode");
    if (desc->memory != TFS_MEMORY_HOST && desc->memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: unknown memory space");
    if (!(desc->seg_radius > 0) || !(desc->blob_radius > 0) || !(desc->mu > 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_create: radii and mu must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TFS_ERR_NO_DEVICE, "tfs_rpy_create: no HIP device visible");
    tfs_rpy* h = new tfs_rpy();
    h->grid.may_alloc = true;
    h->desc = *desc;
    hipGetDevice(&h->device);
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) cus = prop.multiProcessorCount;
    h->num_cus = cus;
    if (desc->stream) {
        h->stream = (hipStream_t)desc->stream;
        h->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete h;
            return hip_fail(e, "tfs_rpy_create: stream");
        }
        h->own_stream = true;
    }
    const size_t ns = desc->num_segs, nb = desc->num_blobs;
    // One allocation for the six AoS arrays (at least one element each).
    const size_t lens[6] = {3 * ns, 6 * ns, 3 * nb, 3 * nb, 6 * ns, 3 * nb};
    size_t total = 0;
    for (size_t l : lens) total += std::max<size_t>(l, 1) + 32;  // keep 256-B alignment
    hipError_t e = hipMalloc(&h->arena, total * sizeof(double));
    if (e != hipSuccess) {
        tfs_rpy_destroy(h);
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_rpy_create: device arrays");
    }
    double* p = h->arena;
    double** dst[6] = {&h->x_segs, &h->f_segs, &h->x_blobs, &h->f_blobs, &h->v_segs, &h->v_blobs};
    for (int i = 0; i < 6; ++i) {
        *dst[i] = p;
        p += ((std::max<size_t>(lens[i], 1) + 31) / 32) * 32;
    }
    e = hipMemsetAsync(h->arena, 0, total * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        tfs_rpy_destroy(h);
        return hip_fail(e, "tfs_rpy_create: clear");
    }
    for (int pr = 0; pr < 2; ++pr)
        h->shape[pr] = preferred_shape(pr, desc->num_segs, desc->num_blobs, cus, desc->seg_dof, 1,
                                       sym_slot_count({BodyRange{0, desc->num_segs, 0, desc->num_blobs}}));
    h->bodies.push_back(BodyRange{0, desc->num_segs, 0, desc->num_blobs});
    e = hipHostMalloc((void**)&h->host_int, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        tfs_rpy_destroy(h);
        return hip_fail(e, "tfs_rpy_create: pinned scratch");
    }
    *out = h;
    return TFS_OK;
}

int tfs_rpy_destroy(tfs_rpy* h) {
    if (!h) return TFS_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (auto& kv : h->plans) {
        hipFree(kv.second.d_tiles);
        hipFree(kv.second.d_items);
    }
    for (auto* v : {&h->ev_used, &h->ev_free})
        for (auto& pr : *v) {
            hipEventDestroy(pr.begin);
            hipEventDestroy(pr.end);
        }
    if (h->partial) hipFree(h->partial);
    if (h->arena) hipFree(h->arena);
    for (auto& nl : h->near) free_near(nl);
    free_grid(h->grid);
    if (h->x_fx) hipFree(h->x_fx);
    if (h->d_body_box) hipFree(h->d_body_box);
    if (h->d_body_range) hipFree(h->d_body_range);
    if (h->host_int) hipHostFree(h->host_int);
    if (h->blob_rep) hipFree(h->blob_rep);
    free_stage(h->f32);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
    delete h;
    return TFS_OK;
}

static int array_info(tfs_rpy* h, int32_t what, double** dev, size_t* len) {
    const size_t ns = h->desc.num_segs, nb = h->desc.num_blobs;
    switch (what) {
        case TFS_SEG_POSITIONS: *dev = h->x_segs; *len = 3 * ns; return TFS_OK;
        case TFS_SEG_FORCES: *dev = h->f_segs; *len = 6 * ns; return TFS_OK;
        case TFS_BLOB_POSITIONS: *dev = h->x_blobs; *len = 3 * nb; return TFS_OK;
        case TFS_BLOB_FORCES: *dev = h->f_blobs; *len = 3 * nb; return TFS_OK;
        case TFS_SEG_VELOCITIES: *dev = h->v_segs; *len = 6 * ns; return TFS_OK;
        case TFS_BLOB_VELOCITIES: *dev = h->v_blobs; *len = 3 * nb; return TFS_OK;
        default: return fail(TFS_ERR_INVALID_ARGUMENT, "unknown array id");
    }
}

int tfs_rpy_upload(tfs_rpy* h, int32_t what, const double* src) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload: null handle");
    if (what > TFS_BLOB_FORCES && what != TFS_SEG_VELOCITIES && what != TFS_BLOB_VELOCITIES)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload: not an input array");
    double* dev;
    size_t len;
    int rc = array_info(h, what, &dev, &len);
    if (rc) return rc;
    if (len == 0) return TFS_OK;
    if (!src) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload: null source");
    if (what == TFS_SEG_POSITIONS || what == TFS_BLOB_POSITIONS) invalidate_near(h);
    hipError_t e = hipMemcpyAsync(dev, src, len * sizeof(double),
                                  h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice
                                                                    : hipMemcpyDeviceToDevice,
                                  h->stream);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_rpy_upload");
}

int tfs_rpy_download(tfs_rpy* h, int32_t what, double* dst) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download: null handle");
    if (what != TFS_SEG_VELOCITIES && what != TFS_BLOB_VELOCITIES)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download: not an output array");
    double* dev;
    size_t len;
    int rc = array_info(h, what, &dev, &len);
    if (rc) return rc;
    if (len == 0) return TFS_OK;
    if (!dst) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download: null destination");
    hipError_t e = hipMemcpyAsync(dst, dev, len * sizeof(double),
                                  h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyDeviceToHost
                                                                    : hipMemcpyDeviceToDevice,
                                  h->stream);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_rpy_download");
}

int tfs_rpy_evaluate(tfs_rpy* h, uint32_t blocks) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_evaluate: null handle");
    if (blocks & ~TFS_BLOCK_ALL) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_evaluate: bad block mask");
    return run_matvec(h, blocks, h->desc.precision, CView{h->f_segs, 6, 1}, CView{h->f_blobs, 3, 1},
                      View{h->v_segs, 6, 1}, View{h->v_blobs, 3, 1}, 6);
}

int tfs_rpy_apply(tfs_rpy* h, uint32_t blocks, const double* f_segs, const double* f_blobs,
                  double* v_segs, double* v_blobs) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply: null handle");
    if (blocks & ~TFS_BLOCK_ALL) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply: bad block mask");
    const bool need_fs = blocks & (TFS_BLOCK_SS | TFS_BLOCK_BS);
    const bool need_fb = blocks & (TFS_BLOCK_SB | TFS_BLOCK_BB);
    const bool out_s = blocks & (TFS_BLOCK_SS | TFS_BLOCK_SB);
    const bool out_b = blocks & (TFS_BLOCK_BB | TFS_BLOCK_BS);
    const bool ns = h->desc.num_segs > 0, nb = h->desc.num_blobs > 0;
    if ((need_fs && ns && !f_segs) || (need_fb && nb && !f_blobs) || (out_s && ns && !v_segs) ||
        (out_b && nb && !v_blobs))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply: missing array for requested block");
    if (h->desc.memory == TFS_MEMORY_DEVICE) {
        return run_matvec(h, blocks, h->desc.precision, CView{f_segs ? f_segs : h->f_segs, 6, 1},
                          CView{f_blobs ? f_blobs : h->f_blobs, 3, 1}, View{v_segs ? v_segs : h->v_segs, 6, 1},
                          View{v_blobs ? v_blobs : h->v_blobs, 3, 1}, 6);
    }
    int rc = TFS_OK;
    if (need_fs && ns) rc = tfs_rpy_upload(h, TFS_SEG_FORCES, f_segs);
    if (!rc && need_fb && nb) rc = tfs_rpy_upload(h, TFS_BLOB_FORCES, f_blobs);
    // Accumulating blocks read the caller's current velocities.
    const bool acc_s = (blocks & TFS_BLOCK_SB) && !(blocks & TFS_BLOCK_SS);
    const bool acc_b = (blocks & TFS_BLOCK_BS) && !(blocks & TFS_BLOCK_BB) &&
                       h->desc.bs_mode == TFS_BS_ACCUMULATE;
    if (!rc && acc_s && ns) rc = tfs_rpy_upload(h, TFS_SEG_VELOCITIES, v_segs);
    if (!rc && acc_b && nb) rc = tfs_rpy_upload(h, TFS_BLOB_VELOCITIES, v_blobs);
    if (!rc) rc = tfs_rpy_evaluate(h, blocks);
    if (!rc && out_s && ns) rc = tfs_rpy_download(h, TFS_SEG_VELOCITIES, v_segs);
    if (!rc && out_b && nb) rc = tfs_rpy_download(h, TFS_BLOB_VELOCITIES, v_blobs);
    return rc;
}

}  // extern "C"

// Checks a barrier descriptor against the handle; fills the kernel's parameters.
static int barrier_params(tfs_rpy* h, const tfs_barrier_desc* desc, bool have_segs, BarrierParams* bp,
                          const std::string& where) {
    if (!h || !desc) return fail(TFS_ERR_INVALID_ARGUMENT, where + ": null handle or descriptor");
    const int64_t ns = h->desc.num_segs;
    if (ns > 0 && !have_segs) return fail(TFS_ERR_INVALID_ARGUMENT, where + ": null segment forces");
    if (!(desc->strength >= 0) || !std::isfinite(desc->strength))
        return fail(TFS_ERR_INVALID_ARGUMENT, where + ": strength must be finite and >= 0");
    if (desc->seg_per_fil < 1 || (ns > 0 && ns % desc->seg_per_fil != 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, where + ": seg_per_fil must divide the segment count");
    if (desc->wall && !(0.5 * desc->dl > h->desc.seg_radius))
        return fail(TFS_ERR_INVALID_ARGUMENT, where + ": the wall term needs dl/2 > segment radius");
    bp->strength = desc->strength;
    bp->nseg = desc->seg_per_fil;
    bp->wall = desc->wall ? 1 : 0;
    bp->dl = desc->dl;
    return TFS_OK;
}

extern "C" {

int tfs_rpy_barrier_forces(tfs_rpy* h, const tfs_barrier_desc* desc, double* f_segs, double* f_blobs_repulsion) {
    BarrierParams bp;
    int rc0 = barrier_params(h, desc, f_segs != nullptr, &bp, "tfs_rpy_barrier_forces");
    if (rc0) return rc0;
    const int64_t ns = h->desc.num_segs, nb = h->desc.num_blobs;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    double* fs = f_segs;
    double* fb = f_blobs_repulsion;
    hipError_t e = hipSuccess;
    if (host) {
        fs = h->f_segs;
        if (ns > 0) e = hipMemcpyAsync(fs, f_segs, 48 * ns, hipMemcpyHostToDevice, h->stream);
        if (f_blobs_repulsion && nb > 0) {
            if (!h->blob_rep && hipMalloc(&h->blob_rep, 24 * nb) != hipSuccess)
                return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_rpy_barrier_forces: device memory");
            fb = h->blob_rep;
        }
    }
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_barrier_forces: upload");
    const MatvecArgs a = matvec_args(h, CView{h->f_segs, 6, 1}, CView{h->f_blobs, 3, 1});
    e = launch_barrier(a, h->grid, bp, fs, nb > 0 ? fb : nullptr, h->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_barrier_forces");
    if (host) {
        if (ns > 0) e = hipMemcpyAsync(f_segs, fs, 48 * ns, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess && f_blobs_repulsion && nb > 0)
            e = hipMemcpyAsync(f_blobs_repulsion, fb, 24 * nb, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(e, "tfs_rpy_barrier_forces: download");
    }
    return TFS_OK;
}

int tfs_rpy_set_domain(tfs_rpy* h, int32_t domain) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_domain: null handle");
    if (domain != TFS_DOMAIN_UNBOUNDED && domain != TFS_DOMAIN_HALF_SPACE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_domain: unknown domain");
    if (domain == TFS_DOMAIN_HALF_SPACE &&
        (h->desc.seg_radius != 1.0 || h->desc.blob_radius != 1.0 || h->desc.mu != 1.0))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_rpy_set_domain: the half-space tensor is for radius 1 and viscosity 1");
    h->domain = domain;
    return TFS_OK;
}

int tfs_rpy_sync(tfs_rpy* h) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_sync: null handle");
    hipError_t e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_rpy_sync");
}

int tfs_rpy_set_precision(tfs_rpy* h, int32_t precision) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_precision: null handle");
    if (precision != TFS_PRECISION_FP64 && precision != TFS_PRECISION_FP32_ACC64)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_precision: unknown precision");
    h->desc.precision = precision;
    return TFS_OK;
}

static int set_shape(tfs_rpy* h, const char* fn, int32_t precision, int32_t block, int32_t per_thread,
                     int32_t chunk, int split) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(fn) + ": null handle");
    if (precision != 0 && precision != 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(fn) + ": unknown precision");
    if (split ? (block != 256 && block != 512) : (block != 64 && block != 128 && block != 256))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    std::string(fn) + (split ? ": block must be 256 or 512" : ": block must be 64, 128 or 256"));
    if (per_thread != 1 && per_thread != 2)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(fn) + ": per_thread must be 1 or 2");
    const bool part_batch = !split && chunk >= 32 && chunk < block && block % chunk == 0;
    if (!part_batch && (chunk < block || chunk % block))
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    std::string(fn) + ": chunk must be a positive multiple of block (or, without the split, a "
                                      "divisor of it of at least 32)");
    const int pers = h->shape[precision].persistent;
    if (split && pers > 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(fn) + ": the source split has no persistent form");
    h->shape[precision] = TileShape{block, per_thread, chunk, pers, split};
    h->shape_chosen[precision] = true;
    return TFS_OK;
}

int tfs_rpy_set_tile_shape(tfs_rpy* h, int32_t precision, int32_t block, int32_t per_thread,
                           int32_t chunk) {
    return set_shape(h, "tfs_rpy_set_tile_shape", precision, block, per_thread, chunk, 0);
}

int tfs_rpy_set_split_shape(tfs_rpy* h, int32_t precision, int32_t block, int32_t per_thread, int32_t chunk) {
    return set_shape(h, "tfs_rpy_set_split_shape", precision, block, per_thread, chunk, 1);
}

int tfs_rpy_set_symmetric(tfs_rpy* h, int32_t precision, int32_t enable) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_symmetric: null handle");
    if (precision != 0 && precision != 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_symmetric: unknown precision");
    if (!enable) {
        if (h->shape[precision].symmetric) {
            int ns = 0, nb = 0;
            largest_body(h, &ns, &nb);
            h->shape[precision] = default_shape(precision, ns, nb, h->num_cus, h->desc.seg_dof, (int)h->bodies.size());
            h->shape_chosen[precision] = true;
        }
        return TFS_OK;
    }
    if (h->desc.seg_dof != 3)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_symmetric: the symmetric form needs 3-DOF segments");
    if (h->shape[precision].persistent > 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_symmetric: the symmetric form has no persistent form");
    TileShape sh{64 * 4, 1, 0, 0, 0};
    sh.symmetric = 1;
    h->shape[precision] = sh;
    h->shape_chosen[precision] = true;
    return TFS_OK;
}

int tfs_rpy_get_symmetric(tfs_rpy* h, int32_t precision, int32_t* enabled) {
    if (!h || !enabled || (precision != 0 && precision != 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_get_symmetric: bad arguments");
    *enabled = h->shape[precision].symmetric;
    return TFS_OK;
}

int tfs_rpy_positions_changed(tfs_rpy* h) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_positions_changed: null handle");
    invalidate_near(h);
    return TFS_OK;
}

int tfs_rpy_set_bodies(tfs_rpy* h, int32_t num_bodies, const int32_t* segs_per_body,
                       const int32_t* blobs_per_body) {
    if (!h || num_bodies < 1 || !segs_per_body || !blobs_per_body)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_bodies: bad arguments");
    std::vector<BodyRange> bodies;
    int64_t s = 0, b = 0;
    for (int i = 0; i < num_bodies; ++i) {
        if (segs_per_body[i] < 0 || blobs_per_body[i] < 0)
            return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_bodies: negative count");
        bodies.push_back(BodyRange{(int32_t)s, (int32_t)(s + segs_per_body[i]), (int32_t)b,
                                   (int32_t)(b + blobs_per_body[i])});
        s += segs_per_body[i];
        b += blobs_per_body[i];
    }
    if (s != h->desc.num_segs || b != h->desc.num_blobs)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_bodies: counts do not add up to the handle's sizes");
    hipStreamSynchronize(h->stream);
    free_plans(h);
    free_persist(h);
    h->bodies = bodies;
    h->body_table_valid = false;
    invalidate_near(h);
    int lns = 0, lnb = 0;
    largest_body(h, &lns, &lnb);
    for (int pr = 0; pr < 2; ++pr)
        if (!h->shape_chosen[pr])
            h->shape[pr] = preferred_shape(pr, lns, lnb, h->num_cus, h->desc.seg_dof, (int)h->bodies.size(),
                                           sym_slot_count(h->bodies));
    return TFS_OK;
}

struct tfs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
};

int tfs_rpy_graph_create(tfs_rpy* h, uint32_t blocks, int32_t evaluations, tfs_graph** out) {
    if (!h || !out || evaluations < 1 || (blocks & ~TFS_BLOCK_ALL))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_graph_create: bad arguments");
    *out = nullptr;
    // Everything that allocates or synchronises (work list, partial buffer,
    // near-pair lists, queue state) is created by one evaluation outside capture.
    int rc = tfs_rpy_evaluate(h, blocks);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_graph_create: warm-up");
    const bool prof = h->profiling;
    h->profiling = false;
    tfs_graph* g = new tfs_graph();
    g->stream = h->stream;
    e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        h->profiling = prof;
        delete g;
        return hip_fail(e, "tfs_rpy_graph_create: begin capture");
    }
    for (int i = 0; i < evaluations && rc == TFS_OK; ++i) rc = tfs_rpy_evaluate(h, blocks);
    hipError_t e2 = hipStreamEndCapture(h->stream, &g->graph);
    h->profiling = prof;
    if (rc || e2 != hipSuccess) {
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return rc ? rc : hip_fail(e2, "tfs_rpy_graph_create: end capture");
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g->graph);
        delete g;
        return hip_fail(e, "tfs_rpy_graph_create: instantiate");
    }
    *out = g;
    return TFS_OK;
}

int tfs_graph_launch(tfs_graph* g) {
    if (!g) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_graph_launch: null graph");
    hipError_t e = hipGraphLaunch(g->exec, g->stream);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_graph_launch");
}

int tfs_graph_destroy(tfs_graph* g) {
    if (!g) return TFS_OK;
    hipStreamSynchronize(g->stream);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return TFS_OK;
}

int tfs_rpy_autotune(tfs_rpy* h, int32_t precision, uint32_t blocks, int32_t reps, double* best_ms) {
    if (!h || (precision != 0 && precision != 1) || reps < 1 || (blocks & ~TFS_BLOCK_ALL))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_autotune: bad arguments");
    const int saved_precision = h->desc.precision;
    h->desc.precision = precision;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(TFS_ERR_HIP, "tfs_rpy_autotune: events");
    const bool prof = h->profiling;
    h->profiling = false;
    TileShape best = h->shape[precision];
    double best_t = 1e300;
    int rc = TFS_OK;
    // Chunk candidates are sized on the largest body (all of it for one body),
    // whose items alone should fill the device.
    int ns = 0, nb = 0;
    largest_body(h, &ns, &nb);
    const double pairs = double(ns + nb) * double(ns + nb);
    auto time_shape = [&](const TileShape& sh) {
        h->shape[precision] = sh;
        rc = tfs_rpy_evaluate(h, blocks);  // builds the plan; warm-up
        if (rc) return;
        hipEventRecord(e0, h->stream);
        for (int i = 0; i < reps && !rc; ++i) rc = tfs_rpy_evaluate(h, blocks);
        hipEventRecord(e1, h->stream);
        if (rc) return;
        hipEventSynchronize(e1);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms / reps < best_t) {
            best_t = ms / reps;
            best = h->shape[precision];
        }
    };
    if (symmetric_applies(h, precision, blocks)) {
        TileShape sym{256, 1, 0, 0, 0};
        sym.symmetric = 1;
        time_shape(sym);
    }
    for (int cand = 0; cand < 8 && !rc; ++cand) {
        const int split = cand / 4, block = 64 << (cand % 4);
        if (split ? block < 256 : block > 256) continue;  // the split needs >= 4 waves
        if (block == 64 && pairs > 1e8) continue;  // large problems: 64-thread groups never win
        for (int r : {1, 2}) {
            TileShape probe{block, r, block, 0, split};
            const int tt = tile_targets(probe);
            // chunk candidates: the largest body's work list at ~16, 32 and 64 items per CU
            const long n_tiles = (ns + tt - 1) / tt + (nb + tt - 1) / tt;
            long prev = -1;
            for (long per_cu : {16L, 32L, 64L}) {
                const long want = per_cu * (h->num_cus > 0 ? h->num_cus : 256);
                const long c0 = n_tiles > 0 ? ((long)(ns + nb) * n_tiles + want - 1) / want : (long)(ns + nb);
                const long chunk = std::max<long>(block, ((c0 + block - 1) / block) * block);
                if (chunk == prev) continue;
                prev = chunk;
                for (int pers : {0, 2}) {
                    if (split && pers) continue;  // no persistent form of the split
                    time_shape(TileShape{block, r, (int)chunk, pers, split});
                    if (rc) break;
                }
                if (rc) break;
            }
            if (rc) break;
        }
        if (rc) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    h->profiling = prof;
    h->desc.precision = saved_precision;
    h->shape[precision] = best;
    h->shape_chosen[precision] = true;
    if (best_ms) *best_ms = best_t;
    return rc;
}

int tfs_rpy_set_persistent(tfs_rpy* h, int32_t precision, int32_t workgroups_per_cu) {
    if (!h || (precision != 0 && precision != 1) || workgroups_per_cu < 0 || workgroups_per_cu > 64)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_set_persistent: bad arguments");
    if (workgroups_per_cu > 0 && (h->shape[precision].split || h->shape[precision].symmetric)) {
        // the split and the symmetric form have no persistent form: a default
        // shape gives way to the plain default
        if (h->shape_chosen[precision])
            return fail(TFS_ERR_INVALID_ARGUMENT, h->shape[precision].split
                                                      ? "tfs_rpy_set_persistent: the source split has no persistent form"
                                                      : "tfs_rpy_set_persistent: the symmetric form has no persistent form");
        int ns = 0, nb = 0;
        largest_body(h, &ns, &nb);
        h->shape[precision] = plain_shape(ns, nb, h->num_cus);
        h->shape_chosen[precision] = true;
    }
    h->shape[precision].persistent = workgroups_per_cu;
    return TFS_OK;
}

int tfs_rpy_get_tile_shape(tfs_rpy* h, int32_t precision, int32_t shape[4]) {
    if (!h || !shape || (precision != 0 && precision != 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_get_tile_shape: bad arguments");
    const TileShape& s = h->shape[precision];
    shape[0] = s.block;
    shape[1] = s.per_thread;
    shape[2] = s.chunk;
    shape[3] = s.persistent;
    return TFS_OK;
}

int tfs_rpy_get_source_split(tfs_rpy* h, int32_t precision, int32_t* split) {
    if (!h || !split || (precision != 0 && precision != 1))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_get_source_split: bad arguments");
    *split = h->shape[precision].split;
    return TFS_OK;
}

int tfs_rpy_body_timing(tfs_rpy* h, int32_t enable) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_body_timing: null handle");
    h->body_timing = enable != 0;
    return TFS_OK;
}

int tfs_rpy_body_times(tfs_rpy* h, double* ms_per_body) {
    if (!h || !ms_per_body) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_body_times: null argument");
    const Plan* p = h->stamped_plan;
    if (!p) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_body_times: no timed launch yet");
    const size_t ni = p->items.size(), nb = h->bodies.size();
    std::vector<int64_t> st(2 * ni);
    hipError_t e = hipMemcpyAsync(st.data(), h->d_item_stamps, sizeof(int64_t) * 2 * ni, hipMemcpyDeviceToHost,
                                  h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    int khz = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device);
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_body_times");
    if (khz <= 0) khz = 100000;
    int64_t t0 = INT64_MAX;
    for (size_t i = 0; i < ni; ++i) t0 = std::min(t0, st[2 * i]);
    std::vector<int64_t> done(nb, t0);
    const int nt = (int)p->tiles.size();
    for (size_t i = 0; i < ni; ++i) {
        if (p->symmetric) {
            // an item pairs tile groups ga <= gb: the bodies both hold
            const int ga = p->items[i].tile, gb = p->items[i].src_cls;
            const int g = kSymGroupTiles;
            const int lo = std::max(p->tiles[g * ga].body, p->tiles[g * gb].body);
            const int hi = std::min(p->tiles[std::min(g * ga + g, nt) - 1].body, p->tiles[std::min(g * gb + g, nt) - 1].body);
            for (int b = lo; b <= hi; ++b) done[b] = std::max(done[b], st[2 * i + 1]);
        } else {
            const int b = p->tiles[p->items[i].tile].body;
            done[b] = std::max(done[b], st[2 * i + 1]);
        }
    }
    for (size_t b = 0; b < nb; ++b) ms_per_body[b] = double(done[b] - t0) / double(khz);
    return TFS_OK;
}

int tfs_rpy_profile(tfs_rpy* h, int32_t enable) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_profile: null handle");
    h->profiling = enable != 0;
    return TFS_OK;
}

int tfs_rpy_profile_read(tfs_rpy* h, double* pair_ms, int32_t* launches) {
    if (!h || !pair_ms || !launches) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_profile_read: null argument");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_profile_read");
    pair_ms[0] = pair_ms[1] = 0.0;
    launches[0] = launches[1] = 0;
    for (auto& pr : h->ev_used) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, pr.begin, pr.end);
        if (e != hipSuccess) return hip_fail(e, "tfs_rpy_profile_read: elapsed");
        pair_ms[pr.precision] += ms;
        launches[pr.precision] += 1;
        h->ev_free.push_back(pr);
    }
    h->ev_used.clear();
    return TFS_OK;
}

int tfs_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_host_alloc: null argument");
    *out = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipHostMalloc(out, bytes, hipHostMallocPortable);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_host_alloc");
}

int tfs_host_free(void* p) {
    if (!p) return TFS_OK;
    hipError_t e = hipHostFree(p);
    return e == hipSuccess ? TFS_OK : hip_fail(e, "tfs_host_free");
}

int tfs_rpy_device_buffer(tfs_rpy* h, int32_t what, double** out) {
    if (!h || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_device_buffer: null argument");
    size_t len;
    return array_info(h, what, out, &len);
}

// ---- FP32 arrays (Real = float builds) -------------------------------------

int tfs_rpy_upload_f32(tfs_rpy* h, int32_t what, const float* src) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload_f32: null handle");
    if (what > TFS_BLOB_FORCES && what != TFS_SEG_VELOCITIES && what != TFS_BLOB_VELOCITIES)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload_f32: not an input array");
    double* dev;
    size_t len;
    int rc = array_info(h, what, &dev, &len);
    if (rc) return rc;
    if (len == 0) return TFS_OK;
    if (!src) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_upload_f32: null source");
    if (what == TFS_SEG_POSITIONS || what == TFS_BLOB_POSITIONS) invalidate_near(h);
    return widen_in(h->f32, src, dev, (int64_t)len, h->desc.memory == TFS_MEMORY_HOST, h->stream,
                    "tfs_rpy_upload_f32");
}

int tfs_rpy_download_f32(tfs_rpy* h, int32_t what, float* dst) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download_f32: null handle");
    if (what != TFS_SEG_VELOCITIES && what != TFS_BLOB_VELOCITIES)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download_f32: not an output array");
    double* dev;
    size_t len;
    int rc = array_info(h, what, &dev, &len);
    if (rc) return rc;
    if (len == 0) return TFS_OK;
    if (!dst) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_download_f32: null destination");
    return narrow_out(h->f32, dev, dst, (int64_t)len, h->desc.memory == TFS_MEMORY_HOST, h->stream,
                      "tfs_rpy_download_f32");
}

int tfs_rpy_apply_f32(tfs_rpy* h, uint32_t blocks, const float* f_segs, const float* f_blobs, float* v_segs,
                      float* v_blobs) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply_f32: null handle");
    if (blocks & ~TFS_BLOCK_ALL) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply_f32: bad block mask");
    const bool need_fs = blocks & (TFS_BLOCK_SS | TFS_BLOCK_BS);
    const bool need_fb = blocks & (TFS_BLOCK_SB | TFS_BLOCK_BB);
    const bool out_s = blocks & (TFS_BLOCK_SS | TFS_BLOCK_SB);
    const bool out_b = blocks & (TFS_BLOCK_BB | TFS_BLOCK_BS);
    const bool ns = h->desc.num_segs > 0, nb = h->desc.num_blobs > 0;
    if ((need_fs && ns && !f_segs) || (need_fb && nb && !f_blobs) || (out_s && ns && !v_segs) ||
        (out_b && nb && !v_blobs))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_rpy_apply_f32: missing array for requested block");
    int rc = TFS_OK;
    if (need_fs && ns) rc = tfs_rpy_upload_f32(h, TFS_SEG_FORCES, f_segs);
    if (!rc && need_fb && nb) rc = tfs_rpy_upload_f32(h, TFS_BLOB_FORCES, f_blobs);
    // Accumulating blocks read the caller's current velocities.
    const bool acc_s = (blocks & TFS_BLOCK_SB) && !(blocks & TFS_BLOCK_SS);
    const bool acc_b = (blocks & TFS_BLOCK_BS) && !(blocks & TFS_BLOCK_BB) && h->desc.bs_mode == TFS_BS_ACCUMULATE;
    if (!rc && acc_s && ns) rc = tfs_rpy_upload_f32(h, TFS_SEG_VELOCITIES, v_segs);
    if (!rc && acc_b && nb) rc = tfs_rpy_upload_f32(h, TFS_BLOB_VELOCITIES, v_blobs);
    if (!rc) rc = tfs_rpy_evaluate(h, blocks);
    if (!rc && out_s && ns) rc = tfs_rpy_download_f32(h, TFS_SEG_VELOCITIES, v_segs);
    if (!rc && out_b && nb) rc = tfs_rpy_download_f32(h, TFS_BLOB_VELOCITIES, v_blobs);
    return rc;
}

int tfs_rpy_barrier_forces_f32(tfs_rpy* h, const tfs_barrier_desc* desc, float* f_segs, float* f_blobs_repulsion) {
    BarrierParams bp;
    int rc = barrier_params(h, desc, f_segs != nullptr, &bp, "tfs_rpy_barrier_forces_f32");
    if (rc) return rc;
    const int64_t ns = h->desc.num_segs, nb = h->desc.num_blobs;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    if (f_blobs_repulsion && nb > 0 && !h->blob_rep && hipMalloc(&h->blob_rep, 24 * nb) != hipSuccess)
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_rpy_barrier_forces_f32: device memory");
    rc = widen_in(h->f32, f_segs, h->f_segs, 6 * ns, host, h->stream, "tfs_rpy_barrier_forces_f32");
    if (rc) return rc;
    const MatvecArgs a = matvec_args(h, CView{h->f_segs, 6, 1}, CView{h->f_blobs, 3, 1});
    hipError_t e = launch_barrier(a, h->grid, bp, h->f_segs, f_blobs_repulsion && nb > 0 ? h->blob_rep : nullptr,
                                  h->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_rpy_barrier_forces_f32");
    rc = narrow_out(h->f32, h->f_segs, f_segs, 6 * ns, host, h->stream, "tfs_rpy_barrier_forces_f32");
    if (!rc && f_blobs_repulsion && nb > 0)
        rc = narrow_out(h->f32, h->blob_rep, f_blobs_repulsion, 3 * nb, host, h->stream, "tfs_rpy_barrier_forces_f32");
    if (!rc && host) rc = hip_fail_if(hipStreamSynchronize(h->stream), "tfs_rpy_barrier_forces_f32");
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// tfs_rpy_apply_noalloc: the matvec of tfs_rpy_apply for callers that cannot
// allocate.  Everything it needs beyond the caller's arrays -- the work list,
// the partial sums, the overlapping-pair lists, the hashed grid they are found
// with and the fixed-point positions of FP32 tiles -- lives in `scratch`, one
// file-scope device array of kScratchBytes that each GPU holds for the life of
// the process.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a host-side stand-in): its device address is looked up
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
namespace {

constexpr int kStaticMaxParticles = 32768;   // segments + blobs per call
constexpr int kStaticTile = 256;             // targets per tile: 256 threads, one target each
constexpr int kStaticNearCap = 32;           // overlapping neighbours per particle
constexpr size_t kStaticPartialBytes = size_t(16) << 20;
constexpr int kStaticMaxItems = (int)(kStaticPartialBytes / (6 * kStaticTile * sizeof(double)));
constexpr int kStaticMaxTiles = 2 * (kStaticMaxParticles / kStaticTile + 1);
constexpr int kStaticTable = 65536;          // grid buckets: a power of two >= 2 kStaticMaxParticles

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kOffPartial = 0;
constexpr size_t kOffNearIdx = align256(kOffPartial + kStaticPartialBytes);
constexpr size_t kOffNearCnt = align256(kOffNearIdx + sizeof(int32_t) * kStaticMaxParticles * kStaticNearCap);
constexpr size_t kOffNearMax = align256(kOffNearCnt + sizeof(int32_t) * kStaticMaxParticles);
constexpr size_t kOffStart = align256(kOffNearMax + sizeof(int32_t));
constexpr size_t kOffFill = align256(kOffStart + sizeof(int32_t) * (kStaticTable + 1));
constexpr size_t kOffBucket = align256(kOffFill + sizeof(int32_t) * kStaticTable);
constexpr size_t kOffOrder = align256(kOffBucket + sizeof(int32_t) * kStaticMaxParticles);
constexpr size_t kOffBodyOf = align256(kOffOrder + sizeof(int32_t) * kStaticMaxParticles);
constexpr size_t kOffFixed = align256(kOffBodyOf + sizeof(int32_t) * kStaticMaxParticles);
constexpr size_t kOffBody = align256(kOffFixed + sizeof(int4) * kStaticMaxParticles);
constexpr size_t kOffBox = align256(kOffBody + sizeof(int4));
constexpr size_t kOffTiles = align256(kOffBox + 6 * sizeof(double));
constexpr size_t kOffItems = align256(kOffTiles + sizeof(MatvecTile) * kStaticMaxTiles);
constexpr size_t kScratchBytes = align256(kOffItems + sizeof(MatvecItem) * kStaticMaxItems);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on first use).
int scratch_address(unsigned char** out, const char* where) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, where);
    if (dev < 0 || dev >= kMaxDevices) return fail(TFS_ERR_NO_DEVICE, std::string(where) + ": device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, where);
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return TFS_OK;
}

// The work list of one body: ts segment tiles (is items each) then tb blob
// tiles (ib items each); a tile's items take its sources in chunks, segments
// first, then blobs -- the order build_worklist gives.  One thread per tile;
// thread 0 also writes the body record.
__global__ __launch_bounds__(256) void static_worklist_kernel(int ns, int nb, int chunk, int ts, int tb, int is_,
                                                              int ib, int seg_write, int blob_write, int ss, int sb,
                                                              int bs, int bb, MatvecTile* __restrict__ tiles,
                                                              MatvecItem* __restrict__ items, int4* __restrict__ body) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t == 0) *body = make_int4(0, ns, 0, nb);
    if (t >= ts + tb) return;
    const bool seg = t < ts;
    const int u = seg ? t : t - ts;
    const int n_t = seg ? ns : nb;
    MatvecTile tl;
    tl.cls = seg ? 0 : 1;
    tl.t_begin = u * kStaticTile;
    tl.t_end = tl.t_begin + kStaticTile < n_t ? tl.t_begin + kStaticTile : n_t;
    tl.first_item = seg ? u * is_ : ts * is_ + u * ib;
    tl.n_items = seg ? is_ : ib;
    tl.write_mode = seg ? seg_write : blob_write;
    tl.body = 0;
    tiles[t] = tl;
    int k = tl.first_item;
    if (seg ? ss : bs)
        for (int s0 = 0; s0 < ns; s0 += chunk) items[k++] = MatvecItem{t, 0, s0, s0 + chunk < ns ? s0 + chunk : ns};
    if (seg ? sb : bb)
        for (int s0 = 0; s0 < nb; s0 += chunk) items[k++] = MatvecItem{t, 1, s0, s0 + chunk < nb ? s0 + chunk : nb};
}

// X = rint(lambda (x - c)) for the one body of the call.
__global__ __launch_bounds__(256) void static_fixed_kernel(const double* __restrict__ xs, const double* __restrict__ xb,
                                                           int ns, int nb, double c0, double c1, double c2, double lam,
                                                           int4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ns + nb) return;
    const double* x = i < ns ? xs + 3 * (int64_t)i : xb + 3 * (int64_t)(i - ns);
    out[i] = make_int4((int)__builtin_rint((x[0] - c0) * lam), (int)__builtin_rint((x[1] - c1) * lam),
                       (int)__builtin_rint((x[2] - c2) * lam), 0);
}

int static_items(int ns, int nb, int chunk, int ts, int tb, bool ss, bool sb, bool bs, bool bb, int* is_, int* ib) {
    const int cs = (ns + chunk - 1) / chunk, cb = (nb + chunk - 1) / chunk;
    *is_ = (ss ? cs : 0) + (sb ? cb : 0);
    *ib = (bs ? cs : 0) + (bb ? cb : 0);
    return ts * *is_ + tb * *ib;
}

}  // namespace

extern "C" {

int tfs_rpy_apply_noalloc(const tfs_rpy_desc* desc, uint32_t blocks, const double* x_segs, const double* x_blobs,
                          const double* f_segs, const double* f_blobs, double* v_segs, double* v_blobs) {
    const char* w = "tfs_rpy_apply_noalloc";
    if (!desc) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": null descriptor");
    const tfs_rpy_desc& D = *desc;
    if (D.num_segs < 0 || D.num_blobs < 0 || (D.seg_dof != 3 && D.seg_dof != 6) ||
        (D.precision != TFS_PRECISION_FP64 && D.precision != TFS_PRECISION_FP32_ACC64) ||
        (D.bs_mode != TFS_BS_OVERWRITE && D.bs_mode != TFS_BS_ACCUMULATE) || !(D.seg_radius > 0) ||
        !(D.blob_radius > 0) || !(D.mu > 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": bad descriptor");
    if (D.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": takes device pointers (memory = TFS_MEMORY_DEVICE)");
    if (blocks & ~TFS_BLOCK_ALL) return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": bad block mask");
    const int ns = D.num_segs, nb = D.num_blobs;
    if ((int64_t)ns + nb > kStaticMaxParticles)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": at most " + std::to_string(kStaticMaxParticles) +
                                                  " particles (segments + blobs) per call; use tfs_rpy_apply");
    // the blocks evaluated and their write rules, as tfs_rpy_apply
    bool ss = blocks & TFS_BLOCK_SS, sb = blocks & TFS_BLOCK_SB, bb = blocks & TFS_BLOCK_BB, bs = blocks & TFS_BLOCK_BS;
    if (bs && bb && D.bs_mode == TFS_BS_OVERWRITE) bb = false;
    const int seg_write = ss ? 0 : (sb ? 1 : -1);
    const int blob_write = bb ? 0 : (bs ? (D.bs_mode == TFS_BS_OVERWRITE ? 0 : 1) : -1);
    const uint32_t eff = (ss ? 1u : 0u) | (sb ? 2u : 0u) | (bb ? 4u : 0u) | (bs ? 8u : 0u);
    if ((ns && !x_segs) || (nb && !x_blobs) || ((ss || bs) && ns && !f_segs) || ((sb || bb) && nb && !f_blobs) ||
        (seg_write >= 0 && ns && !v_segs) || (blob_write >= 0 && nb && !v_blobs))
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": missing array for requested block");
    const int ts = seg_write >= 0 ? (ns + kStaticTile - 1) / kStaticTile : 0;
    const int tb = blob_write >= 0 ? (nb + kStaticTile - 1) / kStaticTile : 0;
    if (ts + tb == 0) return TFS_OK;
    // the smallest chunk (a multiple of the tile) whose items fit the partial buffer
    int chunk = kStaticTile, is_ = 0, ib = 0;
    while (static_items(ns, nb, chunk, ts, tb, ss, sb, bs, bb, &is_, &ib) > kStaticMaxItems) chunk += kStaticTile;
    const int n_items = ts * is_ + tb * ib;
    unsigned char* base = nullptr;
    int rc = scratch_address(&base, w);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)D.stream;
    MatvecTile* tiles = reinterpret_cast<MatvecTile*>(base + kOffTiles);
    MatvecItem* items = reinterpret_cast<MatvecItem*>(base + kOffItems);
    int4* body = reinterpret_cast<int4*>(base + kOffBody);
    hipLaunchKernelGGL(static_worklist_kernel, dim3((unsigned)((ts + tb + 255) / 256)), dim3(256), 0, st, ns, nb, chunk,
                       ts, tb, is_, ib, seg_write, blob_write, (int)ss, (int)sb, (int)bs, (int)bb, tiles, items, body);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, w);

    MatvecArgs a;
    a.xs = {x_segs, 3, 1};
    a.xb = {x_blobs, 3, 1};
    a.fs = {f_segs, 6, 1};
    a.fb = {f_blobs, 3, 1};
    a.ns = ns;
    a.nb = nb;
    a.a_seg = D.seg_radius;
    a.a_blob = D.blob_radius;
    a.mu = D.mu;
    const int n = ns + nb;
    if (D.precision == TFS_PRECISION_FP32_ACC64) {
        // fixed-point frame of the body's bounding box (read back to the stack)
        double* box = reinterpret_cast<double*>(base + kOffBox);
        hipLaunchKernelGGL(body_box_kernel, dim3(1), dim3(256), 0, st, x_segs, x_blobs, (const int4*)body, box);
        double hbox[6], centre[3], lam = 1.0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(hbox, box, sizeof(hbox), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, w);
        rc = fixed_point_frame(hbox, 1, std::min(D.seg_radius, D.blob_radius), centre, &lam);
        if (rc) return rc;
        int4* fx = reinterpret_cast<int4*>(base + kOffFixed);
        hipLaunchKernelGGL(static_fixed_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x_segs, x_blobs,
                           ns, nb, centre[0], centre[1], centre[2], lam, fx);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, w);
        a.xfx = fx;
        a.lam = lam;
    }
    // overlapping pairs, on the grid kept in scratch (sized for n: no allocation)
    NearList near;
    near.cnt = reinterpret_cast<int32_t*>(base + kOffNearCnt);
    near.idx = reinterpret_cast<int32_t*>(base + kOffNearIdx);
    near.max_cnt = reinterpret_cast<int32_t*>(base + kOffNearMax);
    near.cap = kStaticNearCap;
    NearGrid grid;
    grid.start = reinterpret_cast<int32_t*>(base + kOffStart);
    grid.fill = reinterpret_cast<int32_t*>(base + kOffFill);
    grid.bucket_of = reinterpret_cast<int32_t*>(base + kOffBucket);
    grid.order = reinterpret_cast<int32_t*>(base + kOffOrder);
    grid.body_of = reinterpret_cast<int32_t*>(base + kOffBodyOf);
    grid.n = n;
    grid.table = 1024;
    while (grid.table < 2 * n) grid.table <<= 1;
    int max_count = 0, host_count = 0;
    e = build_near_list(D.precision, a, grid, reinterpret_cast<const int32_t*>(body), 1, near, n, &max_count, &host_count,
                        st);
    if (e != hipSuccess) return hip_fail(e, w);
    if (max_count > kStaticNearCap)
        return fail(TFS_ERR_INVALID_ARGUMENT, std::string(w) + ": a particle overlaps more than " +
                                                  std::to_string(kStaticNearCap) + " others; use tfs_rpy_apply");
    OutArgs o;
    o.vs = View{v_segs, 6, 1};
    o.vb = View{v_blobs, 3, 1};
    o.seg_out_slots = 6;
    o.blocks = eff;
    const TileShape sh{kStaticTile, 1, chunk, 0};
    PersistState ps;
    e = launch_matvec(D.precision, D.seg_dof, sh, a, o, tiles, ts + tb, items, n_items,
                      reinterpret_cast<double*>(base + kOffPartial), near, ps, st);
    return e == hipSuccess ? TFS_OK : hip_fail(e, w);
}

}  // extern "C"
