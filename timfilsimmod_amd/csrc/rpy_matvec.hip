// rpy_matvec.hip -- fused, tiled RPY mobility matvec for gfx950.
//
// Replaces the four reference kernels Mss_mult / Msb_mult / Mbb_mult /
// Mbs_mult (src/cilia/cuda_functions.cu:362-757) with one launch over a work
// list.  A work item pairs a tile of targets (segments or blobs) with a chunk
// of sources (segments or blobs).  A workgroup stages its chunk through LDS in
// batches of `BLOCK` sources (one coalesced load per lane, converted to the
// evaluation type as it is stored), every lane then walks the batch with
// broadcast LDS reads and accumulates PER_THREAD targets in registers.  Each
// item writes its partial sums to its own slot; a second kernel adds the slots
// of a tile in a fixed order and applies the reference's write rule (= or +=)
// -- so results are deterministic run to run, unlike atomics.
//
// Splitting the source range into chunks is what fills the chip: 21,800
// targets are only 341 wave64s, far fewer than the 1,024 SIMDs of an MI355X;
// with chunks of ~1k sources the production body gives ~2k workgroups.
//
// Two evaluation types:
//   double : the FP64 path (parity with the FP64 oracle at <= 1e-12).
//   float  : pair tensors in FP32 from double-float positions (x = hi + lo,
//            so near pairs -- the ones that dominate -- get an exact hi
//            difference, by Sterbenz's lemma); per-batch FP32 partial sums are
//            flushed into FP64 accumulators after every LDS batch.
#include "rpy_pair.h"
#include "tfs_internal.h"

namespace tfs {
namespace {

constexpr int kMaxDof = 6;

// LDS record length (in T) of one staged source with FS strength components.
template <typename T, int FS>
constexpr int rec_len() {
    if constexpr (sizeof(T) == 8) {
        return FS == 6 ? 10 : 6;   // x y z f0..f(FS-1) [pad]
    } else {
        return 12;                 // xh yh zh xl yl zl f0..f5 (3-dof: padded)
    }
}

template <int VT, int FS, typename T, int BLOCK, int R>
__device__ __forceinline__ void run_item(const MatvecItem it, const MatvecTile tl, const MatvecArgs& a,
                                         double* __restrict__ slot, T* __restrict__ lds) {
    constexpr int REC = rec_len<T, FS>();
    constexpr int TILE = BLOCK * R;
    constexpr bool F32 = sizeof(T) == 4;
    const int tid = threadIdx.x;

    const CView xt = tl.cls == 0 ? a.xs : a.xb;
    const CView xsrc = it.src_cls == 0 ? a.xs : a.xb;
    const CView fsrc = it.src_cls == 0 ? a.fs : a.fb;
    const double ai = tl.cls == 0 ? a.a_seg : a.a_blob;
    const double aj = it.src_cls == 0 ? a.a_seg : a.a_blob;
    const double mu = a.mu;
    const double PI = 3.14159265358979323846;
    const T sum_a2 = T((ai + aj) * (ai + aj));
    const T s3 = T((ai * ai + aj * aj) / 3.0);
    const T c8 = T(1.0 / (8.0 * PI * mu));

    // Targets of this lane (clamped into the tile; surplus lanes compute a
    // duplicate and their slot entries are never read).
    double xi[R][3];
    T xh[R][3], xl[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int t = tl.t_begin + tid + r * BLOCK;
        t = t < tl.t_end ? t : tl.t_end - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xi[r][k] = xt.p[t * xt.ld_p + k * xt.ld_c];
            if constexpr (F32) {
                xh[r][k] = T(xi[r][k]);
                xl[r][k] = T(xi[r][k] - double(xh[r][k]));
            }
        }
    }

    double acc[R][VT];
    T part[R][VT];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < VT; ++k) {
            acc[r][k] = 0.0;
            part[r][k] = T(0);
        }

    for (int s0 = it.s_begin; s0 < it.s_end; s0 += BLOCK) {
        const int n = (it.s_end - s0) < BLOCK ? (it.s_end - s0) : BLOCK;
        __syncthreads();  // the previous batch has been consumed
        if (tid < n) {
            const int s = s0 + tid;
            T* rec = lds + tid * REC;
            double xs0 = xsrc.p[s * xsrc.ld_p];
            double xs1 = xsrc.p[s * xsrc.ld_p + xsrc.ld_c];
            double xs2 = xsrc.p[s * xsrc.ld_p + 2 * xsrc.ld_c];
            if constexpr (F32) {
                const float h0 = float(xs0), h1 = float(xs1), h2 = float(xs2);
                rec[0] = h0; rec[1] = h1; rec[2] = h2;
                rec[3] = float(xs0 - double(h0));
                rec[4] = float(xs1 - double(h1));
                rec[5] = float(xs2 - double(h2));
#pragma unroll
                for (int k = 0; k < FS; ++k) rec[6 + k] = float(fsrc.p[s * fsrc.ld_p + k * fsrc.ld_c]);
            } else {
                rec[0] = xs0; rec[1] = xs1; rec[2] = xs2;
#pragma unroll
                for (int k = 0; k < FS; ++k) rec[3 + k] = fsrc.p[s * fsrc.ld_p + k * fsrc.ld_c];
            }
        }
        __syncthreads();

#pragma unroll 2
        for (int j = 0; j < n; ++j) {
            const T* rec = lds + j * REC;
            T f[FS];
            if constexpr (F32) {
#pragma unroll
                for (int k = 0; k < FS; ++k) f[k] = rec[6 + k];
            } else {
#pragma unroll
                for (int k = 0; k < FS; ++k) f[k] = rec[3 + k];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                T dx, dy, dz;
                if constexpr (F32) {
                    dx = (xh[r][0] - rec[0]) + (xl[r][0] - rec[3]);
                    dy = (xh[r][1] - rec[1]) + (xl[r][1] - rec[4]);
                    dz = (xh[r][2] - rec[2]) + (xl[r][2] - rec[5]);
                } else {
                    dx = xi[r][0] - rec[0];
                    dy = xi[r][1] - rec[1];
                    dz = xi[r][2] - rec[2];
                }
                const T r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > sum_a2) {
                    const T rm1 = rsqrt_refined(r2);
                    if constexpr (F32) {
                        pair_far<T, VT, FS>(part[r], dx, dy, dz, rm1, f, s3, c8);
                    } else {
                        pair_far<T, VT, FS>(acc[r], dx, dy, dz, rm1, f, s3, c8);
                    }
                } else {
                    double fd[FS];
#pragma unroll
                    for (int k = 0; k < FS; ++k) fd[k] = double(f[k]);
                    double ddx, ddy, ddz;
                    if constexpr (F32) {
                        ddx = xi[r][0] - (double(rec[0]) + double(rec[3]));
                        ddy = xi[r][1] - (double(rec[1]) + double(rec[4]));
                        ddz = xi[r][2] - (double(rec[2]) + double(rec[5]));
                    } else {
                        ddx = dx; ddy = dy; ddz = dz;
                    }
                    pair_near<VT, FS>(acc[r], ddx, ddy, ddz, fd, ai, aj, mu);
                }
            }
        }
        if constexpr (F32) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int k = 0; k < VT; ++k) {
                    acc[r][k] += double(part[r][k]);
                    part[r][k] = T(0);
                }
        }
    }

#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < VT; ++k) slot[k * TILE + tid + r * BLOCK] = acc[r][k];
}

template <typename T, int SEGDOF, int BLOCK, int R>
__global__ __launch_bounds__(BLOCK) void rpy_pairs_kernel(const MatvecItem* __restrict__ items,
                                                          const MatvecTile* __restrict__ tiles,
                                                          MatvecArgs a, double* __restrict__ partial) {
    constexpr int REC = rec_len<T, 6>() > rec_len<T, 3>() ? rec_len<T, 6>() : rec_len<T, 3>();
    __shared__ __align__(16) T lds[BLOCK * REC];
    const MatvecItem it = items[blockIdx.x];
    const MatvecTile tl = tiles[it.tile];
    double* slot = partial + (size_t)blockIdx.x * (kMaxDof * BLOCK * R);
    if (tl.cls == 0) {
        if (it.src_cls == 0) run_item<SEGDOF, SEGDOF, T, BLOCK, R>(it, tl, a, slot, lds);
        else run_item<SEGDOF, 3, T, BLOCK, R>(it, tl, a, slot, lds);
    } else {
        if (it.src_cls == 0) run_item<3, SEGDOF, T, BLOCK, R>(it, tl, a, slot, lds);
        else run_item<3, 3, T, BLOCK, R>(it, tl, a, slot, lds);
    }
}

// Adds the partial slots of each tile in item order and applies the write rule.
template <int SEGDOF>
__global__ void rpy_reduce_kernel(const MatvecTile* __restrict__ tiles, const double* __restrict__ partial,
                                  int tile_targets, OutArgs o) {
    const MatvecTile tl = tiles[blockIdx.x];
    const int vt = tl.cls == 0 ? SEGDOF : 3;
    const int slots = tl.cls == 0 ? o.seg_out_slots : 3;
    const View out = tl.cls == 0 ? o.vs : o.vb;
    const size_t slot_len = (size_t)kMaxDof * tile_targets;
    for (int l = threadIdx.x; l < tile_targets; l += blockDim.x) {
        const int t = tl.t_begin + l;
        if (t >= tl.t_end) break;
        for (int k = 0; k < slots; ++k) {
            double s = 0.0;
            if (k < vt) {
                const double* p = partial + (size_t)tl.first_item * slot_len + (size_t)k * tile_targets + l;
                for (int m = 0; m < tl.n_items; ++m) s += p[(size_t)m * slot_len];
            }
            double* dst = out.p + t * out.ld_p + k * out.ld_c;
            if (tl.write_mode == 0) *dst = s;
            else if (k < vt) *dst += s;
        }
    }
}

template <typename T, int SEGDOF, int BLOCK, int R>
hipError_t launch_pairs(const MatvecArgs& a, const MatvecItem* d_items, int n_items,
                        const MatvecTile* d_tiles, double* partial, hipStream_t stream) {
    hipLaunchKernelGGL((rpy_pairs_kernel<T, SEGDOF, BLOCK, R>), dim3(n_items), dim3(BLOCK), 0, stream,
                       d_items, d_tiles, a, partial);
    return hipGetLastError();
}

template <typename T, int SEGDOF>
hipError_t dispatch_shape(const TileShape& sh, const MatvecArgs& a, const MatvecItem* d_items, int n_items,
                          const MatvecTile* d_tiles, double* partial, hipStream_t stream) {
#define TFS_SHAPE(B, R)                                                                       \
    if (sh.block == B && sh.per_thread == R)                                                 \
        return launch_pairs<T, SEGDOF, B, R>(a, d_items, n_items, d_tiles, partial, stream);
    TFS_SHAPE(64, 1) TFS_SHAPE(64, 2) TFS_SHAPE(128, 1) TFS_SHAPE(128, 2) TFS_SHAPE(256, 1)
    TFS_SHAPE(256, 2)
#undef TFS_SHAPE
    return hipErrorInvalidValue;
}

}  // namespace

void build_worklist(int ns, int nb, int tile_targets, int chunk, bool ss, bool sb, bool bb, bool bs,
                    int seg_write, int blob_write, std::vector<MatvecTile>& tiles,
                    std::vector<MatvecItem>& items) {
    tiles.clear();
    items.clear();
    auto add_tiles = [&](int cls, int n, bool from_segs, bool from_blobs, int write) {
        if (write < 0) return;
        for (int t0 = 0; t0 < n; t0 += tile_targets) {
            MatvecTile tl;
            tl.cls = cls;
            tl.t_begin = t0;
            tl.t_end = t0 + tile_targets < n ? t0 + tile_targets : n;
            tl.first_item = (int)items.size();
            tl.write_mode = write;
            const int tile_id = (int)tiles.size();
            auto add_src = [&](int src_cls, int m) {
                for (int s0 = 0; s0 < m; s0 += chunk)
                    items.push_back({tile_id, src_cls, s0, s0 + chunk < m ? s0 + chunk : m});
            };
            if (from_segs) add_src(0, ns);
            if (from_blobs) add_src(1, nb);
            tl.n_items = (int)items.size() - tl.first_item;
            tiles.push_back(tl);
        }
    };
    add_tiles(0, ns, ss, sb, seg_write);
    add_tiles(1, nb, bs, bb, blob_write);
}

hipError_t launch_matvec(int precision, int seg_dof, const TileShape& shape, const MatvecArgs& a,
                         const OutArgs& o, const MatvecTile* d_tiles, int n_tiles,
                         const MatvecItem* d_items, int n_items, double* partial, hipStream_t stream,
                         hipEvent_t ev_begin, hipEvent_t ev_end) {
    if (n_tiles == 0) return hipSuccess;
    if (n_items > 0) {
        hipError_t e;
        if (ev_begin) TFS_HIP_CHECK(hipEventRecord(ev_begin, stream));
        if (precision == 0) {
            e = seg_dof == 6 ? dispatch_shape<double, 6>(shape, a, d_items, n_items, d_tiles, partial, stream)
                             : dispatch_shape<double, 3>(shape, a, d_items, n_items, d_tiles, partial, stream);
        } else {
            e = seg_dof == 6 ? dispatch_shape<float, 6>(shape, a, d_items, n_items, d_tiles, partial, stream)
                             : dispatch_shape<float, 3>(shape, a, d_items, n_items, d_tiles, partial, stream);
        }
        if (e != hipSuccess) return e;
        if (ev_end) TFS_HIP_CHECK(hipEventRecord(ev_end, stream));
    }
    const int tt = shape.block * shape.per_thread;
    const int rb = tt < 256 ? tt : 256;
    if (seg_dof == 6)
        hipLaunchKernelGGL(rpy_reduce_kernel<6>, dim3(n_tiles), dim3(rb), 0, stream, d_tiles, partial, tt, o);
    else
        hipLaunchKernelGGL(rpy_reduce_kernel<3>, dim3(n_tiles), dim3(rb), 0, stream, d_tiles, partial, tt, o);
    return hipGetLastError();
}

}  // namespace tfs
