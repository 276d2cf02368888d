// flow_field.hip -- flow-field velocity at tracer points from segment and blob
// sources, for gfx950.
//
// Two modes:
//   Stokeslet + rotlet  -- the reference evaluator flow_velocities
//                          (src/flow_field/flow_field_evaluator.cu:13-228):
//                          Stokeslet of every segment force, rotlet of every
//                          segment torque, Stokeslet of every blob force, the
//                          Blakelet image system when there are no blobs
//                          (:83-116), optional removal of each body's net-force
//                          Stokeslet placed at the body centre (:125-143,
//                          :191-209), scaled by 1/(8 pi mu) (:549).
//   RPY                 -- the point-tracer limit (target radius 0) of the RPY
//                          pair tensor, <3,6> from segments and <3,3> from blobs.
//
// Two targets per lane, all sources streamed through LDS in batches of BLOCK;
// the 4.2M-point plane is 8k workgroups, so no source splitting is needed.  FP32 evaluation uses double-float positions and
// flushes FP32 partial sums into FP64 accumulators after every batch.
#include "rpy_pair.h"
#include "tfs_internal.h"

namespace tfs {
namespace {

template <typename T>
constexpr int flow_rec() {
    // FP64: x y z f0..f5 (+pad).  FP32: xh yh zh xl yl zl f0..f5.  (Staging the
    // negated torque too, so that the rotlet terms become accumulates, measured
    // slower: the extra 16 B per source outweighs the saved issue slots.)
    return sizeof(T) == 8 ? 10 : 12;
}

// Image system of a Stokeslet above the no-slip wall z = 0 (Blake 1971):
// subtracts the image Stokeslet and adds the Stokeslet doublet and source
// dipole terms of flow_field_evaluator.cu:83-116.  acc += image velocity.
__device__ __forceinline__ void blakelet(double* acc, double zi, double zj, double ddx, double ddy, double fx,
                                         double fy, double fz) {
    const double Rz = zi + zj;
    const double R2 = ddx * ddx + ddy * ddy + Rz * Rz;
    const double Rm1 = __builtin_sqrt(1.0 / R2);
    const double kk = (ddx * fx + ddy * fy + Rz * fz) * Rm1 * Rm1;
    acc[0] -= Rm1 * (fx + ddx * kk);
    acc[1] -= Rm1 * (fy + ddy * kk);
    acc[2] -= Rm1 * (fz + Rz * kk);
    const double Dxx = (3.0 * ddx * ddx - R2) * zi;
    const double Dxy = 3.0 * ddx * ddy * zi;
    const double Dxz = (R2 - 3.0 * Rz * zi) * ddx;
    const double Dyy = (3.0 * ddy * ddy - R2) * zi;
    const double Dyz = (R2 - 3.0 * Rz * zi) * ddy;
    const double Dzx = (3.0 * Rz * zi + R2) * ddx;
    const double Dzy = (3.0 * Rz * zi + R2) * ddy;
    const double Dzz = (R2 - 3.0 * Rz * Rz) * zi;
    const double g = 2.0 * zj * Rm1 * Rm1 * Rm1 * Rm1 * Rm1;
    acc[0] += g * (Dxx * fx + Dxy * fy + Dxz * fz);
    acc[1] += g * (Dxy * fx + Dyy * fy + Dyz * fz);
    acc[2] += g * (Dzx * fx + Dzy * fy + Dzz * fz);
}

template <typename T, int MODE, bool BLAKE, int FS, int BLOCK, int R>
__device__ __forceinline__ void flow_sources(const FlowArgs& a, CView xsrc, CView fsrc, int nsrc, double aj,
                                             const double (&xi)[R][3], const T (&xh)[R][3],
                                             const T (&xl)[R][3], double (&acc)[R][3], T* lds) {
    constexpr int REC = flow_rec<T>();
    constexpr bool F32 = sizeof(T) == 4;
    const int tid = threadIdx.x;
    const double PI = 3.14159265358979323846;
    const T s3 = T(aj * aj / 3.0);
    const T c8 = T(1.0 / (8.0 * PI * a.mu));
    const T aj2 = T(aj * aj);
    T part[R][3];
    // FP32 Stokeslet mode: the rotlet's negative cross-product halves, summed
    // apart and subtracted at the flush (every operation an accumulate)
    T negp[R][3];
    for (int s0 = 0; s0 < nsrc; s0 += BLOCK) {
        const int n = (nsrc - s0) < BLOCK ? (nsrc - s0) : BLOCK;
        __syncthreads();
        if (tid < n) {
            const int64_t s = s0 + tid;
            T* rec = lds + tid * REC;
            const double p0 = xsrc.p[s * xsrc.ld_p], p1 = xsrc.p[s * xsrc.ld_p + xsrc.ld_c],
                         p2 = xsrc.p[s * xsrc.ld_p + 2 * xsrc.ld_c];
            if constexpr (F32) {
                const float h0 = float(p0), h1 = float(p1), h2 = float(p2);
                rec[0] = h0; rec[1] = h1; rec[2] = h2;
                rec[3] = float(p0 - double(h0)); rec[4] = float(p1 - double(h1)); rec[5] = float(p2 - double(h2));
#pragma unroll
                for (int k = 0; k < FS; ++k) rec[6 + k] = float(fsrc.p[s * fsrc.ld_p + k * fsrc.ld_c]);

            } else {
                rec[0] = p0; rec[1] = p1; rec[2] = p2;
#pragma unroll
                for (int k = 0; k < FS; ++k) rec[3 + k] = fsrc.p[s * fsrc.ld_p + k * fsrc.ld_c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            part[r][0] = part[r][1] = part[r][2] = T(0);
            negp[r][0] = negp[r][1] = negp[r][2] = T(0);
        }
        for (int j = 0; j < n; ++j) {
            const T* rec = lds + j * REC;
            constexpr int F0 = F32 ? 6 : 3;
            T f[FS];
#pragma unroll
            for (int k = 0; k < FS; ++k) f[k] = rec[F0 + k];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                T dx, dy, dz;
                if constexpr (F32) {
                    dx = (xh[r][0] - rec[0]) + (xl[r][0] - rec[3]);
                    dy = (xh[r][1] - rec[1]) + (xl[r][1] - rec[4]);
                    dz = (xh[r][2] - rec[2]) + (xl[r][2] - rec[5]);
                } else {
                    dx = T(xi[r][0]) - rec[0];
                    dy = T(xi[r][1]) - rec[1];
                    dz = T(xi[r][2]) - rec[2];
                }
                const T r2 = dx * dx + dy * dy + dz * dz;
                if constexpr (MODE == 0 && F32) {
                    // Stokeslet (+ rotlet): flow_field_evaluator.cu:66-78, 119-124, written
                    // as accumulates only: v += f/r + (d.f) d / r^3 + (tau x d) / r^3.
                    const T rm1 = rsqrt_refined(r2);
                    const T rm3 = rm1 * (rm1 * rm1);
                    const T sd = rm3 * fma_t(dz, f[2], fma_t(dy, f[1], dx * f[0]));
                    part[r][0] = fma_t(sd, dx, fma_t(rm1, f[0], part[r][0]));
                    part[r][1] = fma_t(sd, dy, fma_t(rm1, f[1], part[r][1]));
                    part[r][2] = fma_t(sd, dz, fma_t(rm1, f[2], part[r][2]));
                    if constexpr (FS == 6) {
                        // (tau x d) / r^3 as + and - halves
                        const T ex = rm3 * dx, ey = rm3 * dy, ez = rm3 * dz;
                        part[r][0] = fma_t(f[4], ez, part[r][0]);
                        negp[r][0] = fma_t(f[5], ey, negp[r][0]);
                        part[r][1] = fma_t(f[5], ex, part[r][1]);
                        negp[r][1] = fma_t(f[3], ez, negp[r][1]);
                        part[r][2] = fma_t(f[3], ey, part[r][2]);
                        negp[r][2] = fma_t(f[4], ex, negp[r][2]);
                    }
                    if constexpr (BLAKE) {
                        // Blakelet image system, flow_field_evaluator.cu:83-116, evaluated in
                        // double (the image distance of a wall-adjacent pair is small).
                        const double zi = xi[r][2];
                        const double zj = double(rec[2]) + double(rec[5]);
                        blakelet(acc[r], zi, zj, double(dx), double(dy), double(f[0]), double(f[1]), double(f[2]));
                    }
                } else if constexpr (MODE == 0) {
                    // Stokeslet (+ rotlet): flow_field_evaluator.cu:66-78, 119-124
                    const T rm1 = rsqrt_refined(r2);
                    const T rm2 = rm1 * rm1;
                    const T k = rm2 * (dx * f[0] + dy * f[1] + dz * f[2]);
                    part[r][0] += rm1 * (f[0] + dx * k);
                    part[r][1] += rm1 * (f[1] + dy * k);
                    part[r][2] += rm1 * (f[2] + dz * k);
                    if constexpr (FS == 6) {
                        const T rm3 = rm1 * rm2;
                        part[r][0] += rm3 * (f[4] * dz - f[5] * dy);
                        part[r][1] += rm3 * (f[5] * dx - f[3] * dz);
                        part[r][2] += rm3 * (f[3] * dy - f[4] * dx);
                    }
                    if constexpr (BLAKE) {
                        // Blakelet image system, flow_field_evaluator.cu:83-116, evaluated in
                        // double (the image distance of a wall-adjacent pair is small).
                        const double zi = xi[r][2];
                        double zj;
                        if constexpr (F32) zj = double(rec[2]) + double(rec[5]);
                        else zj = rec[2];
                        blakelet(acc[r], zi, zj, double(dx), double(dy), double(f[0]), double(f[1]), double(f[2]));
                    }
                } else {
                    // RPY point tracer: the pair tensor with ai = 0.
                    if (r2 > aj2) {
                        pair_far<T, 3, FS>(part[r], dx, dy, dz, rsqrt_refined(r2), f, s3, c8);
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
                        pair_near<3, FS>(acc[r], ddx, ddy, ddz, fd, 0.0, aj, a.mu);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[r][k] += double(part[r][k]) - double(negp[r][k]);
    }
}

template <typename T, int MODE, bool BLAKE, int BLOCK, int R>
__global__ __launch_bounds__(BLOCK) void flow_kernel(FlowArgs a, const double* __restrict__ net) {
    __shared__ __align__(16) T lds[BLOCK * flow_rec<T>()];
    constexpr bool F32 = sizeof(T) == 4;
    const int64_t t0 = (int64_t)blockIdx.x * (BLOCK * R) + threadIdx.x;
    double xi[R][3];
    T xh[R][3], xl[R][3];
    double acc[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t t = t0 + (int64_t)r * BLOCK;
        t = t < a.nt ? t : a.nt - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xi[r][k] = a.xt.p[t * a.xt.ld_p + k * a.xt.ld_c];
            if constexpr (F32) {
                xh[r][k] = T(xi[r][k]);
                xl[r][k] = T(xi[r][k] - double(xh[r][k]));
            } else {
                xh[r][k] = xl[r][k] = T(0);
            }
            acc[r][k] = 0.0;
        }
    }
    if (a.ns_total > 0)
        flow_sources<T, MODE, BLAKE, 6, BLOCK, R>(a, a.xs, a.fs, a.ns_total, a.a_seg, xi, xh, xl, acc, lds);
    if (a.nb_total > 0)
        flow_sources<T, MODE, false, 3, BLOCK, R>(a, a.xb, a.fb, a.nb_total, a.a_blob, xi, xh, xl, acc, lds);
    const double PI = 3.14159265358979323846;
    const double prefac = MODE == 0 ? 1.0 / (8.0 * PI * a.mu) : 1.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t t = t0 + (int64_t)r * BLOCK;
        if (t >= a.nt) continue;
        if (MODE == 0 && a.remove_net_force) {
            // Each body's total force as one Stokeslet at its centre
            // (the per-source sums of flow_field_evaluator.cu:125-143, 191-209).
            for (int b = 0; b < a.num_bodies; ++b) {
                const double dx = xi[r][0] - a.x_bodies[3 * b], dy = xi[r][1] - a.x_bodies[3 * b + 1],
                             dz = xi[r][2] - a.x_bodies[3 * b + 2];
                const double fx = net[3 * b], fy = net[3 * b + 1], fz = net[3 * b + 2];
                const double rm2 = 1.0 / (dx * dx + dy * dy + dz * dz);
                const double rm1 = __builtin_sqrt(rm2);
                const double k = rm2 * (dx * fx + dy * fy + dz * fz);
                acc[r][0] -= rm1 * (fx + dx * k);
                acc[r][1] -= rm1 * (fy + dy * k);
                acc[r][2] -= rm1 * (fz + dz * k);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) a.vt.p[t * a.vt.ld_p + k * a.vt.ld_c] = prefac * acc[r][k];
    }
}

// Per-body total force (segments + blobs), one thread per body.
__global__ void net_force_kernel(FlowArgs a, double* __restrict__ net, int nbod) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbod) return;
    double f[3] = {0, 0, 0};
    for (int64_t s = (int64_t)b * a.ns_per_body; s < (int64_t)(b + 1) * a.ns_per_body; ++s)
        for (int k = 0; k < 3; ++k) f[k] += a.fs.p[s * a.fs.ld_p + k * a.fs.ld_c];
    for (int64_t s = (int64_t)b * a.nb_per_body; s < (int64_t)(b + 1) * a.nb_per_body; ++s)
        for (int k = 0; k < 3; ++k) f[k] += a.fb.p[s * a.fb.ld_p + k * a.fb.ld_c];
    for (int k = 0; k < 3; ++k) net[3 * b + k] = f[k];
}

template <typename T, int MODE, bool BLAKE>
hipError_t launch_typed(const FlowArgs& a, const double* net, int block, hipStream_t s) {
    const int R = 2;  // two targets per lane share every LDS source read
    const int64_t tile = (int64_t)block * R;
    const dim3 grid((unsigned)((a.nt + tile - 1) / tile));
    switch (block) {
        case 64: hipLaunchKernelGGL((flow_kernel<T, MODE, BLAKE, 64, 2>), grid, dim3(64), 0, s, a, net); break;
        case 128: hipLaunchKernelGGL((flow_kernel<T, MODE, BLAKE, 128, 2>), grid, dim3(128), 0, s, a, net); break;
        case 256: hipLaunchKernelGGL((flow_kernel<T, MODE, BLAKE, 256, 2>), grid, dim3(256), 0, s, a, net); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_flow_net(const FlowArgs& a, double* net, int nbod, hipStream_t s) {
    if (nbod <= 0) return hipSuccess;
    hipLaunchKernelGGL(net_force_kernel, dim3((nbod + 63) / 64), dim3(64), 0, s, a, net, nbod);
    return hipGetLastError();
}

hipError_t launch_flow_eval(int precision, const FlowArgs& a, const double* net, int block, hipStream_t s) {
    if (a.nt <= 0) return hipSuccess;
    const bool blake = a.mode == 0 && a.nb_per_body == 0 && a.ns_total > 0;
    if (precision == 0) {
        if (a.mode == 1) return launch_typed<double, 1, false>(a, net, block, s);
        return blake ? launch_typed<double, 0, true>(a, net, block, s) : launch_typed<double, 0, false>(a, net, block, s);
    }
    if (a.mode == 1) return launch_typed<float, 1, false>(a, net, block, s);
    return blake ? launch_typed<float, 0, true>(a, net, block, s) : launch_typed<float, 0, false>(a, net, block, s);
}

}  // namespace tfs
