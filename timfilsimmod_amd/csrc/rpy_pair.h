// rpy_pair.h -- Rotne-Prager-Yamakawa pair couplings on the device (gfx950).
//
// Restates the unbounded branch of rpy_interaction<vel_dof, force_dof>
// (src/cilia/cuda_functions.cu:207-354) for one target i and one source j:
//
//   far field  r > ai + aj            : closed RPY forms written with the
//                                       unnormalised separation d = xi - xj
//                                       and 1/r from a refined rsqrt;
//   overlap    |ai - aj| < r <= ai+aj : Zuk et al. (2014) forms, evaluated
//                                       exactly as the reference does (r from
//                                       sqrt, 1/(r + 1e-20));
//   inner      r <= |ai - aj|         : rigid-body limit (includes r = 0).
//
// The far branch is the hot one; the other two are taken by a handful of
// neighbours per particle, so they sit behind a wave-uniform-skippable branch.
// Template parameters: T is the evaluation type (double or float), VT the
// number of velocity components written (3: v, 6: v and omega), FS the
// number of source components read (3: force, 6: force and torque).
#pragma once
#include <hip/hip_runtime.h>

namespace tfs {

// Pair-independent constants, computed once on the host for given radii and mu.
template <typename T>
struct PairConsts {
    T sum_a2;      // (ai + aj)^2          far-field threshold on r^2
    T s3;          // (ai^2 + aj^2)/3
    T c8;          // 1/(8 pi mu)
    // overlap / inner branch (evaluated in double for both evaluation types)
    double ai, aj, mu;
};

__device__ __forceinline__ double rsqrt_refined(double x) {
    // v_rsq_f64 seed (measured on gfx950: relative error <= 5.3e-8 over
    // x in [4, 4e6]), then one Newton-Raphson step y <- y + y/2 (1 - x y^2),
    // which squares it to <= 4.3e-15 -- far inside the 1e-12 FP64 bar.
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

__device__ __forceinline__ float rsqrt_refined(float x) {
    return __builtin_amdgcn_rsqf(x);
}

// Far field, r^2 > (ai+aj)^2.  rm1 = 1/r.
template <typename T, int VT, int FS>
__device__ __forceinline__ void pair_far(T* __restrict__ v, T dx, T dy, T dz, T rm1,
                                         const T* __restrict__ f, T s3, T c8) {
    const T rm2 = rm1 * rm1;
    const T c = c8 * rm1;                       // 1/(8 pi mu r)
    const T A = c + (c * s3) * rm2;             // (1 + (ai^2+aj^2)/(3 r^2)) / (8 pi mu r)
    const T B = (c - T(3) * (c * s3) * rm2) * rm2;
    const T dF = dx * f[0] + dy * f[1] + dz * f[2];
    const T Bd = B * dF;
    v[0] += A * f[0] + Bd * dx;
    v[1] += A * f[1] + Bd * dy;
    v[2] += A * f[2] + Bd * dz;
    if constexpr (FS == 6 || VT == 6) {
        const T C = c * rm2;                    // 1/(8 pi mu r^3)
        if constexpr (FS == 6) {
            // translation-rotation: v += C (tau x d)
            v[0] += C * (f[4] * dz - f[5] * dy);
            v[1] += C * (f[5] * dx - f[3] * dz);
            v[2] += C * (f[3] * dy - f[4] * dx);
        }
        if constexpr (VT == 6) {
            // rotation-translation: w += C (F x d)
            v[3] += C * (f[1] * dz - f[2] * dy);
            v[4] += C * (f[2] * dx - f[0] * dz);
            v[5] += C * (f[0] * dy - f[1] * dx);
            if constexpr (FS == 6) {
                // rotation-rotation: w += -C/2 [tau - 3 (tau.d) d / r^2]
                const T Ar = T(-0.5) * C;
                const T Br = T(-3) * Ar * rm2 * (dx * f[3] + dy * f[4] + dz * f[5]);
                v[3] += Ar * f[3] + Br * dx;
                v[4] += Ar * f[4] + Br * dy;
                v[5] += Ar * f[5] + Br * dz;
            }
        }
    }
}

__device__ __forceinline__ float fmax_t(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double fmax_t(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Far field with the source strengths already multiplied by 1/(8 pi mu), in
// the fewest VALU operations: A = 1/r + s3/r^3, B = (1 - 3 s3/r^2)/r^3,
// C = 1/r^3, and every term is fused into the accumulators.  The same tensor
// as pair_far.
// fneg (FP32 only, else unused): the negated torque (FS = 6) followed by the
// negated force (VT = 6) of the source, so that the cross products are
// accumulates.
template <typename T, int VT, int FS>
__device__ __forceinline__ void pair_far_scaled(T* __restrict__ v, T dx, T dy, T dz, T rm1,
                                                const T* __restrict__ f, const T* __restrict__ fneg, T s3,
                                                T m3s3) {
    // Written so that every operation is a two-source multiply or an
    // accumulate into a dying value (VOP2 forms, the double-rate FP32 issue on
    // gfx950): A accumulates into 1/r, B into 1/r^3.  FP32 callers pass
    // d, 1/r, s3 in a frame scaled by up to 2^31 (fixed-point coordinates),
    // where 1/r^5 would underflow: there B = r^-3 + (m3s3 r^-2) r^-3, and the
    // rotation-rotation term multiplies (tau.d) r^-2 before r^-3.
    constexpr bool F32 = sizeof(T) == 4;
    const T rm2 = rm1 * rm1;
    const T rm3 = rm1 * rm2;
    const T A = fma_t(s3, rm3, rm1);
    T B;
    if constexpr (F32) {
        const T u = m3s3 * rm2;
        B = fma_t(u, rm3, rm3);
    } else {
        const T rm5 = rm3 * rm2;
        B = fma_t(m3s3, rm5, rm3);
    }
    const T Bd = B * fma_t(dz, f[2], fma_t(dy, f[1], dx * f[0]));
    v[0] = fma_t(Bd, dx, fma_t(A, f[0], v[0]));
    v[1] = fma_t(Bd, dy, fma_t(A, f[1], v[1]));
    v[2] = fma_t(Bd, dz, fma_t(A, f[2], v[2]));
    if constexpr (FS == 6) {
        // translation-rotation: v += C (tau x d)
        if constexpr (F32) {
            v[0] = fma_t(rm3, fma_t(fneg[2], dy, f[4] * dz), v[0]);
            v[1] = fma_t(rm3, fma_t(fneg[0], dz, f[5] * dx), v[1]);
            v[2] = fma_t(rm3, fma_t(fneg[1], dx, f[3] * dy), v[2]);
        } else {
            v[0] = fma_t(rm3, f[4] * dz - f[5] * dy, v[0]);
            v[1] = fma_t(rm3, f[5] * dx - f[3] * dz, v[1]);
            v[2] = fma_t(rm3, f[3] * dy - f[4] * dx, v[2]);
        }
    }
    if constexpr (VT == 6) {
        // rotation-translation: w += C (F x d)
        if constexpr (F32) {
            const T* nf = fneg + (FS == 6 ? 3 : 0);
            v[3] = fma_t(rm3, fma_t(nf[2], dy, f[1] * dz), v[3]);
            v[4] = fma_t(rm3, fma_t(nf[0], dz, f[2] * dx), v[4]);
            v[5] = fma_t(rm3, fma_t(nf[1], dx, f[0] * dy), v[5]);
        } else {
            v[3] = fma_t(rm3, f[1] * dz - f[2] * dy, v[3]);
            v[4] = fma_t(rm3, f[2] * dx - f[0] * dz, v[4]);
            v[5] = fma_t(rm3, f[0] * dy - f[1] * dx, v[5]);
        }
        if constexpr (FS == 6) {
            // rotation-rotation: w += -C/2 tau + (3C/2) (tau.d) d / r^2
            const T Ar = T(-0.5) * rm3;
            const T Br = T(1.5) * rm3 * (rm2 * fma_t(dz, f[5], fma_t(dy, f[4], dx * f[3])));
            v[3] = fma_t(Br, dx, fma_t(Ar, f[3], v[3]));
            v[4] = fma_t(Br, dy, fma_t(Ar, f[4], v[4]));
            v[5] = fma_t(Br, dz, fma_t(Ar, f[5], v[5]));
        }
    }
}

// Overlap and inner branches: a term-by-term restatement of
// cuda_functions.cu:207-354, in double.  Accumulates into v.
template <int VT, int FS>
__device__ __forceinline__ void pair_near(double* __restrict__ v, double dx, double dy, double dz,
                                       const double* __restrict__ f, double ai, double aj,
                                       double mu) {
    const double PI = 3.14159265358979323846;
    const double r = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    const double inv_r = 1.0 / (r + 1e-20);
    const double da = __builtin_fabs(ai - aj);
    const double ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
    double A, B;
    const double e_dot = ex * f[0] + ey * f[1] + ez * f[2];
    if (r > ai + aj) {
        const double s = inv_r * inv_r * (ai * ai + aj * aj) / 3.0;
        const double c = inv_r / (8.0 * PI * mu);
        A = (1.0 + s) * c;
        B = (1.0 - 3.0 * s) * c * e_dot;
    } else if (r > da) {
        const double den = 32.0 * r * r * r;
        const double p = da * da + 3.0 * r * r;
        const double q = da * da - r * r;
        const double c = 1.0 / (6.0 * PI * mu * ai * aj);
        A = (0.5 * (ai + aj) - p * p / den) * c;
        B = (3.0 * q * q / den) * c * e_dot;
    } else {
        A = 1.0 / (6.0 * PI * mu * (ai > aj ? ai : aj));
        B = 0.0;
    }
    v[0] += A * f[0] + B * ex;
    v[1] += A * f[1] + B * ey;
    v[2] += A * f[2] + B * ez;
    if constexpr (FS == 6) {
        double C;
        if (r > ai + aj) {
            C = inv_r * inv_r / (8.0 * PI * mu);
        } else if (r > da) {
            const double g = aj - ai + r;
            C = inv_r * inv_r * g * g * (ai * ai + 2.0 * ai * (r + aj) - 3.0 * (aj - r) * (aj - r))
                / (128.0 * PI * mu * ai * aj * aj * aj);
        } else {
            C = (aj > ai ? 1.0 : 0.0) * r / (8.0 * PI * mu * aj * aj * aj);
        }
        v[0] += C * (f[4] * ez - f[5] * ey);
        v[1] += C * (f[5] * ex - f[3] * ez);
        v[2] += C * (f[3] * ey - f[4] * ex);
    }
    if constexpr (VT == 6) {
        double C;
        if (r > ai + aj) {
            C = inv_r * inv_r / (8.0 * PI * mu);
        } else if (r > da) {
            const double g = ai - aj + r;
            C = inv_r * inv_r * g * g * (aj * aj + 2.0 * aj * (r + ai) - 3.0 * (ai - r) * (ai - r))
                / (128.0 * PI * mu * aj * ai * ai * ai);
        } else {
            C = (ai > aj ? 1.0 : 0.0) * r / (8.0 * PI * mu * ai * ai * ai);
        }
        v[3] += C * (f[1] * ez - f[2] * ey);
        v[4] += C * (f[2] * ex - f[0] * ez);
        v[5] += C * (f[0] * ey - f[1] * ex);
        if constexpr (FS == 6) {
            const double t_dot = ex * f[3] + ey * f[4] + ez * f[5];
            double Ar, Br;
            if (r > ai + aj) {
                Ar = -inv_r * inv_r * inv_r / (16.0 * PI * mu);
                Br = -3.0 * t_dot * Ar;
            } else if (r > da) {
                const double r2 = r * r, r3 = r2 * r;
                const double ai2 = ai * ai, aj2 = aj * aj;
                const double d4 = da * da * da * da;
                const double na = 5.0 * r3 * r3 - 27.0 * r2 * r2 * (ai2 + aj2)
                                + 32.0 * r3 * (ai2 * ai + aj2 * aj)
                                - 9.0 * r2 * (ai2 - aj2) * (ai2 - aj2)
                                - d4 * (ai2 + aj2 + 4.0 * ai * aj);
                const double w = (ai - aj) * (ai - aj) - r2;
                const double nb = 3.0 * w * w * (ai2 + aj2 + 4.0 * ai * aj - r2);
                const double c = 1.0 / (512.0 * PI * mu * r3 * ai2 * ai * aj2 * aj);
                Ar = na * c;
                Br = nb * c * t_dot;
            } else {
                const double am = ai > aj ? ai : aj;
                Ar = 1.0 / (8.0 * PI * mu * am * am * am);
                Br = 0.0;
            }
            v[3] += Ar * f[3] + Br * ex;
            v[4] += Ar * f[4] + Br * ey;
            v[5] += Ar * f[5] + Br * ez;
        }
    }
}

// Half-space pair tensor (no-slip wall at z = 0; cuda_functions.cu:12-203,
// Swan & Brady 2007, radius 1 and viscosity 1 as the reference assumes),
// evaluated in T (double, or float as the reference's Real = float build
// does).  Accumulates into v the velocity (VT = 3) or velocity and angular
// velocity (VT = 6) of particle i due to the force (FS = 3) or force and
// torque (FS = 6) f on particle j, given their separation (dx, dy, dz) =
// xi - xj and heights zi, zj above the wall; `same` marks i and j as one
// particle (the wall self-mobility).  Distinct particles take the unbounded
// far-field form plus the image corrections; as in the reference, the
// rotation-translation wall terms enter with the opposite sign to the
// paper's printed alternating-tensor convention.
template <typename T, int VT, int FS>
__device__ inline void pair_wall(T* __restrict__ v, bool same, T dx, T dy, T dz, T zi, T zj,
                                 const T* __restrict__ f) {
    const double PI = 3.14159265358979323846;
    const T k6 = T(1.0 / (6.0 * PI)), k8 = T(1.0 / (8.0 * PI));
    if (same) {
        const T g1 = T(1) / zi, g3 = g1 * g1 * g1, g4 = g3 * g1, g5 = g4 * g1;
        const T tpar = k6 - (T(9.0) * g1 - T(2.0) * g3 + g5) * (k6 / T(16.0));
        const T tperp = k6 - (T(9.0) * g1 - T(4.0) * g3 + g5) * (k6 / T(8.0));
        const T cpl = g4 * (k6 * T(3.0) / T(32.0));
        v[0] += tpar * f[0];
        v[1] += tpar * f[1];
        v[2] += tperp * f[2];
        if constexpr (FS == 6) {
            v[0] += cpl * f[4];
            v[1] -= cpl * f[3];
        }
        if constexpr (VT == 6) {
            v[3] -= cpl * f[1];
            v[4] += cpl * f[0];
            if constexpr (FS == 6) {
                const T rpar = k8 - g3 * (k6 * T(15.0) / T(64.0));
                const T rperp = k8 - g3 * (k6 * T(3.0) / T(32.0));
                v[3] += rpar * f[3];
                v[4] += rpar * f[4];
                v[5] += rperp * f[5];
            }
        }
        return;
    }
    // unbounded far field between the particles
    {
        const T q = rsqrt_refined(dx * dx + dy * dy + dz * dz);
        const T ux = dx * q, uy = dy * q, uz = dz * q;
        const T uf = ux * f[0] + uy * f[1] + uz * f[2];
        const T ca = k8 * q, cb = T(1.0) + (T(2.0) / T(3.0)) * q * q, cc = (T(1.0) - T(2.0) * q * q) * uf;
        v[0] += ca * (cb * f[0] + cc * ux);
        v[1] += ca * (cb * f[1] + cc * uy);
        v[2] += ca * (cb * f[2] + cc * uz);
        const T c2 = ca * q;
        if constexpr (FS == 6) {
            v[0] += c2 * (f[4] * uz - f[5] * uy);
            v[1] += c2 * (f[5] * ux - f[3] * uz);
            v[2] += c2 * (f[3] * uy - f[4] * ux);
        }
        if constexpr (VT == 6) {
            v[3] += c2 * (f[1] * uz - f[2] * uy);
            v[4] += c2 * (f[2] * ux - f[0] * uz);
            v[5] += c2 * (f[0] * uy - f[1] * ux);
            if constexpr (FS == 6) {
                const T c3 = T(0.5) * c2 * q, ut = ux * f[3] + uy * f[4] + uz * f[5];
                v[3] += c3 * (T(3.0) * ut * ux - f[3]);
                v[4] += c3 * (T(3.0) * ut * uy - f[4]);
                v[5] += c3 * (T(3.0) * ut * uz - f[5]);
            }
        }
    }
    // image corrections: separation from the image of j at (xj, yj, -zj)
    const T Rx = dx, Ry = dy, Rz = zi + zj;
    const T p1 = rsqrt_refined(Rx * Rx + Ry * Ry + Rz * Rz);
    const T p2 = p1 * p1, p3 = p2 * p1, p4 = p3 * p1, p5 = p4 * p1;
    const T ex = Rx * p1, ey = Ry * p1, ez = Rz * p1, ez2 = ez * ez;
    const T h = zj / Rz;
    const T ef = ex * f[0] + ey * f[1] + ez * f[2];
    {
        const T t0 = -(T(0.75) * p1 * (T(1.0) + T(2.0) * h * (T(1.0) - h) * ez2) +
                            T(0.5) * (p3 * (T(1.0) - T(3.0) * ez2) - p5 * (T(1.0) - T(5.0) * ez2)));
        const T t12 = -(T(0.75) * p1 * (T(1.0) - T(6.0) * h * (T(1.0) - h) * ez2) - T(1.5) * p3 * (T(1.0) - T(5.0) * ez2) +
                             T(2.5) * p5 * (T(1.0) - T(7.0) * ez2)) * ef +
                           ez * f[2] *
                               (T(1.5) * h * p1 * (T(1.0) - T(6.0) * (T(1.0) - h) * ez2) - T(3.0) * p3 * (T(1.0) - T(5.0) * ez2) +
                                T(5.0) * p5 * (T(2.0) - T(7.0) * ez2));
        const T t3 = (T(1.5) * h * p1 - T(5.0) * p5) * ef;
        const T t4 = -((T(3.0) * h * h * p1 + T(3.0) * p3) * ez2 + (T(2.0) - T(15.0) * ez2) * p5) * f[2];
        v[0] += k6 * (t0 * f[0] + t12 * ex);
        v[1] += k6 * (t0 * f[1] + t12 * ey);
        v[2] += k6 * (t0 * f[2] + (t12 + t3) * ez + t4);
    }
    if constexpr (VT == 6) {
        const T u0 = T(0.75) * p2;
        const T u12 = f[2] * (T(9.0) * h * ez2 * p2 + (T(1.5) - T(15.0) * ez2) * p4) - ez * ef * (T(4.5) * h * p2 - T(7.5) * p4);
        const T u3 = -T(1.5) * ez * (h * p2 - p4);
        v[3] -= k6 * (u0 * (f[1] * ez - f[2] * ey) - u12 * ey + u3 * f[1]);
        v[4] -= k6 * (u0 * (f[2] * ex - f[0] * ez) + u12 * ex - u3 * f[0]);
        v[5] -= k6 * u0 * (f[0] * ey - f[1] * ex);
        if constexpr (FS == 6) {
            const T et = ex * f[3] + ey * f[4] + ez * f[5];
            const T w0 = p3 * (T(0.375) - T(2.25) * ez2), w1 = -T(1.125) * p3 * et, w2 = T(2.25) * ez * p3 * et;
            const T w3 = T(2.25) * p3 * (ex * f[4] - ey * f[3]);
            v[3] += k6 * (w0 * f[3] + w1 * ex - w3 * ey);
            v[4] += k6 * (w0 * f[4] + w1 * ey + w3 * ex);
            v[5] += k6 * (w0 * f[5] + w1 * ez + w2);
        }
    }
    if constexpr (FS == 6) {
        // translation from torque: transpose of the rotation-from-force block
        // seen from j (image weight 1 - h, e_x and e_y reversed)
        const T hh = T(1.0) - h, fx = -ex, fy = -ey;
        const T ext = fx * f[4] - fy * f[3];
        const T q0 = T(0.75) * p2;
        const T q1 = (T(9.0) * hh * ez2 * p2 + (T(1.5) - T(15.0) * ez2) * p4) * ext;
        const T q2 = -ez * (T(4.5) * hh * p2 - T(7.5) * p4) * ext;
        const T q3 = -T(1.5) * ez * (hh * p2 - p4);
        v[0] -= k6 * (q0 * (f[5] * fy - f[4] * ez) + q2 * fx - q3 * f[4]);
        v[1] -= k6 * (q0 * (f[3] * ez - f[5] * fx) + q2 * fy + q3 * f[3]);
        v[2] -= k6 * (q0 * (f[4] * fx - f[3] * fy) + q2 * ez + q1);
    }
}

}  // namespace tfs
