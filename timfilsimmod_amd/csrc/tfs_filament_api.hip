// tfs_filament_api.hip -- the Cosserat filament model on the device
// (include/tfs_filament.h).
//
// The reference runs these per filament in host loops over `matrix` objects,
// between every pair of mobility solves of a Broyden iteration
// (cilia_sim_main.cu:394-502).  Here all filaments of all swimmers go at once,
// in FP64: forces one thread per segment (each adds the two links it touches,
// so no atomics and a fixed summation order), the approximate Jacobian one
// workgroup per filament (each thread owns the position and Lie-algebra rows
// of one segment), the sequential robot arm one thread per filament.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/tfs_broyden.h"
#include "../../include/tfs_filament.h"
#include "../../include/tfs_solver.h"
#include "tfs_api_internal.h"

using namespace tfs;

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Q {
    double w, x, y, z;
};

__device__ __forceinline__ Q qmul(const Q& p, const Q& q) {
    return Q{p.w * q.w - (p.x * q.x + p.y * q.y + p.z * q.z), p.w * q.x + q.w * p.x + p.y * q.z - p.z * q.y,
             p.w * q.y + q.w * p.y + p.z * q.x - p.x * q.z, p.w * q.z + q.w * p.z + p.x * q.y - p.y * q.x};
}

__device__ __forceinline__ Q qconj(const Q& q) { return Q{q.w, -q.x, -q.y, -q.z}; }

// Principal square root (quaternion.cpp:245-270).
__device__ __forceinline__ Q qsqrt(const Q& q) {
    if (q.w < -0.999999999) return Q{0.0, 0.0, 0.0, 1.0};
    const double s = sqrt(0.5 * (1.0 + q.w));
    const double d = 2.0 * s;
    return Q{s, q.x / d, q.y / d, q.z / d};
}

// (q2 q1*)^(1/2) q1 (quaternion.cpp:555-564).
__device__ __forceinline__ Q midpoint(const Q& q1, const Q& q2) { return qmul(qsqrt(qmul(q2, qconj(q1))), q1); }

// exp map of the Lie algebra (quaternion.cpp:522-545).
__device__ __forceinline__ Q lie_exp(const double* u) {
    const double th = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (th < 1e-6) return Q{1.0, 0.0, 0.0, 0.0};
    const double sn = sin(0.5 * th) / th;
    return Q{cos(0.5 * th), sn * u[0], sn * u[1], sn * u[2]};
}

__device__ __forceinline__ Q load_q(const double* p) { return Q{p[0], p[1], p[2], p[3]}; }
__device__ __forceinline__ void store_q(double* p, const Q& q) {
    p[0] = q.w;
    p[1] = q.x;
    p[2] = q.y;
    p[3] = q.z;
}

// Columns of the rotation matrix: tangent, normal, binormal (quaternion.cpp:285-307).
__device__ __forceinline__ void tangent(const Q& q, double* t) {
    t[0] = 1.0 - 2.0 * (q.y * q.y + q.z * q.z);
    t[1] = 2.0 * (q.x * q.y + q.w * q.z);
    t[2] = 2.0 * (q.x * q.z - q.w * q.y);
}
__device__ __forceinline__ void normal(const Q& q, double* n) {
    n[0] = 2.0 * (q.x * q.y - q.w * q.z);
    n[1] = 1.0 - 2.0 * (q.x * q.x + q.z * q.z);
    n[2] = 2.0 * (q.y * q.z + q.w * q.x);
}
__device__ __forceinline__ void binormal(const Q& q, double* b) {
    b[0] = 2.0 * (q.x * q.z + q.w * q.y);
    b[1] = 2.0 * (q.y * q.z - q.w * q.x);
    b[2] = 1.0 - 2.0 * (q.x * q.x + q.y * q.y);
}

// 3 x 3 row-major helpers.
struct M3 {
    double a[9];
};
__device__ __forceinline__ M3 zero3() {
    M3 m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.a[i] = 0.0;
    return m;
}
// rcross(v): M x = x cross v (matrix.cpp:665-679).
__device__ __forceinline__ M3 rcross(const double* v) {
    M3 m;
    m.a[0] = 0.0;
    m.a[1] = v[2];
    m.a[2] = -v[1];
    m.a[3] = -v[2];
    m.a[4] = 0.0;
    m.a[5] = v[0];
    m.a[6] = v[1];
    m.a[7] = -v[0];
    m.a[8] = 0.0;
    return m;
}
__device__ __forceinline__ M3 mm(const M3& A, const M3& B) {
    M3 C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C.a[3 * i + j] = A.a[3 * i] * B.a[j] + A.a[3 * i + 1] * B.a[3 + j] + A.a[3 * i + 2] * B.a[6 + j];
    return C;
}
__device__ __forceinline__ M3 axpy(double s, const M3& A, const M3& B) {  // s A + B
    M3 C;
#pragma unroll
    for (int i = 0; i < 9; ++i) C.a[i] = s * A.a[i] + B.a[i];
    return C;
}
__device__ __forceinline__ M3 scale(double s, const M3& A) {
    M3 C;
#pragma unroll
    for (int i = 0; i < 9; ++i) C.a[i] = s * A.a[i];
    return C;
}
__device__ __forceinline__ void mv(const M3& A, const double* v, double* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = A.a[3 * i] * v[0] + A.a[3 * i + 1] * v[1] + A.a[3 * i + 2] * v[2];
}
// Rotation matrix of q (columns t, n, b).
__device__ __forceinline__ M3 rot(const Q& q) {
    double t[3], n[3], b[3];
    tangent(q, t);
    normal(q, n);
    binormal(q, b);
    M3 R;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R.a[3 * i] = t[i];
        R.a[3 * i + 1] = n[i];
        R.a[3 * i + 2] = b[i];
    }
    return R;
}

__device__ __forceinline__ double dexpinv_fac(const double* u) {
    const double th2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double th = sqrt(th2);
    if (th < 1e-5) return -1.0 / 12.0;
    return (0.5 * th * cos(0.5 * th) / sin(0.5 * th) - 1.0) / th2;
}

__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// dexpinv(u)^T v (quaternion.cpp:635-660).
__device__ __forceinline__ void dexpinv_transpose(const double* u, const double* v, double* out) {
    const double fac = dexpinv_fac(u);
    double c[3], cc[3];
    cross(u, v, c);
    cross(u, c, cc);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = v[i] + 0.5 * c[i] - fac * cc[i];
}

struct FilParams {
    int nfil, nseg, fil_per_swim;
    double dl, kb, kt, mu, a, dt, end_force;
    double st[3];
    int pert_a, pert_b;
    int base_mode;  // TFS_BASE_CLAMPED or TFS_BASE_DRIVEN
};

// Body-frame elastic moment of the link q1 -> q2 (filament.cpp:1451-1465).
__device__ void body_frame_moment(const Q& q1, const Q& q2, const FilParams& P, double* M) {
    const Q mid = qconj(midpoint(q1, q2));
    const double inv = 1.0 / P.dl;
    const Q d{(q2.w - q1.w) * inv, (q2.x - q1.x) * inv, (q2.y - q1.y) * inv, (q2.z - q1.z) * inv};
    const Q m = qmul(mid, d);
    M[0] = P.kt * (2.0 * m.x - P.st[0]);
    M[1] = P.kb * (2.0 * m.y - P.st[1]);
    M[2] = P.kb * (2.0 * m.z - P.st[2]);
}

// d(body-frame moment)/d(rotation of q2 (upper) or q1) (filament.cpp:9-55):
// column c from the quaternion A_c = +-psi(q) e_c / DL.
__device__ M3 moment_lie_derivative(const Q& q1, const Q& q2, bool upper, const FilParams& P) {
    const Q mid = qconj(midpoint(q1, q2));
    const double inv = 1.0 / P.dl;
    const Q dq{(q2.w - q1.w) * inv, (q2.x - q1.x) * inv, (q2.y - q1.y) * inv, (q2.z - q1.z) * inv};
    const Q qs = upper ? q2 : q1;
    const double sa = (upper ? 0.5 : -0.5) * inv;     // psi carries the factor 1/2
    const double sb = upper ? 0.5 * P.dl : -0.5 * P.dl;
    const double kscale[3] = {P.kt, P.kb, P.kb};
    M3 out;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Q e{0.0, c == 0 ? 1.0 : 0.0, c == 1 ? 1.0 : 0.0, c == 2 ? 1.0 : 0.0};
        Q A = qmul(e, qs);
        A = Q{sa * A.w, sa * A.x, sa * A.y, sa * A.z};
        const Q t1 = qmul(mid, A);
        const Q t2 = qmul(dq, Q{sb * A.w, sb * A.x, sb * A.y, sb * A.z});
        out.a[c] = 2.0 * (t1.x + t2.x) * kscale[0];
        out.a[3 + c] = 2.0 * (t1.y + t2.y) * kscale[1];
        out.a[6 + c] = 2.0 * (t1.z + t2.z) * kscale[2];
    }
    return out;
}

// K0 rcross(t) + K1 rcross(n) + K2 rcross(b) of q.
__device__ M3 frame_cross(const Q& q, const double* K) {
    double t[3], n[3], b[3];
    tangent(q, t);
    normal(q, n);
    binormal(q, b);
    double v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = K[0] * t[i];
    M3 m = rcross(v);
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = K[1] * n[i];
    M3 m2 = rcross(v);
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = K[2] * b[i];
    M3 m3 = rcross(v);
#pragma unroll
    for (int i = 0; i < 9; ++i) m.a[i] += m2.a[i] + m3.a[i];
    return m;
}

struct FilState {
    double *q, *qm1, *u, *um1, *x, *xm1, *xm2;
    double *mult, *mult_m1, *mult_m2;  // 3 (nseg + 1) per filament
    double *bases;                     // 3 per filament
    double *pert;                      // 6 per filament (may be null)
    double *clampblk;                  // 18 per filament: elastic clamping blocks 1, 2 (row-major)
    double *base_val;                  // 3 per filament (may be null): clamped base: its turn per
                                       // step; driven base: the torque on segment 0
    // A free body per swimmer (all null while the bodies are held still):
    // position at n+1, n, n-1; Lie-algebra element at n+1, n; quaternion at
    // n+1, n; Q_init = R(q) of the initial guess (3, 3, 3, 3, 3, 4, 4, 9 each).
    double *bx, *bxm1, *bxm2, *bu, *bum1, *bq, *bqm1, *bQ;
    double *brefs;                     // 3 per filament: its anchor in the body frame
};

// The anchor of filament f on its (free) body: X + R(q) ref.
__device__ void body_anchor(const FilState& S, const FilParams& P, int f, double* base) {
    const int sw = f / P.fil_per_swim;
    const M3 R = rot(load_q(S.bq + 4 * sw));
    mv(R, S.brefs + 3 * f, base);
#pragma unroll
    for (int k = 0; k < 3; ++k) base[k] += S.bx[3 * sw + k];
}

// ---- kernels ---------------------------------------------------------------
__device__ void robot_arm(const FilState& S, const FilParams& P, int f) {
    const int64_t s0 = (int64_t)f * P.nseg;
    double t1[3], t2[3];
    tangent(load_q(S.q + 4 * s0), t1);
    for (int i = 1; i < P.nseg; ++i) {
        tangent(load_q(S.q + 4 * (s0 + i)), t2);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            S.x[3 * (s0 + i) + k] = S.x[3 * (s0 + i - 1) + k] + 0.5 * P.dl * (t1[k] + t2[k]);
            t1[k] = t2[k];
        }
    }
}

__global__ void set_state_kernel(FilState S, FilParams P) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nfil) return;
    const int64_t s0 = (int64_t)f * P.nseg;
    for (int i = 0; i < P.nseg; ++i) {
        const int64_t g = s0 + i;
#pragma unroll
        for (int k = 0; k < 4; ++k) S.qm1[4 * g + k] = S.q[4 * g + k];
#pragma unroll
        for (int k = 0; k < 3; ++k) S.u[3 * g + k] = S.um1[3 * g + k] = 0.0;
    }
    for (int k = 0; k < 3; ++k) S.x[3 * s0 + k] = S.bases[3 * f + k];
    robot_arm(S, P, f);
    for (int i = 0; i < 3 * P.nseg; ++i) S.xm1[3 * s0 + i] = S.xm2[3 * s0 + i] = S.x[3 * s0 + i];
    const int nm = 3 * (P.nseg + 1);
    for (int i = 0; i < nm; ++i) S.mult[(int64_t)f * nm + i] = S.mult_m1[(int64_t)f * nm + i] = S.mult_m2[(int64_t)f * nm + i] = 0.0;
}

// filament::initial_guess for a body held still.
__global__ void initial_guess_kernel(FilState S, FilParams P, int nt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nfil) return;
    const int64_t s0 = (int64_t)f * P.nseg;
    for (int i = 0; i < P.nseg; ++i) {
        const int64_t g = s0 + i;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            S.xm2[3 * g + k] = S.xm1[3 * g + k];
            S.xm1[3 * g + k] = S.x[3 * g + k];
        }
        if (nt <= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) S.um1[3 * g + k] = S.u[3 * g + k];
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double t = 2.0 * S.u[3 * g + k] - S.um1[3 * g + k];
                S.um1[3 * g + k] = S.u[3 * g + k];
                S.u[3 * g + k] = t;
            }
        }
        const Q q = load_q(S.q + 4 * g);
        store_q(S.qm1 + 4 * g, q);
        if (i == 0 && !S.bx && P.base_mode == TFS_BASE_DRIVEN) {
            // a driven base segment is free: its u was extrapolated above
            store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), q));
        } else if (i == 0 && !S.bx && S.base_val) {
            // clamped, turned by a fixed element every step
#pragma unroll
            for (int k = 0; k < 3; ++k) S.u[3 * g + k] = S.base_val[3 * f + k];
            store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), q));
        } else if (i == 0 && !S.bx) {
            // clamped to a body that does not turn: u_0 = 0, q_0 unchanged
#pragma unroll
            for (int k = 0; k < 3; ++k) S.u[3 * g + k] = 0.0;
        } else if (i == 0) {
            // accept_state_from_rigid_body (filament.cpp:649-672): the base
            // segment turns with the body, u_0 = u_body
            const int sw = f / P.fil_per_swim;
#pragma unroll
            for (int k = 0; k < 3; ++k) S.u[3 * g + k] = S.bu[3 * sw + k];
            store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), q));
        } else {
            store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), q));
        }
    }
    if (S.bx) {
        body_anchor(S, P, f, S.x + 3 * s0);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) S.x[3 * s0 + k] = S.bases[3 * f + k];
    }
    const int nm = 3 * (P.nseg + 1);
    double* m = S.mult + (int64_t)f * nm;
    double* m1 = S.mult_m1 + (int64_t)f * nm;
    double* m2 = S.mult_m2 + (int64_t)f * nm;
    for (int i = 0; i < nm; ++i) {
        if (nt == 1) {
            m1[i] = m[i];
        } else if (nt == 2) {
            m2[i] = m1[i];
            m1[i] = m[i];
            m[i] = 2.0 * m1[i] - m2[i];
        } else if (nt > 2) {
            const double t = 3.0 * (m[i] - m1[i]) + m2[i];
            m2[i] = m1[i];
            m1[i] = m[i];
            m[i] = t;
        }
    }
    robot_arm(S, P, f);
}

// filament::internal_forces_and_torques, one thread per segment.
__global__ void forces_kernel(FilState S, FilParams P, int nt, double* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)P.nfil * P.nseg) return;
    const int f = (int)(g / P.nseg), i = (int)(g % P.nseg), N = P.nseg;
    const int nm = 3 * (N + 1);
    const double* mult = S.mult + (int64_t)f * nm;
    const Q qi = load_q(S.q + 4 * g);
    double t[3];
    tangent(qi, t);
    double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    if (i > 0) {  // link i-1 -> i
        const double* lam = mult + 3 * i;
        const Q qp = load_q(S.q + 4 * (g - 1));
        double c[3], Mb[3], M[3];
        cross(t, lam, c);
        body_frame_moment(qp, qi, P, Mb);
        mv(rot(midpoint(qp, qi)), Mb, M);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            F[k] += lam[k];
            T[k] += -0.5 * P.dl * c[k] - M[k];
        }
    }
    if (i + 1 < N) {  // link i -> i+1
        const double* lam = mult + 3 * (i + 1);
        const Q qn = load_q(S.q + 4 * (g + 1));
        double c[3], Mb[3], M[3];
        cross(t, lam, c);
        body_frame_moment(qi, qn, P, Mb);
        mv(rot(midpoint(qi, qn)), Mb, M);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            F[k] -= lam[k];
            T[k] += -0.5 * P.dl * c[k] + M[k];
        }
    }
    if (i == 0) {
        // tether force; the clamp torque (filament.cpp:1555-1571) or the
        // driving torque of a free base (:1523-1531)
        double ct[3] = {0.0, 0.0, 0.0};
        if (P.base_mode == TFS_BASE_DRIVEN) {
            if (S.base_val)
                for (int k = 0; k < 3; ++k) ct[k] = S.base_val[3 * f + k];
        } else {
            dexpinv_transpose(S.u + 3 * g, mult + 3 * N, ct);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            F[k] += mult[k];
            T[k] += ct[k];
        }
    }
    if (i == N - 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) F[k] -= P.end_force * t[k];
    }
    if (nt == 0 && S.pert) {
        if (i == P.pert_a)
            for (int k = 0; k < 3; ++k) F[k] += S.pert[6 * f + k];
        if (i == P.pert_b)
            for (int k = 0; k < 3; ++k) F[k] += S.pert[6 * f + 3 + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out[6 * g + k] = F[k];
        out[6 * g + 3 + k] = T[k];
    }
}

__device__ __forceinline__ void put3(double* J, int64_t ld, int r, int c, const M3& B) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[(int64_t)(r + i) * ld + c + j] = B.a[3 * i + j];
}

// Approximate Jacobian of one filament (filament.cpp:1577-1894), one
// workgroup per filament; thread k owns rows 3k.. (positions) and 3N+3k..
// (Lie algebra) of segment k.
__global__ __launch_bounds__(64) void jacobian_kernel(FilState S, FilParams P, int bdf2,
                                                      const double* __restrict__ forces, double* __restrict__ blocks) {
    const int f = blockIdx.x;
    const int N = P.nseg, n6 = 6 * N;
    const int swim = f / P.fil_per_swim, fs = f % P.fil_per_swim;
    double* J = blocks + (int64_t)swim * ((int64_t)P.fil_per_swim * n6 * n6 + 36) + (int64_t)fs * n6 * n6;
    for (int64_t e = threadIdx.x; e < (int64_t)n6 * n6; e += blockDim.x) J[e] = 0.0;
    __syncthreads();
    const int64_t s0 = (int64_t)f * N;
    const int nm = 3 * (N + 1);
    const double* mult = S.mult + (int64_t)f * nm;
    const double pi = kPi, mu = P.mu, a = P.a;
    const double fac = bdf2 ? -P.dt / (9.0 * pi * mu * a) : -P.dt / (6.0 * pi * mu * a);
    const double fac2 = bdf2 ? P.dl * P.dt / (24.0 * pi * mu * a * a * a) : P.dl * P.dt / (16.0 * pi * mu * a * a * a);
    const double fac3 = bdf2 ? -P.dt / (12.0 * pi * mu * a * a * a) : -P.dt / (8.0 * pi * mu * a * a * a);
    const double tsf = bdf2 ? 2.0 * P.dt / 3.0 : P.dt;
    const double Tfac = 1.0 / (8.0 * pi * mu * a * a * a);
    M3 I3 = zero3();
    I3.a[0] = I3.a[4] = I3.a[8] = 1.0;
    const bool driven = P.base_mode == TFS_BASE_DRIVEN;
    double t0[3];
    tangent(load_q(S.q + 4 * s0), t0);
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        const int64_t g = s0 + k;
        const Q qk = load_q(S.q + 4 * g);
        double t[3];
        tangent(qk, t);
        // position rows: multipliers (diagonal fac, next -fac), then the
        // rotations of segments 1..k
        for (int c = 0; c < 3; ++c) {
            J[(int64_t)(3 * k + c) * n6 + 3 * k + c] = fac;
            if (k < N - 1) J[(int64_t)(3 * k + c) * n6 + 3 * k + 3 + c] = -fac;
        }
        for (int j = 1; j <= k; ++j) {
            double tj[3];
            tangent(load_q(S.q + 4 * (s0 + j)), tj);
            M3 B = scale(j == k ? 0.5 * P.dl : P.dl, rcross(tj));
            if (k == N - 1 && j == N - 1) B = axpy(-fac * P.end_force, rcross(tj), B);
            put3(J, n6, 3 * k, 3 * (N + j), B);
        }
        // a free base turns every later segment with its tangent
        // (filament.cpp:1789-1800)
        if (driven && k >= 1) put3(J, n6, 3 * k, 3 * N, scale(0.5 * P.dl, rcross(t0)));
        // Lie rows, multiplier columns 3k (k >= 1) and 3(k+1) (k+1 <= N-1)
        const M3 um = rcross(S.u + 3 * g);
        const M3 tm = scale(-1.0, rcross(t));
        const M3 ut = mm(um, tm);
        M3 LB = axpy(0.5, ut, tm);
        LB = axpy(1.0 / 12.0, mm(um, ut), LB);
        LB = scale(fac2, LB);
        if (k >= 1) put3(J, n6, 3 * N + 3 * k, 3 * k, LB);
        if (k + 1 <= N - 1) put3(J, n6, 3 * N + 3 * k, 3 * (k + 1), LB);
        // Lie-Lie blocks (jacobian_lie_algebra_block)
        const M3 L = scale(-1.0, rcross(S.u + 3 * g));
        double om[3];
        for (int c = 0; c < 3; ++c) om[c] = Tfac * forces[6 * g + 3 + c];
        const M3 Om = rcross(om);
        M3 elastic = zero3();
        double lamsum[3] = {0.0, 0.0, 0.0};
        if (k + 1 < N) {
            const Q qn = load_q(S.q + 4 * (g + 1));
            double Kph[3];
            body_frame_moment(qk, qn, P, Kph);
            const M3 R = rot(midpoint(qk, qn));
            elastic = axpy(0.5, frame_cross(qk, Kph), mm(R, moment_lie_derivative(qk, qn, false, P)));
            const M3 right = scale(Tfac, axpy(0.5, frame_cross(qn, Kph), mm(R, moment_lie_derivative(qk, qn, true, P))));
            M3 B = axpy(-0.5, mm(L, right), right);
            B = axpy(1.0 / 12.0, mm(L, mm(L, right)), B);
            put3(J, n6, 3 * N + 3 * k, 3 * N + 3 * k + 3, scale(-tsf, B));
            for (int c = 0; c < 3; ++c) lamsum[c] += mult[3 * (k + 1) + c];
        }
        if (k > 0) {
            const Q qp = load_q(S.q + 4 * (g - 1));
            double Kmh[3];
            body_frame_moment(qp, qk, P, Kmh);
            const M3 R = rot(midpoint(qp, qk));
            elastic = axpy(-1.0, axpy(0.5, frame_cross(qk, Kmh), mm(R, moment_lie_derivative(qp, qk, true, P))),
                           elastic);
            const M3 left =
                scale(-Tfac, axpy(0.5, frame_cross(qp, Kmh), mm(R, moment_lie_derivative(qp, qk, false, P))));
            M3 B = axpy(-0.5, mm(L, left), left);
            B = axpy(1.0 / 12.0, mm(L, mm(L, left)), B);
            // segment 1's block on the clamp column is zero (filament.cpp:1887);
            // its value before that is elastic_clamping_block2
            if (k != 1 || driven) {
                put3(J, n6, 3 * N + 3 * k, 3 * N + 3 * k - 3, scale(-tsf, B));
            } else if (S.clampblk) {
                const M3 c2 = scale(-tsf, B);
#pragma unroll
                for (int e = 0; e < 9; ++e) S.clampblk[18 * (int64_t)f + 9 + e] = c2.a[e];
            }
            for (int c = 0; c < 3; ++c) lamsum[c] += mult[3 * k + c];
        }
        if (k == 0 && !driven) {
            // the clamp multiplier replaces u_0 (filament.cpp:1866-1888)
            const M3 blk = axpy(1.0 / 12.0, mm(um, um), axpy(0.5, um, I3));
            const M3 blkT = axpy(1.0 / 12.0, mm(um, um), axpy(-0.5, um, I3));
            put3(J, n6, 3 * N, 3 * N, scale(fac3, mm(blk, blkT)));
        }
        if (k > 0 || S.clampblk) {
            // the Lie-algebra diagonal block; for segment 0 it is overwritten
            // above and kept as elastic_clamping_block1
            const M3 constraint = scale(-0.5 * P.dl, mm(rcross(lamsum), rcross(t)));
            const M3 dOm = scale(Tfac, axpy(1.0, elastic, constraint));
            double ou[3];
            mv(Om, S.u + 3 * g, ou);
            M3 B = axpy(-0.5, axpy(1.0, Om, mm(L, dOm)), dOm);
            const M3 hi = axpy(1.0, rcross(ou), axpy(1.0, mm(L, Om), mm(L, mm(L, dOm))));
            B = axpy(1.0 / 12.0, hi, B);
            const M3 diag = axpy(-tsf, B, I3);
            if (k > 0 || driven) {
                put3(J, n6, 3 * N + 3 * k, 3 * N + 3 * k, diag);
            } else {
#pragma unroll
                for (int e = 0; e < 9; ++e) S.clampblk[18 * (int64_t)f + e] = diag.a[e];
            }
        }
    }
}

// filament::update (filament.cpp:1896-1930), one thread per segment: the
// segment's Lie-algebra element and orientation, and its three multiplier
// entries (segment 0 also the clamp's); the robot arm follows in
// robot_arm_kernel.
__global__ void update_kernel(FilState S, FilParams P, const double* __restrict__ du,
                              const int32_t* __restrict__ stop) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)P.nfil * P.nseg || (stop && *stop)) return;
    const int N = P.nseg, n6 = 6 * N;
    const int f = (int)(g / N), i = (int)(g % N);
    const int swim = f / P.fil_per_swim, fs = f % P.fil_per_swim;
    const double* d = du + (int64_t)swim * ((int64_t)P.fil_per_swim * n6 + 6) + (int64_t)fs * n6;
    const int nm = 3 * (N + 1);
    double* m = S.mult + (int64_t)f * nm;
#pragma unroll
    for (int k = 0; k < 3; ++k) m[3 * i + k] += d[3 * i + k];   // tether (i = 0) or lambda_{i-1}
    if (i == 0 && P.base_mode == TFS_BASE_DRIVEN) {
        // a free base: u_0 takes the clamp's place (filament.cpp:1918-1927)
#pragma unroll
        for (int k = 0; k < 3; ++k) S.u[3 * g + k] += d[3 * N + k];
        store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), load_q(S.qm1 + 4 * g)));
        return;
    }
    if (i == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) m[3 * N + k] += d[3 * N + k];   // clamp
        if (S.bx) {
            // swimmer::update: the anchor and base rotation follow the body
            // updated just before (accept_state_from_rigid_body)
            const int sw = f / P.fil_per_swim;
#pragma unroll
            for (int k = 0; k < 3; ++k) S.u[3 * g + k] = S.bu[3 * sw + k];
            store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), load_q(S.qm1 + 4 * g)));
            body_anchor(S, P, f, S.x + 3 * g);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) S.u[3 * g + k] += d[3 * (N + i) + k];
    store_q(S.q + 4 * g, qmul(lie_exp(S.u + 3 * g), load_q(S.qm1 + 4 * g)));
}

__global__ void robot_arm_kernel(FilState S, FilParams P) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nfil) return;
    robot_arm(S, P, f);
}

// The robot arm for filaments of at most 64 segments: one wave per filament,
// lane i the segment's step 0.5 DL (t_{i-1} + t_i), the positions x_0 plus an
// inclusive scan of the steps over the lanes.
__global__ __launch_bounds__(64) void robot_arm_wave_kernel(FilState S, FilParams P) {
    const int f = blockIdx.x, i = threadIdx.x, N = P.nseg;
    const int64_t s0 = (int64_t)f * N;
    double t[3] = {0.0, 0.0, 0.0};
    if (i < N) tangent(load_q(S.q + 4 * (s0 + i)), t);
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double tp = __shfl_up(t[k], 1, 64);
        d[k] = i >= 1 && i < N ? 0.5 * P.dl * (tp + t[k]) : 0.0;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double v = __shfl_up(d[k], off, 64);
            d[k] += i >= off ? v : 0.0;
        }
    if (i >= 1 && i < N) {
#pragma unroll
        for (int k = 0; k < 3; ++k) S.x[3 * (s0 + i) + k] = S.x[3 * s0 + k] + d[k];
    }
}

// rigid_body::initial_guess (rigid_body.cpp:395-447), one thread per body.
__global__ void body_guess_kernel(FilState S, int nswim, int nt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nswim) return;
    double* x = S.bx + 3 * b;
    double* xm1 = S.bxm1 + 3 * b;
    double* xm2 = S.bxm2 + 3 * b;
    double* u = S.bu + 3 * b;
    double* um1 = S.bum1 + 3 * b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double gx = nt <= 1 ? 2.0 * x[k] - xm1[k] : 3.0 * (x[k] - xm1[k]) + xm2[k];
        const double gu = nt <= 1 ? u[k] : 2.0 * u[k] - um1[k];
        xm2[k] = xm1[k];
        xm1[k] = x[k];
        x[k] = gx;
        um1[k] = u[k];
        u[k] = gu;
    }
    const Q q = load_q(S.bq + 4 * b);
    store_q(S.bqm1 + 4 * b, q);
    const Q qn = qmul(lie_exp(u), q);
    store_q(S.bq + 4 * b, qn);
    const M3 R = rot(qn);
#pragma unroll
    for (int e = 0; e < 9; ++e) S.bQ[9 * b + e] = R.a[e];
}

// The blob forces' extrapolation at a step's start (rigid_body.cpp:447-478).
__global__ void blob_force_guess_kernel(double* __restrict__ fb, double* __restrict__ m1, double* __restrict__ m2,
                                        int64_t n, int nt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (nt == 1) {
        m1[i] = fb[i];
    } else if (nt == 2) {
        m2[i] = m1[i];
        m1[i] = fb[i];
        fb[i] = 2.0 * m1[i] - m2[i];
    } else if (nt > 2) {
        const double t = 3.0 * (fb[i] - m1[i]) + m2[i];
        m2[i] = m1[i];
        m1[i] = fb[i];
        fb[i] = t;
    }
}

// rigid_body::update's blob part (rigid_body.cpp:506-520): swimmer s's blob
// forces take the entries after its body's 6 in the Broyden update.
__global__ void blob_force_update_kernel(double* __restrict__ fb, const double* __restrict__ du, int64_t nb3,
                                         int nswim, int64_t per, int64_t off, const int32_t* __restrict__ stop) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb3 * nswim || (stop && *stop)) return;
    const int64_t s = i / nb3, k = i % nb3;
    fb[i] += du[s * per + off + k];
}

// Q_init = R(q) of the current pose, one thread per body.
__global__ void body_rot_kernel(FilState S, int nswim) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nswim) return;
    const M3 R = rot(load_q(S.bq + 4 * b));
#pragma unroll
    for (int e = 0; e < 9; ++e) S.bQ[9 * b + e] = R.a[e];
}

// rigid_body::update (rigid_body.cpp:506-535) with the body's 6 entries of the
// Broyden update (after its filaments' blocks).
__global__ void body_update_kernel(FilState S, FilParams P, int nswim, const double* __restrict__ du,
                                   const int32_t* __restrict__ stop) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nswim || (stop && *stop)) return;
    const int64_t n6 = 6LL * P.nseg;
    const double* d = du + (int64_t)b * (P.fil_per_swim * n6 + 6) + P.fil_per_swim * n6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S.bx[3 * b + k] += d[k];
        S.bu[3 * b + k] += d[3 + k];
    }
    store_q(S.bq + 4 * b, qmul(lie_exp(S.bu + 3 * b), load_q(S.bqm1 + 4 * b)));
}

// rigid_body::blob_positions (rigid_body.cpp:482-500): X + R(q) ref per blob.
__global__ void blob_positions_kernel(FilState S, const double* __restrict__ refs, int64_t nblob, int blobs_per_swim,
                                      double* __restrict__ xb) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nblob) return;
    const int b = (int)(m / blobs_per_swim);
    const M3 R = rot(load_q(S.bq + 4 * b));
    double p[3];
    mv(R, refs + 3 * m, p);
#pragma unroll
    for (int k = 0; k < 3; ++k) xb[3 * m + k] = S.bx[3 * b + k] + p[k];
}

// The body mobility at the pose of the initial guess: Q N_ref Q^T per 3 x 3
// block (swimmer.cpp:808-813), one thread per body.
__global__ void rotate_mobility_kernel(FilState S, int nswim, const double* __restrict__ n_ref,
                                       double* __restrict__ n_rot) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nswim) return;
    M3 Q;
#pragma unroll
    for (int e = 0; e < 9; ++e) Q.a[e] = S.bQ[9 * b + e];
    M3 Qt;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Qt.a[3 * r + c] = Q.a[3 * c + r];
    for (int br = 0; br < 2; ++br)
        for (int bc = 0; bc < 2; ++bc) {
            M3 blk;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) blk.a[3 * r + c] = n_ref[36 * b + 6 * (3 * br + r) + 3 * bc + c];
            const M3 o = mm(Q, mm(blk, Qt));
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) n_rot[36 * b + 6 * (3 * br + r) + 3 * bc + c] = o.a[3 * r + c];
        }
}

__device__ __forceinline__ M3 block_of(const double* n6x6, int br, int bc) {
    M3 m;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) m.a[3 * r + c] = n6x6[6 * (3 * br + r) + 3 * bc + c];
    return m;
}

__device__ __forceinline__ M3 identity3() {
    M3 m = zero3();
    m.a[0] = m.a[4] = m.a[8] = 1.0;
    return m;
}

// RPY couplings between segment n and the body taken as a sphere of radius
// R (swimmer.cpp:818-843).
__device__ void sphere_terms(const FilState& S, const FilParams& P, int64_t g, int sw, double R, M3& Mtt, M3& Mtr,
                             M3& Mrr, M3& tx) {
    double rv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) rv[k] = S.x[3 * g + k] - S.bx[3 * sw + k];
    const double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
    double rh[3] = {rv[0] / r, rv[1] / r, rv[2] / r};
    const double a2 = P.a * P.a, R2 = R * R;
    const double ct = 1.0 / (8.0 * kPi * P.mu * r), cr = 1.0 / (8.0 * kPi * P.mu * r * r),
                 crr = 1.0 / (16.0 * kPi * P.mu * r * r * r);
    const double d1 = 1.0 + (R2 + a2) / (3.0 * r * r), d2 = 1.0 - (R2 + a2) / (r * r);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double rr = rh[i] * rh[j], id = i == j ? 1.0 : 0.0;
            Mtt.a[3 * i + j] = (d1 * id + d2 * rr) * ct;
            Mrr.a[3 * i + j] = (3.0 * rr - id) * crr;
        }
    Mtr = scale(cr, rcross(rh));
    double t[3];
    tangent(load_q(S.q + 4 * g), t);
    tx = rcross(t);
}

struct CouplingOut {
    const double* n_rot;   // 36 per body
    double* A;             // per filament 6N x 6, swimmer layout
    double* B;             // per filament 6 x 6N
    double* part;          // 9 per filament: its term of omega_lie_deriv (or of S, bfe)
    double sphere;         // radius of the sphere correction (0: none)
    int bdf2;
    int bfe;               // Broyden-for-everything form: B is the force and torque balance's D
};

// swimmer::prepare_jacobian_inv, GMRES branch with a free body
// (swimmer.cpp:688-880): the coupling blocks of one filament per workgroup;
// thread c owns column block c of B_i (3 columns) and row block c of A_i.
__global__ __launch_bounds__(64) void coupling_kernel(FilState S, FilParams P, CouplingOut O) {
    const int f = blockIdx.x;
    const int N = P.nseg, n6 = 6 * N;
    const int sw = f / P.fil_per_swim;
    const int64_t s0 = (int64_t)f * N;
    const double dt_fac = O.bdf2 ? -2.0 * P.dt / 3.0 : -P.dt;
    M3 Qi;
#pragma unroll
    for (int e = 0; e < 9; ++e) Qi.a[e] = S.bQ[9 * sw + e];
    double d[3];
    mv(Qi, S.brefs + 3 * f, d);
    const double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double dd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dd[k] = (0.5 * P.dl + nd) * d[k] / nd;
    const M3 Dm = rcross(dd);
    const M3 I3 = identity3();
    const M3 L = scale(-1.0, rcross(S.bu + 3 * sw));
    const M3 LL = mm(L, L);
    const M3 Dinv = axpy(-0.5, L, axpy(1.0 / 12.0, LL, I3));
    const M3 DinvT = axpy(0.5, L, axpy(1.0 / 12.0, LL, I3));
    double tN[3], efN[3];
    tangent(load_q(S.q + 4 * (s0 + N - 1)), tN);
#pragma unroll
    for (int k = 0; k < 3; ++k) efN[k] = P.end_force * tN[k];
    const M3 T = rcross(efN);
    const double* nr = O.bfe ? nullptr : O.n_rot + 36 * sw;
    const M3 N00 = nr ? block_of(nr, 0, 0) : zero3(), N01 = nr ? block_of(nr, 0, 1) : zero3(),
             N10 = nr ? block_of(nr, 1, 0) : zero3(), N11 = nr ? block_of(nr, 1, 1) : zero3();
    const int swim_f = f % P.fil_per_swim;
    double* Bf = O.B + (int64_t)sw * P.fil_per_swim * 6 * n6 + (int64_t)swim_f * 6 * n6;
    double* Af = O.A + (int64_t)sw * P.fil_per_swim * 6 * n6 + (int64_t)swim_f * 6 * n6;
    const double* mult = S.mult + (int64_t)f * 3 * (N + 1);
    for (int c = threadIdx.x; c < 2 * N; c += blockDim.x) {
        M3 top = zero3(), bot = zero3();
        if (c == 0) {
            top = scale(-1.0, I3);
            bot = rcross(d);
        } else if (c == N) {
            bot = scale(-1.0, DinvT);
        } else if (c == 2 * N - 1) {
            top = T;
            double v[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = 0.5 * P.dl * tN[k] - (S.x[3 * (s0 + N - 1) + k] - S.bx[3 * sw + k]);
            bot = mm(rcross(v), T);
        } else if (c > N) {
            double tm[3];
            tangent(load_q(S.q + 4 * (s0 + c - N)), tm);
#pragma unroll
            for (int k = 0; k < 3; ++k) tm[k] *= P.dl;
            bot = mm(T, rcross(tm));
        }
        if (O.bfe) {
            // the force and torque balance's dependence (swimmer.cpp:605-640):
            // D = -[top; bot]
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    Bf[(int64_t)r * n6 + 3 * c + j] = -top.a[3 * r + j];
                    Bf[(int64_t)(3 + r) * n6 + 3 * c + j] = -bot.a[3 * r + j];
                }
        } else {
        // body_mobility * B
        M3 nt_ = axpy(1.0, mm(N01, bot), mm(N00, top));
        M3 nb_ = axpy(1.0, mm(N11, bot), mm(N10, top));
        if (O.sphere > 0.0) {
            M3 Mtt, Mtr, Mrr, tx;
            if (c == 0) {
                sphere_terms(S, P, s0, sw, O.sphere, Mtt, Mtr, Mrr, tx);
                nt_ = axpy(1.0, Mtt, nt_);
                nb_ = axpy(1.0, Mtr, nb_);
            } else if (c == N) {
                sphere_terms(S, P, s0, sw, O.sphere, Mtt, Mtr, Mrr, tx);
                nt_ = axpy(1.0, mm(Mtr, DinvT), nt_);
                nb_ = axpy(1.0, mm(Mrr, DinvT), nb_);
            } else if (c < N) {
                // lambda_{c-1}: segment c-1 from the left, segment c from the right
                sphere_terms(S, P, s0 + c - 1, sw, O.sphere, Mtt, Mtr, Mrr, tx);
                nt_ = axpy(-1.0, Mtt, axpy(0.5 * P.dl, mm(Mtr, tx), nt_));
                nb_ = axpy(-1.0, Mtr, axpy(0.5 * P.dl, mm(Mrr, tx), nb_));
                sphere_terms(S, P, s0 + c, sw, O.sphere, Mtt, Mtr, Mrr, tx);
                nt_ = axpy(1.0, Mtt, axpy(0.5 * P.dl, mm(Mtr, tx), nt_));
                nb_ = axpy(1.0, Mtr, axpy(0.5 * P.dl, mm(Mrr, tx), nb_));
            } else if (c == 2 * N - 1) {
                sphere_terms(S, P, s0 + N - 1, sw, O.sphere, Mtt, Mtr, Mrr, tx);
                nt_ = axpy(-P.end_force, mm(Mtt, tx), nt_);
                nb_ = axpy(-P.end_force, mm(Mtr, tx), nb_);
            }
        }
        nb_ = mm(Dinv, nb_);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Bf[(int64_t)r * n6 + 3 * c + j] = dt_fac * nt_.a[3 * r + j];
                Bf[(int64_t)(3 + r) * n6 + 3 * c + j] = dt_fac * nb_.a[3 * r + j];
            }
        }
        // A: [I | Dm] on the position rows, the clamping blocks on the Lie
        // rows of segments 0 and 1
        M3 al = zero3(), ar = zero3();
        if (c < N) {
            al = I3;
            ar = Dm;
        } else if (c == N || c == N + 1) {
#pragma unroll
            for (int e = 0; e < 9; ++e) ar.a[e] = S.clampblk[18 * (int64_t)f + 9 * (c - N) + e];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Af[(int64_t)(3 * c + r) * 6 + j] = al.a[3 * r + j];
                Af[(int64_t)(3 * c + r) * 6 + 3 + j] = ar.a[3 * r + j];
            }
    }
    if (threadIdx.x == 0 && O.bfe) {
        // this filament's terms of the Schur matrix (swimmer.cpp:597-617)
        const double* tether = mult;
        const double* clamp = mult + 3 * N;
        const M3 U = rcross(S.bu + 3 * sw);
        const M3 rc = rcross(clamp);
        double uxl[3];
        cross(S.bu + 3 * sw, clamp, uxl);
        M3 part = mm(rcross(tether), rcross(d));
        part = axpy(0.5, rc, part);
        part = axpy(1.0 / 12.0, axpy(-1.0, mm(U, rc), rcross(uxl)), part);
        double tb[3];
        tangent(load_q(S.q + 4 * s0), tb);
#pragma unroll
        for (int k = 0; k < 3; ++k) tb[k] *= 0.5 * P.dl;
        part = axpy(-1.0, mm(T, rcross(tb)), part);   // rcross(-EF t_end) rcross(DL t_base / 2)
#pragma unroll
        for (int e = 0; e < 9; ++e) O.part[9 * (int64_t)f + e] = part.a[e];
    } else if (threadIdx.x == 0) {
        const double* tether = mult;
        const double* clamp = mult + 3 * N;
        M3 part = scale(-1.0, mm(rcross(tether), rcross(d)));
        double uxl[3];
        cross(S.bu + 3 * sw, clamp, uxl);
        const M3 rc = rcross(clamp);
        part = axpy(0.5, rc, part);
        part = axpy(-1.0 / 12.0, axpy(1.0, rcross(uxl), mm(L, rc)), part);
        part = axpy(1.0, mm(T, Dm), part);
#pragma unroll
        for (int e = 0; e < 9; ++e) O.part[9 * (int64_t)f + e] = part.a[e];
    }
}

// The Schur matrix of the Broyden-for-everything form before its filament
// corrections (swimmer.cpp:565-596): the blob terms with the diagonal blob
// mobility and the blob forces' torque derivative, plus the filaments' terms;
// and the blob displacements d_m = Q_init ref_m.  One workgroup per swimmer.
__global__ __launch_bounds__(256) void bfe_body_block_kernel(FilState S, FilParams P, int nblob,
                                                             const double* __restrict__ blob_refs,
                                                             const double* __restrict__ f_blobs,
                                                             const double* __restrict__ part, double bmf,
                                                             double* __restrict__ S0, double* __restrict__ disp) {
    const int sw = blockIdx.x;
    M3 Qi;
#pragma unroll
    for (int e = 0; e < 9; ++e) Qi.a[e] = S.bQ[9 * sw + e];
    double acc[27];
#pragma unroll
    for (int e = 0; e < 27; ++e) acc[e] = 0.0;
    for (int m = threadIdx.x; m < nblob; m += blockDim.x) {
        const int64_t g = (int64_t)sw * nblob + m;
        double d[3];
        mv(Qi, blob_refs + 3 * g, d);
#pragma unroll
        for (int k = 0; k < 3; ++k) disp[3 * g + k] = d[k];
        const M3 rd = rcross(d);
        const M3 fr = mm(rcross(f_blobs + 3 * g), rd);
        const M3 rr = mm(rd, rd);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            acc[e] += fr.a[e];
            acc[9 + e] += rd.a[e];
            acc[18 + e] += rr.a[e];
        }
    }
    __shared__ double sm[27][256];
#pragma unroll
    for (int e = 0; e < 27; ++e) sm[e][threadIdx.x] = acc[e];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int e = 0; e < 27; ++e) sm[e][threadIdx.x] += sm[e][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* s = S0 + 36 * sw;
        for (int e = 0; e < 36; ++e) s[e] = 0.0;
        for (int k = 0; k < 3; ++k) s[7 * k] = -nblob * bmf;
        M3 rot;
        for (int e = 0; e < 9; ++e) rot.a[e] = sm[e][0] + bmf * sm[18 + e][0];
        for (int i = 0; i < P.fil_per_swim; ++i)
            for (int e = 0; e < 9; ++e) rot.a[e] += part[9 * ((int64_t)sw * P.fil_per_swim + i) + e];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                s[6 * r + 3 + c] = -bmf * sm[9 + 3 * r + c][0];
                s[6 * (3 + r) + c] = bmf * sm[9 + 3 * r + c][0];
                s[6 * (3 + r) + 3 + c] = rot.a[3 * r + c];
            }
    }
}

// The body block S0 (swimmer.cpp:692-780), one thread per body.
__global__ void body_block_kernel(FilState S, FilParams P, int nswim, const double* __restrict__ n_rot,
                                  const double* __restrict__ part, int bdf2, double* __restrict__ S0) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nswim) return;
    const double dt_fac = bdf2 ? -2.0 * P.dt / 3.0 : -P.dt;
    const M3 L = scale(-1.0, rcross(S.bu + 3 * b));
    const M3 cbo = scale(-1.0 / dt_fac, L);
    M3 old = zero3();
    for (int i = 0; i < P.fil_per_swim; ++i)
#pragma unroll
        for (int e = 0; e < 9; ++e) old.a[e] += part[9 * ((int64_t)b * P.fil_per_swim + i) + e];
    old = mm(block_of(n_rot + 36 * b, 1, 1), old);
    M3 lie = axpy(-0.5, cbo, scale(1.0 / 12.0, mm(L, cbo)));
    M3 om = axpy(1.0 / 12.0, mm(L, mm(L, old)), axpy(-0.5, mm(L, old), old));
    lie = axpy(1.0, om, lie);
    double* s = S0 + 36 * b;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) s[6 * r + c] = r == c ? 1.0 : 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s[6 * (3 + r) + 3 + c] += dt_fac * lie.a[3 * r + c];
}

}  // namespace

struct tfs_filament {
    tfs_filament_desc desc{};
    FilParams P{};
    FilState S{};
    double* arena = nullptr;
    double* stage = nullptr;  // host-memory staging (forces, blocks, updates)
    size_t stage_len = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // free bodies (tfs_filament_set_body)
    double* body_arena = nullptr;
    double *blob_refs = nullptr, *n_ref = nullptr, *n_rot = nullptr, *part = nullptr;
    double* base_slab = nullptr;  // 3 per filament, in the arena (tfs_filament_set_base)
    int64_t nblob = 0;
    int32_t blobs_per_swim = 0;
    bool jacobian_done = false;
    // while tfs_elastic_step captures its iterations: a device flag that
    // turns the state's update into no-ops once the iterations have stopped
    const int32_t* update_stop = nullptr;
};

namespace {

int check(hipError_t e, const char* where) { return e == hipSuccess ? TFS_OK : hip_fail(e, where); }

int64_t num_segs(const tfs_filament* h) { return (int64_t)h->P.nfil * h->P.nseg; }
int64_t update_len(const tfs_filament* h) {
    return (int64_t)h->desc.num_swim * ((int64_t)h->desc.fil_per_swim * 6 * h->P.nseg + 6);
}
int64_t blocks_len(const tfs_filament* h) {
    const int64_t n6 = 6 * h->P.nseg;
    return (int64_t)h->desc.num_swim * ((int64_t)h->desc.fil_per_swim * n6 * n6 + 36);
}

double* staging(tfs_filament* h, size_t n) {
    if (h->stage_len < n) {
        if (h->stage) (void)hipFree(h->stage);
        h->stage = nullptr;
        h->stage_len = 0;
        if (hipMalloc(&h->stage, n * 8) != hipSuccess) return nullptr;
        h->stage_len = n;
    }
    return h->stage;
}

unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" {

int tfs_filament_create(const tfs_filament_desc* desc, tfs_filament** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_create: null argument");
    *out = nullptr;
    const tfs_filament_desc& D = *desc;
    if (D.num_swim < 1 || D.fil_per_swim < 1 || D.num_seg < 2)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_create: need swimmers, filaments and >= 2 segments");
    if (!(D.dl > 0) || !(D.mu > 0) || !(D.radius > 0) || !(D.dt > 0) || !std::isfinite(D.kb) ||
        !std::isfinite(D.kt) || !std::isfinite(D.end_force))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_create: dl, mu, radius, dt must be positive");
    if (D.memory != TFS_MEMORY_HOST && D.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_create: unknown memory space");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(TFS_ERR_NO_DEVICE, "tfs_filament_create: no GPU");
    tfs_filament* h = new tfs_filament();
    h->desc = D;
    FilParams& P = h->P;
    P.nfil = D.num_swim * D.fil_per_swim;
    P.nseg = D.num_seg;
    P.fil_per_swim = D.fil_per_swim;
    P.dl = D.dl;
    P.kb = D.kb;
    P.kt = D.kt;
    P.mu = D.mu;
    P.a = D.radius;
    P.dt = D.dt;
    P.end_force = D.end_force;
    for (int k = 0; k < 3; ++k) P.st[k] = D.strain_twist[k];
    P.pert_a = P.pert_b = -1;
    P.base_mode = TFS_BASE_CLAMPED;
    const int64_t ns = num_segs(h), nf = P.nfil, nm = 3LL * (P.nseg + 1) * nf;
    std::vector<std::pair<double**, int64_t>> parts = {
        {&h->S.q, 4 * ns},   {&h->S.qm1, 4 * ns},  {&h->S.u, 3 * ns},      {&h->S.um1, 3 * ns},
        {&h->S.x, 3 * ns},   {&h->S.xm1, 3 * ns},  {&h->S.xm2, 3 * ns},    {&h->S.mult, nm},
        {&h->S.mult_m1, nm}, {&h->S.mult_m2, nm},  {&h->S.bases, 3 * nf},  {&h->S.pert, 6 * nf},
        {&h->S.clampblk, 18 * nf}, {&h->base_slab, 3 * nf}};
    int64_t total = 0;
    for (auto& p : parts) total += (p.second + 31) / 32 * 32;
    if (hipMalloc(&h->arena, total * 8) != hipSuccess) {
        delete h;
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_create: device memory");
    }
    double* p = h->arena;
    for (auto& q : parts) {
        *q.first = p;
        p += (q.second + 31) / 32 * 32;
    }
    h->S.pert = nullptr;  // set by tfs_filament_set_perturbation
    if (D.stream) {
        h->stream = (hipStream_t)D.stream;
    } else {
        if (hipStreamCreate(&h->stream) != hipSuccess) {
            (void)hipFree(h->arena);
            delete h;
            return fail(TFS_ERR_HIP, "tfs_filament_create: stream");
        }
        h->own_stream = true;
    }
    if (hipMemsetAsync(h->arena, 0, total * 8, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        tfs_filament_destroy(h);
        return fail(TFS_ERR_HIP, "tfs_filament_create: clear");
    }
    *out = h;
    return TFS_OK;
}

int tfs_filament_destroy(tfs_filament* h) {
    if (!h) return TFS_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->arena) (void)hipFree(h->arena);
    if (h->body_arena) (void)hipFree(h->body_arena);
    if (h->stage) (void)hipFree(h->stage);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return TFS_OK;
}

int tfs_filament_set_state(tfs_filament* h, const double* bases, const double* q) {
    if (!h || !bases || !q) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_state: null argument");
    const hipMemcpyKind k = h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    hipError_t e = hipMemcpyAsync(h->S.bases, bases, 24LL * h->P.nfil, k, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->S.q, q, 32 * num_segs(h), k, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(set_state_kernel, dim3(grid_for(h->P.nfil, 64)), dim3(64), 0, h->stream, h->S, h->P);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_set_state");
}

int tfs_filament_set_perturbation(tfs_filament* h, int32_t seg_a, int32_t seg_b, const double* forces) {
    if (!h || !forces) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_perturbation: null argument");
    if (seg_a < 0 || seg_a >= h->P.nseg || seg_b < 0 || seg_b >= h->P.nseg)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_perturbation: segment index out of range");
    // the perturbation slab follows the bases in the arena
    double* pert = h->S.bases + (3LL * h->P.nfil + 31) / 32 * 32;
    const hipMemcpyKind k = h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    hipError_t e = hipMemcpyAsync(pert, forces, 48LL * h->P.nfil, k, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "tfs_filament_set_perturbation");
    h->S.pert = pert;
    h->P.pert_a = seg_a;
    h->P.pert_b = seg_b;
    return TFS_OK;
}

int tfs_filament_set_base(tfs_filament* h, int32_t mode, const double* values) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_base: null handle");
    if (mode != TFS_BASE_CLAMPED && mode != TFS_BASE_DRIVEN)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_base: unknown base mode");
    if (mode == TFS_BASE_DRIVEN && h->S.bx)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_base: a driven base needs a body held still");
    hipError_t e = hipSuccess;
    if (values) {
        const hipMemcpyKind k = h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        e = hipMemcpyAsync(h->base_slab, values, 24LL * h->P.nfil, k, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(e, "tfs_filament_set_base");
    }
    h->S.base_val = values ? h->base_slab : nullptr;
    h->P.base_mode = mode;
    h->jacobian_done = false;
    return TFS_OK;
}

int tfs_filament_initial_guess(tfs_filament* h, int32_t nt) {
    if (!h || nt < 0) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_initial_guess: null handle or nt < 0");
    h->jacobian_done = false;
    if (h->S.bx) {
        // swimmer::initial_guess: the body, then its filaments from the new pose
        hipLaunchKernelGGL(body_guess_kernel, dim3(grid_for(h->desc.num_swim, 64)), dim3(64), 0, h->stream, h->S,
                           (int)h->desc.num_swim, (int)nt);
        const hipError_t e0 = hipGetLastError();
        if (e0 != hipSuccess) return hip_fail(e0, "tfs_filament_initial_guess: body");
    }
    hipLaunchKernelGGL(initial_guess_kernel, dim3(grid_for(h->P.nfil, 64)), dim3(64), 0, h->stream, h->S, h->P,
                       (int)nt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && h->desc.memory == TFS_MEMORY_HOST) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_initial_guess");
}

int tfs_filament_forces(tfs_filament* h, int32_t nt, double* f_segs) {
    if (!h || !f_segs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_forces: null argument");
    const int64_t ns = num_segs(h);
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    double* out = host ? staging(h, 6 * ns) : f_segs;
    if (!out) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_forces: staging memory");
    hipLaunchKernelGGL(forces_kernel, dim3(grid_for(ns, 256)), dim3(256), 0, h->stream, h->S, h->P, (int)nt, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && host) e = hipMemcpyAsync(f_segs, out, 48 * ns, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_forces");
}

int tfs_filament_jacobian(tfs_filament* h, int32_t nt, const double* f_segs, double* blocks) {
    if (!h || !f_segs || !blocks) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_jacobian: null argument");
    const int64_t ns = num_segs(h), nb = blocks_len(h);
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    const double* fd = f_segs;
    double* out = blocks;
    hipError_t e = hipSuccess;
    if (host) {
        double* st = staging(h, 6 * ns + nb);
        if (!st) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_jacobian: staging memory");
        e = hipMemcpyAsync(st, f_segs, 48 * ns, hipMemcpyHostToDevice, h->stream);
        // body blocks are the caller's: start from the caller's array
        if (e == hipSuccess) e = hipMemcpyAsync(st + 6 * ns, blocks, 8 * nb, hipMemcpyHostToDevice, h->stream);
        fd = st;
        out = st + 6 * ns;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(jacobian_kernel, dim3((unsigned)h->P.nfil), dim3(64), 0, h->stream, h->S, h->P,
                           nt >= 1 ? 1 : 0, fd, out);
        e = hipGetLastError();
    }
    h->jacobian_done = e == hipSuccess;
    if (e == hipSuccess && host) e = hipMemcpyAsync(blocks, out, 8 * nb, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_jacobian");
}

int tfs_filament_update(tfs_filament* h, const double* update) {
    if (!h || !update) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_update: null argument");
    const int64_t n = update_len(h);
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    const double* du = update;
    hipError_t e = hipSuccess;
    if (host) {
        double* st = staging(h, n);
        if (!st) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_update: staging memory");
        e = hipMemcpyAsync(st, update, 8 * n, hipMemcpyHostToDevice, h->stream);
        du = st;
    }
    if (e == hipSuccess && h->S.bx) {
        hipLaunchKernelGGL(body_update_kernel, dim3(grid_for(h->desc.num_swim, 64)), dim3(64), 0, h->stream, h->S,
                           h->P, (int)h->desc.num_swim, du, h->update_stop);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(update_kernel, dim3(grid_for(num_segs(h), 256)), dim3(256), 0, h->stream, h->S, h->P, du,
                           h->update_stop);
        if (h->P.nseg <= 64)
            hipLaunchKernelGGL(robot_arm_wave_kernel, dim3((unsigned)h->P.nfil), dim3(64), 0, h->stream, h->S, h->P);
        else
            hipLaunchKernelGGL(robot_arm_kernel, dim3(grid_for(h->P.nfil, 64)), dim3(64), 0, h->stream, h->S, h->P);
        e = hipGetLastError();
    }
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_update");
}

int tfs_filament_sync(tfs_filament* h) {
    if (!h) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_sync: null handle");
    return check(hipStreamSynchronize(h->stream), "tfs_filament_sync");
}

static int array_of(tfs_filament* h, int32_t which, double** p, int64_t* len) {
    const int64_t ns = num_segs(h);
    switch (which) {
        case TFS_FIL_X: *p = h->S.x; *len = 3 * ns; return TFS_OK;
        case TFS_FIL_XM1: *p = h->S.xm1; *len = 3 * ns; return TFS_OK;
        case TFS_FIL_XM2: *p = h->S.xm2; *len = 3 * ns; return TFS_OK;
        case TFS_FIL_U: *p = h->S.u; *len = 3 * ns; return TFS_OK;
        case TFS_FIL_UM1: *p = h->S.um1; *len = 3 * ns; return TFS_OK;
        case TFS_FIL_Q: *p = h->S.q; *len = 4 * ns; return TFS_OK;
        case TFS_FIL_MULT: *p = h->S.mult; *len = 3LL * (h->P.nseg + 1) * h->P.nfil; return TFS_OK;
        case TFS_FIL_BODY_X: case TFS_FIL_BODY_XM1: case TFS_FIL_BODY_XM2: case TFS_FIL_BODY_U:
        case TFS_FIL_BODY_UM1: case TFS_FIL_BODY_Q: {
            if (!h->S.bx) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament: no free body (tfs_filament_set_body)");
            double* const arr[] = {h->S.bx, h->S.bxm1, h->S.bxm2, h->S.bu, h->S.bum1, h->S.bq};
            *p = arr[which - TFS_FIL_BODY_X];
            *len = (which == TFS_FIL_BODY_Q ? 4LL : 3LL) * h->desc.num_swim;
            return TFS_OK;
        }
        default: return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament: unknown array");
    }
}

int tfs_filament_device_array(tfs_filament* h, int32_t which, double** ptr) {
    if (!h || !ptr) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_device_array: null argument");
    int64_t len;
    return array_of(h, which, ptr, &len);
}

int tfs_filament_read(tfs_filament* h, int32_t which, double* out) {
    if (!h || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_read: null argument");
    double* p;
    int64_t len;
    int rc = array_of(h, which, &p, &len);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(out, p, 8 * len, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_read");
}

int tfs_filament_set_body(tfs_filament* h, const double* x_bodies, const double* q_bodies, const double* fil_refs,
                          int32_t blobs_per_swim, const double* blob_refs) {
    if (!h || !x_bodies || !q_bodies || !fil_refs || (blobs_per_swim > 0 && !blob_refs) || blobs_per_swim < 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_body: null argument or negative blob count");
    const int64_t nsw = h->desc.num_swim, nf = h->P.nfil, nb = nsw * blobs_per_swim;
    if (h->P.base_mode == TFS_BASE_DRIVEN)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_body: a driven base needs a body held still");
    if (h->body_arena && nb != h->nblob)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_set_body: the blob count cannot change");
    if (!h->body_arena) {
        const int64_t parts[] = {3 * nsw, 3 * nsw, 3 * nsw, 3 * nsw, 3 * nsw, 4 * nsw, 4 * nsw, 9 * nsw, 3 * nf,
                                 3 * nb,  36 * nsw, 36 * nsw, 9 * nf};
        int64_t total = 0;
        for (int64_t l : parts) total += (l + 31) / 32 * 32;
        if (hipMalloc(&h->body_arena, total * 8) != hipSuccess)
            return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_set_body: device memory");
        double** dst[] = {&h->S.bx, &h->S.bxm1, &h->S.bxm2, &h->S.bu, &h->S.bum1, &h->S.bq, &h->S.bqm1,
                          &h->S.bQ, &h->S.brefs, &h->blob_refs, &h->n_ref, &h->n_rot, &h->part};
        double* p = h->body_arena;
        for (size_t i = 0; i < sizeof(parts) / sizeof(parts[0]); ++i) {
            *dst[i] = p;
            p += (parts[i] + 31) / 32 * 32;
        }
        h->nblob = nb;
        h->blobs_per_swim = blobs_per_swim;
    }
    const hipMemcpyKind k = h->desc.memory == TFS_MEMORY_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    hipError_t e = hipSuccess;
    for (double* d : {h->S.bx, h->S.bxm1, h->S.bxm2})
        if (e == hipSuccess) e = hipMemcpyAsync(d, x_bodies, 24 * nsw, k, h->stream);
    for (double* d : {h->S.bq, h->S.bqm1})
        if (e == hipSuccess) e = hipMemcpyAsync(d, q_bodies, 32 * nsw, k, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->S.bu, 0, 24 * nsw, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->S.bum1, 0, 24 * nsw, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->S.brefs, fil_refs, 24 * nf, k, h->stream);
    if (e == hipSuccess && nb) e = hipMemcpyAsync(h->blob_refs, blob_refs, 24 * nb, k, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(body_rot_kernel, dim3(grid_for(nsw, 64)), dim3(64), 0, h->stream, h->S, (int)nsw);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_set_body");
}

int tfs_filament_blob_positions(tfs_filament* h, double* x_blobs) {
    if (!h || !x_blobs) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_blob_positions: null argument");
    if (!h->S.bx) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_blob_positions: no free body");
    if (!h->nblob) return TFS_OK;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    double* out = host ? staging(h, 3 * h->nblob) : x_blobs;
    if (!out) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_blob_positions: staging memory");
    hipLaunchKernelGGL(blob_positions_kernel, dim3(grid_for(h->nblob, 256)), dim3(256), 0, h->stream, h->S,
                       (const double*)h->blob_refs, h->nblob, (int)h->blobs_per_swim, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && host) e = hipMemcpyAsync(x_blobs, out, 24 * h->nblob, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_blob_positions");
}

int tfs_filament_body_coupling(tfs_filament* h, int32_t nt, const double* body_mobility, double sphere_radius,
                               double* A, double* B, double* S0) {
    if (!h || !body_mobility || !A || !B || !S0 || nt < 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling: null argument or nt < 0");
    if (!h->S.bx) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling: no free body");
    if (!h->jacobian_done)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_filament_body_coupling: call tfs_filament_jacobian at this state first");
    if (!(sphere_radius >= 0.0) || !std::isfinite(sphere_radius))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling: sphere radius must be >= 0");
    const int64_t nsw = h->desc.num_swim, n6 = 6LL * h->P.nseg, lab = (int64_t)h->P.nfil * 6 * n6;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    const hipMemcpyKind k = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    double *dA = A, *dB = B, *dS = S0;
    if (host) {
        double* st = staging(h, 2 * lab + 36 * nsw);
        if (!st) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_body_coupling: staging memory");
        dA = st;
        dB = st + lab;
        dS = st + 2 * lab;
    }
    hipError_t e = hipMemcpyAsync(h->n_ref, body_mobility, 288 * nsw, k, h->stream);
    const int bdf2 = nt >= 1 ? 1 : 0;  // NUM_EULER_STEPS = 1
    if (e == hipSuccess) {
        hipLaunchKernelGGL(rotate_mobility_kernel, dim3(grid_for(nsw, 64)), dim3(64), 0, h->stream, h->S, (int)nsw,
                           (const double*)h->n_ref, h->n_rot);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        CouplingOut o{h->n_rot, dA, dB, h->part, sphere_radius, bdf2};
        hipLaunchKernelGGL(coupling_kernel, dim3((unsigned)h->P.nfil), dim3(64), 0, h->stream, h->S, h->P, o);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(body_block_kernel, dim3(grid_for(nsw, 64)), dim3(64), 0, h->stream, h->S, h->P, (int)nsw,
                           (const double*)h->n_rot, (const double*)h->part, bdf2, dS);
        e = hipGetLastError();
    }
    if (e == hipSuccess && host) e = hipMemcpyAsync(A, dA, 8 * lab, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipMemcpyAsync(B, dB, 8 * lab, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipMemcpyAsync(S0, dS, 288 * nsw, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_body_coupling");
}

int tfs_filament_body_coupling_blobs(tfs_filament* h, int32_t nt, const double* f_blobs, double blob_radius,
                                     double* A, double* B, double* S0, double* blob_disp) {
    if (!h || !f_blobs || !A || !B || !S0 || !blob_disp || nt < 0)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling_blobs: null argument or nt < 0");
    if (!h->S.bx || h->blobs_per_swim < 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling_blobs: no free body with blobs");
    if (!h->jacobian_done)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_filament_body_coupling_blobs: call tfs_filament_jacobian at this state first");
    if (!(blob_radius > 0.0) || !std::isfinite(blob_radius))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_filament_body_coupling_blobs: blob radius must be > 0");
    const int64_t nsw = h->desc.num_swim, n6 = 6LL * h->P.nseg, lab = (int64_t)h->P.nfil * 6 * n6;
    const int64_t nb3 = 3 * nsw * h->blobs_per_swim;
    const bool host = h->desc.memory == TFS_MEMORY_HOST;
    double *dA = A, *dB = B, *dS = S0, *dD = blob_disp;
    const double* dF = f_blobs;
    hipError_t e = hipSuccess;
    if (host) {
        double* st = staging(h, 2 * lab + 36 * nsw + 2 * nb3);
        if (!st) return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_filament_body_coupling_blobs: staging memory");
        dA = st;
        dB = st + lab;
        dS = st + 2 * lab;
        dD = dS + 36 * nsw;
        double* fst = dD + nb3;
        e = hipMemcpyAsync(fst, f_blobs, 8 * nb3, hipMemcpyHostToDevice, h->stream);
        dF = fst;
    }
    // -6 pi mu a / DT (backward Euler) or -9 pi mu a / DT (BDF2), swimmer.cpp:572
    const int bdf2 = nt >= 1 ? 1 : 0;
    const double bmf = (bdf2 ? -9.0 : -6.0) * kPi * h->P.mu * blob_radius / h->P.dt;
    if (e == hipSuccess) {
        CouplingOut o{nullptr, dA, dB, h->part, 0.0, bdf2, 1};
        hipLaunchKernelGGL(coupling_kernel, dim3((unsigned)h->P.nfil), dim3(64), 0, h->stream, h->S, h->P, o);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bfe_body_block_kernel, dim3((unsigned)nsw), dim3(256), 0, h->stream, h->S, h->P,
                           (int)h->blobs_per_swim, (const double*)h->blob_refs, dF, (const double*)h->part, bmf, dS,
                           dD);
        e = hipGetLastError();
    }
    if (e == hipSuccess && host) e = hipMemcpyAsync(A, dA, 8 * lab, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipMemcpyAsync(B, dB, 8 * lab, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipMemcpyAsync(S0, dS, 288 * nsw, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipMemcpyAsync(blob_disp, dD, 8 * nb3, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && host) e = hipStreamSynchronize(h->stream);
    return check(e, "tfs_filament_body_coupling_blobs");
}

// ---- elastic implicit timestep --------------------------------------------

namespace {

__global__ void set_flags_kernel(int32_t* __restrict__ c, int32_t runs, int32_t pending, int32_t stop) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c[0] = runs;
    c[1] = pending;
    c[2] = stop;
}

}  // namespace

struct tfs_elastic {
    tfs_elastic_desc desc{};
    double* arena = nullptr;
    double *f_segs = nullptr, *v_segs = nullptr, *du = nullptr, *blocks = nullptr, *zeros = nullptr;
    int64_t n = 0;
    // free bodies: blob positions, body velocities, the coupling blocks of J0
    bool free_bodies = false;
    double *x_blobs = nullptr, *v_bodies = nullptr, *A = nullptr, *B = nullptr, *S0 = nullptr;
    // Broyden for everything: the plain matvec, blob velocities and force
    // history, J0 and the update in the Broyden layout (blobs after each body)
    int64_t nb_swim = 0, per = 0;
    tfs_rpy* rpy = nullptr;
    double *v_blobs = nullptr, *fbm1 = nullptr, *fbm2 = nullptr, *bfe_blocks = nullptr, *bfe_du = nullptr;
    double* bdisp = nullptr;  // free bodies in that form: the blobs' displacements Q_init ref
    bool graphs = true;       // Broyden iterations as captured hipGraphs where the handles allow
    int64_t graph_iters = 0, eager_iters = 0;
    // captured iterations' bookkeeping: [0] iterations run, [1] one awaits its
    // end of iteration, [2] stopped (2: residual not finite); device and a
    // pinned host image
    int32_t* d_flags = nullptr;
    int32_t* h_flags = nullptr;
};

int tfs_elastic_create(const tfs_elastic_desc* desc, tfs_elastic** out) {
    g_last_error.clear();
    if (!desc || !out || !desc->filaments || !desc->broyden || !desc->f_blobs ||
        (!desc->velocities && desc->broyden_blobs <= 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: null handle or array");
    *out = nullptr;
    const tfs_filament* fh = desc->filaments;
    const bool free_bodies = fh->S.bx != nullptr;
    const bool bfe = desc->broyden_blobs > 0;
    if (bfe && !(desc->blob_radius > 0))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: the Broyden-for-everything form needs blob_radius > 0");
    if (bfe && free_bodies && fh->blobs_per_swim != desc->broyden_blobs)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_elastic_create: broyden_blobs must equal the blobs given to tfs_filament_set_body");
    if (!free_bodies && (!desc->x_bodies || !desc->v_bodies || !desc->x_blobs))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: null body or blob array");
    if (free_bodies && !bfe && (!desc->body_mobility || !(desc->sphere_radius >= 0.0)))
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: free bodies need the body mobility (and radius >= 0)");
    if (fh->desc.memory != TFS_MEMORY_DEVICE)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: the filament handle must use device memory");
    if (!(desc->tol > 0) || desc->max_iter < 1)
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_create: tol > 0 and max_iter >= 1");
    tfs_elastic* e = new tfs_elastic();
    e->desc = *desc;
    const int64_t ns = num_segs(fh), nswim = fh->desc.num_swim;
    e->n = update_len(fh);
    const int64_t nb = blocks_len(fh);
    e->free_bodies = free_bodies;
    const int64_t lab = free_bodies ? (int64_t)fh->P.nfil * 36 * fh->P.nseg : 0;
    const int64_t n6_ = 6LL * fh->P.nseg, fil_ = fh->desc.fil_per_swim;
    e->nb_swim = bfe ? desc->broyden_blobs : 0;
    e->per = fil_ * n6_ + 6 + 3 * e->nb_swim;
    const int64_t nbt = nswim * e->nb_swim;
    const int64_t bfe_len = bfe ? nswim * (fil_ * n6_ * n6_ + 36 + 9 * e->nb_swim) : 0;
    std::vector<std::pair<double**, int64_t>> parts = {
        {&e->f_segs, 6 * ns}, {&e->v_segs, 6 * ns},        {&e->du, e->n},  {&e->blocks, nb},
        {&e->zeros, 6 * nswim}, {&e->x_blobs, 3 * fh->nblob}, {&e->v_bodies, 6 * nswim}, {&e->A, lab},
        {&e->B, lab},          {&e->S0, 36 * nswim},        {&e->v_blobs, 3 * nbt},   {&e->fbm1, 3 * nbt},
        {&e->fbm2, 3 * nbt},   {&e->bfe_blocks, bfe_len},   {&e->bfe_du, bfe ? nswim * e->per : 0},
        {&e->bdisp, free_bodies && bfe ? 3 * nbt : 0}};
    int64_t total = 0;
    for (auto& p : parts) total += (p.second + 31) / 32 * 32;
    if (hipMalloc(&e->arena, total * 8) != hipSuccess || hipMalloc(&e->d_flags, 4 * sizeof(int32_t)) != hipSuccess ||
        hipHostMalloc((void**)&e->h_flags, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        if (e->arena) (void)hipFree(e->arena);
        if (e->d_flags) (void)hipFree(e->d_flags);
        delete e;
        return fail(TFS_ERR_OUT_OF_MEMORY, "tfs_elastic_create: device memory");
    }
    double* p = e->arena;
    for (auto& q : parts) {
        *q.first = p;
        p += (q.second + 31) / 32 * 32;
    }
    // J0's body blocks: the identity (the body rows' residuals vanish for a
    // body held still); the filament blocks are rewritten every step.
    std::vector<double> host(nb, 0.0);
    const int64_t n6 = 6 * fh->P.nseg, per = (int64_t)fh->desc.fil_per_swim * n6 * n6 + 36;
    for (int64_t s = 0; s < nswim; ++s)
        for (int k = 0; k < 6; ++k) host[s * per + fh->desc.fil_per_swim * n6 * n6 + 7 * k] = 1.0;
    hipError_t err = hipMemcpy(e->blocks, host.data(), 8 * nb, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(e->zeros, 0, 48 * nswim);
    if (err == hipSuccess) err = hipMemset(e->v_bodies, 0, 48 * nswim);
    if (err == hipSuccess && free_bodies && fh->nblob && desc->velocities &&
        tfs_velocities_set_body_frame(desc->velocities, fh->blob_refs) != TFS_OK) {
        const std::string msg = g_last_error;
        tfs_elastic_destroy(e);
        g_last_error = msg;
        return TFS_ERR_INVALID_ARGUMENT;
    }
    if (err == hipSuccess && bfe) {
        // J0 in the Broyden layout: filament blocks copied in every step; the
        // body's identity and the blobs' 3 x 3 blocks (1 / (-6 pi mu a)) I fixed
        std::vector<double> hb(bfe_len, 0.0);
        const double inv_fac = -1.0 / (6.0 * kPi * fh->P.mu * desc->blob_radius);
        const int64_t sblk = fil_ * n6_ * n6_ + 36 + 9 * e->nb_swim;
        for (int64_t s = 0; s < nswim; ++s) {
            double* b0 = hb.data() + s * sblk + fil_ * n6_ * n6_;
            for (int k = 0; k < 6; ++k) b0[7 * k] = 1.0;
            for (int64_t m = 0; m < e->nb_swim; ++m)
                for (int k = 0; k < 3; ++k) b0[36 + 9 * m + 4 * k] = inv_fac;
        }
        err = hipMemcpy(e->bfe_blocks, hb.data(), 8 * bfe_len, hipMemcpyHostToDevice);
        if (err == hipSuccess) err = hipMemset(e->fbm1, 0, 24 * nbt);
        if (err == hipSuccess) err = hipMemset(e->fbm2, 0, 24 * nbt);
        if (err == hipSuccess) {
            tfs_rpy_desc rd{};
            rd.num_segs = (int32_t)ns;
            rd.num_blobs = (int32_t)nbt;
            rd.seg_dof = 6;
            rd.precision = TFS_PRECISION_FP64;
            rd.bs_mode = TFS_BS_ACCUMULATE;
            rd.memory = TFS_MEMORY_DEVICE;
            rd.seg_radius = fh->P.a;
            rd.blob_radius = desc->blob_radius;
            rd.mu = fh->P.mu;
            rd.stream = fh->stream;
            if (tfs_rpy_create(&rd, &e->rpy) != TFS_OK) {
                const std::string msg = g_last_error;
                tfs_elastic_destroy(e);
                g_last_error = msg;
                return TFS_ERR_HIP;
            }
        }
    }
    if (err != hipSuccess) {
        tfs_elastic_destroy(e);
        return hip_fail(err, "tfs_elastic_create");
    }
    *out = e;
    return TFS_OK;
}

int tfs_elastic_destroy(tfs_elastic* e) {
    if (!e) return TFS_OK;
    if (e->rpy) tfs_rpy_destroy(e->rpy);
    if (e->arena) (void)hipFree(e->arena);
    if (e->d_flags) (void)hipFree(e->d_flags);
    if (e->h_flags) (void)hipHostFree(e->h_flags);
    delete e;
    return TFS_OK;
}

int tfs_elastic_step(tfs_elastic* e, int32_t nt, tfs_elastic_stats* stats) {
    if (!e || nt < 0) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_step: null handle or nt < 0");
    const tfs_elastic_desc& D = e->desc;
    tfs_filament* fh = D.filaments;
    if ((fh->S.bx != nullptr) != e->free_bodies)
        return fail(TFS_ERR_INVALID_ARGUMENT,
                    "tfs_elastic_step: the filament handle's bodies were set after tfs_elastic_create");
    tfs_elastic_stats local{};
    tfs_elastic_stats& S = stats ? *stats : local;
    S = tfs_elastic_stats{};
    tfs_broyden_state st{};
    st.num_swim = fh->desc.num_swim;
    st.num_fil = fh->desc.fil_per_swim;
    st.num_seg = fh->P.nseg;
    st.bdf2 = nt >= 1 ? 1 : 0;  // NUM_EULER_STEPS = 1
    st.dt = fh->P.dt;
    st.dl = fh->P.dl;
    st.tol = D.tol;
    st.seg_x = fh->S.x;
    st.seg_xm1 = fh->S.xm1;
    st.seg_xm2 = fh->S.xm2;
    st.seg_u = fh->S.u;
    st.seg_um1 = fh->S.um1;
    st.v_segs = e->v_segs;
    st.body_x = st.body_xm1 = st.body_xm2 = st.body_u = st.body_um1 = e->zeros;
    st.v_bodies = e->zeros;
    if (e->free_bodies) {
        st.body_x = fh->S.bx;
        st.body_xm1 = fh->S.bxm1;
        st.body_xm2 = fh->S.bxm2;
        st.body_u = fh->S.bu;
        st.body_um1 = fh->S.bum1;
        st.v_bodies = e->v_bodies;
    }
    if (e->rpy) {
        st.blobs_per_swim = (int32_t)e->nb_swim;
        st.x_blobs = e->free_bodies ? e->x_blobs : D.x_blobs;
        st.v_blobs = e->v_blobs;
        if (e->free_bodies) {
            st.free_bodies = 1;
            st.f_segs = e->f_segs;
            st.f_blobs = D.f_blobs;
        }
    }
    const int64_t nswim = fh->desc.num_swim, fil = fh->desc.fil_per_swim, n6 = 6LL * fh->P.nseg;
    const int64_t nb3 = 3 * e->nb_swim;
    // How the velocities handle learns whether the blobs moved, for this step
    // only: a free body's blobs move at every evaluation; a body held still
    // has its blobs compared once, at the step's first evaluation (the caller
    // may have moved them since the last step), and taken as fixed after
    // that.  The handle's own mode is restored on return.
    struct MotionScope {
        tfs_velocities* v = nullptr;
        int32_t saved = TFS_BLOBS_CHECK;
        ~MotionScope() {
            if (v) (void)tfs_velocities_set_blob_motion(v, saved);
        }
    } motion;
    if (D.velocities) {
        motion.saved = tfs::velocities_blob_motion(D.velocities);
        const int rc0 = tfs_velocities_set_blob_motion(D.velocities, e->free_bodies ? TFS_BLOBS_MOVING : TFS_BLOBS_CHECK);
        if (rc0) return rc0;
        motion.v = D.velocities;
    }
    // Work queued on stream `from` is ordered before later work on `to`: the
    // handles may run on streams of their own (then the host waits), and one
    // shared stream needs no waits.
    const hipStream_t s_fil = fh->stream;
    const hipStream_t s_vel = e->rpy ? e->rpy->stream : tfs::velocities_stream(D.velocities);
    const hipStream_t s_bro = tfs::broyden_stream(D.broyden);
    auto order = [](hipStream_t from, hipStream_t to) -> int {
        return from == to ? TFS_OK : check(hipStreamSynchronize(from), "tfs_elastic_step: stream order");
    };
    // forces -> velocities -> residual in `slot` (enqueue_only: the residual's
    // flags are left for tfs::broyden_errors_result, nothing waits)
    auto evaluate = [&](int slot, int32_t* too_large, bool enqueue_only = false, bool copy_flags = true) -> int {
        int rc = e->free_bodies ? tfs_filament_blob_positions(fh, e->x_blobs) : TFS_OK;
        if (!rc) rc = tfs_filament_forces(fh, nt, e->f_segs);
        if (!rc) rc = order(s_fil, s_vel);
        tfs_solve_stats vs{};
        if (!rc && e->rpy) {
            // USE_BROYDEN_FOR_EVERYTHING: one plain matvec with the blob forces given
            rc = tfs_rpy_upload(e->rpy, TFS_SEG_POSITIONS, fh->S.x);
            if (!rc) rc = tfs_rpy_upload(e->rpy, TFS_BLOB_POSITIONS, e->free_bodies ? e->x_blobs : D.x_blobs);
            if (!rc) rc = tfs_rpy_apply(e->rpy, TFS_BLOCK_ALL, e->f_segs, D.f_blobs, e->v_segs, e->v_blobs);
        } else if (!rc && e->free_bodies)
            rc = tfs::velocities_compute_ordered(D.velocities, fh->S.x, e->f_segs, e->x_blobs, fh->S.bx, nullptr,
                                                 fh->S.bq, e->v_segs, D.f_blobs, e->v_bodies, &vs, true);
        else if (!rc)
            rc = tfs::velocities_compute_ordered(D.velocities, fh->S.x, e->f_segs, D.x_blobs, D.x_bodies,
                                                 D.v_bodies, nullptr, e->v_segs, D.f_blobs, nullptr, &vs, false);
        if (!enqueue_only) {
            S.gmres_iterations += vs.iterations;
            S.velocity_solves++;
        }
        if (!rc) rc = order(s_vel, s_bro);
        if (!rc)
            rc = enqueue_only ? tfs::broyden_errors_enqueue(D.broyden, &st, slot, copy_flags)
                              : tfs_broyden_compute_errors(D.broyden, &st, slot, too_large);
        return rc;
    };
    // The near-pair lists rebuilt in an evaluation have their capacity checked
    // at the residual's own wait (one host round trip per Broyden iteration);
    // an overflow grows the lists and repeats the evaluation, a pure function
    // of the state.
    tfs_rpy* rpyh = e->rpy ? e->rpy : tfs::velocities_rpy(D.velocities);
    struct DeferScope {
        tfs_rpy* h = nullptr;
        bool saved = false;
        ~DeferScope() {
            if (!h) return;
            bool overflow = false;
            (void)hipStreamSynchronize(h->stream);
            (void)tfs::rpy_near_check(h, h->stream, &overflow);
            h->defer_near_check = saved;
        }
    } defer;
    defer.h = rpyh;
    defer.saved = rpyh->defer_near_check;
    rpyh->defer_near_check = true;
    // captured iterations are replayed several per wait when no near list can
    // overflow: a capacity of every particle (all-pairs lists, small bodies)
    if (e->graphs && s_fil == s_vel && s_vel == s_bro && rpyh->desc.precision == TFS_PRECISION_FP64 &&
        rpyh->desc.num_segs + rpyh->desc.num_blobs <= kNearPairsMax) {
        const int rc0 = tfs::rpy_near_reserve(rpyh, 0, rpyh->desc.num_segs + rpyh->desc.num_blobs);
        if (rc0) return rc0;
    }
    auto evaluate_checked = [&](int slot, int32_t* too_large) -> int {
        for (int attempt = 0; attempt < 8; ++attempt) {
            int rc = evaluate(slot, too_large);
            bool overflow = false;
            if (!rc) rc = tfs::rpy_near_check(rpyh, s_bro, &overflow);
            if (rc || !overflow) return rc;
        }
        return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_step: too many overlapping neighbours");
    };
    int rc = tfs_filament_initial_guess(fh, nt);
    if (!rc && e->rpy && nb3) {
        hipLaunchKernelGGL(blob_force_guess_kernel, dim3(grid_for(nswim * nb3, 256)), dim3(256), 0, fh->stream,
                           D.f_blobs, e->fbm1, e->fbm2, nswim * nb3, (int)nt);
        rc = check(hipGetLastError(), "blob force guess");
    }
    int32_t big = 0;
    if (!rc) rc = evaluate_checked(0, &big);
    if (!rc && D.velocities && !e->free_bodies) rc = tfs_velocities_set_blob_motion(D.velocities, TFS_BLOBS_FIXED);
    if (!rc && big) {
        rc = tfs_filament_jacobian(fh, nt, e->f_segs, e->blocks);
        if (!rc) rc = tfs_filament_sync(fh);
        if (!rc && e->rpy) {
            // the filament blocks into the Broyden layout (blob blocks fixed)
            const int64_t src = fil * n6 * n6 + 36, dst = src + 3 * nb3;
            for (int64_t s = 0; s < nswim && !rc; ++s)
                rc = check(hipMemcpyAsync(e->bfe_blocks + s * dst, e->blocks + s * src, 8 * fil * n6 * n6,
                                          hipMemcpyDeviceToDevice, fh->stream),
                           "J0 blocks");
            if (!rc) rc = tfs_filament_sync(fh);
        }
        if (!rc) rc = tfs_broyden_set_jacobian(D.broyden, e->rpy ? e->bfe_blocks : e->blocks);
        if (!rc && e->free_bodies && e->rpy) {
            // swimmer::prepare_jacobian_inv, Broyden for everything: the blobs
            // eliminated with their diagonal mobility, the body's Schur matrix
            const double fac = -6.0 * kPi * fh->P.mu * D.blob_radius;
            const double dt_fac = (nt >= 1 ? 1.5 : 1.0) / fh->P.dt;
            rc = tfs_filament_body_coupling_blobs(fh, nt, D.f_blobs, D.blob_radius, e->A, e->B, e->S0, e->bdisp);
            if (!rc) rc = tfs_filament_sync(fh);
            if (!rc)
                rc = tfs_broyden_set_body_coupling_blobs(D.broyden, fh->desc.num_swim, fh->desc.fil_per_swim,
                                                         (int32_t)e->nb_swim, e->A, e->B, e->S0, e->bdisp, fac,
                                                         dt_fac);
        } else if (!rc && e->free_bodies) {
            // swimmer::prepare_jacobian_inv: the body's Schur complement
            rc = tfs_filament_body_coupling(fh, nt, D.body_mobility, D.sphere_radius, e->A, e->B, e->S0);
            if (!rc) rc = tfs_filament_sync(fh);
            if (!rc)
                rc = tfs_broyden_set_body_coupling(D.broyden, fh->desc.num_swim, fh->desc.fil_per_swim, e->A, e->B,
                                                   e->S0);
        }
    }
    if (!rc) rc = tfs_broyden_start(D.broyden, nullptr);
    // find_update -> the state's update (device_count: the engine's iteration
    // count read on the device)
    auto advance = [&](bool device_count) -> int {
        int rc2;
        if (e->rpy) {
            // the filament and body entries to the filament layout, the blob
            // entries onto the blob forces
            rc2 = device_count ? tfs::broyden_dev_find_update(D.broyden, e->bfe_du)
                               : tfs_broyden_find_update(D.broyden, e->bfe_du);
            if (!rc2) rc2 = order(s_bro, s_fil);
            for (int64_t s = 0; s < nswim && !rc2; ++s)
                rc2 = check(hipMemcpyAsync(e->du + s * (fil * n6 + 6), e->bfe_du + s * e->per, 8 * (fil * n6 + 6),
                                           hipMemcpyDeviceToDevice, fh->stream),
                            "update");
            if (!rc2 && nb3) {
                hipLaunchKernelGGL(blob_force_update_kernel, dim3(grid_for(nswim * nb3, 256)), dim3(256), 0,
                                   fh->stream, D.f_blobs, (const double*)e->bfe_du, nb3, (int)nswim, e->per,
                                   fil * n6 + 6, fh->update_stop);
                rc2 = check(hipGetLastError(), "blob force update");
            }
        } else {
            rc2 = device_count ? tfs::broyden_dev_find_update(D.broyden, e->du)
                               : tfs_broyden_find_update(D.broyden, e->du);
        }
        if (!rc2 && !e->rpy) rc2 = order(s_bro, s_fil);
        if (!rc2) rc2 = tfs_filament_update(fh, e->du);
        return rc2;
    };
    int32_t iter = 0;
    // Captured iterations: one stream for every handle, device memory, fixed
    // launches (FP64 pair blocks, the dense direct blob solve or the plain
    // matvec), an engine small enough for its device-counted form.  The
    // captured iteration is the engine's end of the previous iteration (while
    // one is pending), find_update, the state's update (a no-op once stopped),
    // forces, velocities, residual and the bookkeeping that counts the
    // iteration and stops the iterations; every other launch recomputes the
    // same values from an unchanged state, so the iteration is replayed
    // `batch` times per host wait.  A last graph ends the last iteration.
    // With near lists of a capacity no count can exceed, the batch is the
    // number of iterations the residual's observed rate of decrease still
    // needs to reach the tolerance (1 to 8; replays past the stop cost one
    // evaluation each); otherwise 1, with the capacity checked at each wait.
    const bool graph_mode = !rc && big && e->graphs && s_fil == s_vel && s_vel == s_bro &&
                            tfs::broyden_dev_supported(D.broyden) && rpyh->desc.precision == TFS_PRECISION_FP64 &&
                            rpyh->desc.memory == TFS_MEMORY_DEVICE &&
                            (e->rpy || tfs::velocities_graph_ready(D.velocities));
    if (graph_mode) {
        struct Graphs {
            hipGraph_t g[2] = {nullptr, nullptr};
            hipGraphExec_t x[2] = {nullptr, nullptr};
            void clear() {
                for (int i = 0; i < 2; ++i) {
                    if (x[i]) (void)hipGraphExecDestroy(x[i]);
                    if (g[i]) (void)hipGraphDestroy(g[i]);
                    x[i] = nullptr;
                    g[i] = nullptr;
                }
            }
            ~Graphs() { clear(); }
        } G;
        auto capture = [&](int which) -> int {
            hipError_t ce = hipStreamBeginCapture(s_fil, hipStreamCaptureModeThreadLocal);
            if (ce != hipSuccess) return hip_fail(ce, "tfs_elastic_step: begin capture");
            fh->update_stop = e->d_flags + 2;
            tfs::broyden_set_bookkeeping(D.broyden, e->d_flags, D.max_iter);
            int rc2 = tfs::broyden_dev_end_of_iter(D.broyden, e->d_flags + 1);
            if (!rc2 && which == 0) rc2 = advance(true);
            if (!rc2 && which == 0) rc2 = evaluate(1, nullptr, true, false);
            tfs::broyden_set_bookkeeping(D.broyden, nullptr, 0);
            fh->update_stop = nullptr;
            hipError_t ee = hipStreamEndCapture(s_fil, &G.g[which]);
            if (rc2) return rc2;
            if (ee != hipSuccess) return hip_fail(ee, "tfs_elastic_step: end capture");
            ce = hipGraphInstantiate(&G.x[which], G.g[which], nullptr, nullptr, 0);
            return ce == hipSuccess ? TFS_OK : hip_fail(ce, "tfs_elastic_step: instantiate");
        };
        const bool roomy = rpyh->near[0].cap >= rpyh->desc.num_segs + rpyh->desc.num_blobs;
        rc = tfs::broyden_dev_begin(D.broyden);
        if (!rc) {
            hipLaunchKernelGGL(set_flags_kernel, dim3(1), dim3(64), 0, s_fil, e->d_flags, 0, 0, 0);
            rc = check(hipGetLastError(), "tfs_elastic_step: bookkeeping");
        }
        int32_t runs = 0, stop = 0, runs_prev = -1;
        double measure_prev = 0.0;
        while (!rc && !stop) {
            if (!G.x[0]) rc = capture(0);
            const double measure = tfs::broyden_last_measure(D.broyden);
            int batch = 1;
            if (roomy && runs_prev >= 0 && runs > runs_prev && measure > 1.0 && measure < measure_prev) {
                const double per_iter = std::log(measure_prev / measure) / (runs - runs_prev);
                const double needed = std::log(measure) / per_iter;
                batch = (int)std::max(1.0, std::min(8.0, std::floor(needed)));
            }
            runs_prev = runs;
            measure_prev = measure;
            for (int r = 0; r < batch && !rc; ++r)
                rc = check(hipGraphLaunch(G.x[0], s_fil), "tfs_elastic_step: iteration graph");
            if (!rc)
                rc = check(hipMemcpyAsync(e->h_flags, e->d_flags, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s_fil),
                           "tfs_elastic_step: bookkeeping");
            if (!rc) rc = tfs::broyden_copy_error_flags(D.broyden);
            if (!rc) rc = check(hipStreamSynchronize(s_fil), "tfs_elastic_step: iteration graph");
            if (rc) break;
            const int32_t before = runs;
            runs = e->h_flags[0];
            stop = e->h_flags[2];
            S.velocity_solves += runs - before;
            e->graph_iters += runs - before;
            rc = tfs::broyden_errors_result(D.broyden, &big);
            if (rc) break;
            bool overflow = false;
            rpyh->near_pending[0] = !roomy;
            rc = tfs::rpy_near_check(rpyh, s_fil, &overflow);
            if (!rc && overflow) {
                // (one iteration per wait here) grown lists: the evaluation
                // again, eagerly, its bookkeeping redone, and new graphs
                G.clear();
                rc = evaluate_checked(1, &big);
                S.velocity_solves++;
                stop = !big || runs >= D.max_iter ? 1 : 0;
                if (!rc) {
                    hipLaunchKernelGGL(set_flags_kernel, dim3(1), dim3(64), 0, s_fil, e->d_flags, runs, 1, stop);
                    rc = check(hipGetLastError(), "tfs_elastic_step: bookkeeping");
                }
            }
        }
        if (!rc && runs > 0) {
            if (!G.x[1]) rc = capture(1);
            if (!rc) rc = check(hipGraphLaunch(G.x[1], s_fil), "tfs_elastic_step: end-of-iteration graph");
        }
        iter = runs;
        if (!rc) tfs::broyden_dev_finish(D.broyden, iter, 1);
        if (!rc && stop == 2) rc = fail(TFS_ERR_NON_FINITE, "tfs_broyden_compute_errors: residual not finite");
    }
    while (!graph_mode && !rc && big && iter < D.max_iter) {
        rc = advance(false);
        if (!rc) rc = evaluate_checked(1, &big);
        if (!rc) rc = tfs_broyden_end_of_iter(D.broyden, nullptr, big);
        if (!rc) rc = tfs_broyden_status(D.broyden, &iter, nullptr);
        e->eager_iters++;
    }
    double norm = 0.0;
    if (!rc) rc = tfs::broyden_latest_norm(D.broyden, &norm);
    S.broyden_iterations = iter;
    S.error_norm = norm;
    S.converged = (!rc && !big) ? 1 : 0;
    if (rc) return rc;
    return big ? fail(TFS_ERR_NOT_CONVERGED, "tfs_elastic_step: Broyden iterations exhausted") : TFS_OK;
}

int tfs_elastic_set_graphs(tfs_elastic* e, int32_t enable) {
    if (!e) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_set_graphs: null handle");
    e->graphs = enable != 0;
    return TFS_OK;
}

int tfs_elastic_iteration_counts(tfs_elastic* e, int64_t* graph_iterations, int64_t* eager_iterations) {
    if (!e) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_iteration_counts: null handle");
    if (graph_iterations) *graph_iterations = e->graph_iters;
    if (eager_iterations) *eager_iterations = e->eager_iters;
    return TFS_OK;
}

}  // extern "C"

namespace {

// Host images of the device state, packed in the backup order or unpacked from it.
struct Backup {
    std::vector<double> q, x, xm1, xm2, u, um1, m, m1, m2, bx, bxm1, bxm2, bu, bum1, bq, fb, fb1, fb2;
};

int64_t backup_len(const tfs_elastic* e) {
    const tfs_filament* fh = e->desc.filaments;
    const int64_t N = fh->P.nseg, nsw = fh->desc.num_swim, fil = fh->desc.fil_per_swim;
    return nsw * (fil * (28 * N + 9) + 19 + 9 * e->nb_swim);
}

int transfer(tfs_elastic* e, Backup& B, bool to_host) {
    tfs_filament* fh = e->desc.filaments;
    const int64_t ns = num_segs(fh), nf = fh->P.nfil, nm = 3LL * (fh->P.nseg + 1) * nf, nsw = fh->desc.num_swim;
    const int64_t nb3 = 3 * nsw * e->nb_swim;
    struct Item {
        std::vector<double>* v;
        double* d;
        int64_t n;
    };
    std::vector<Item> items = {{&B.q, fh->S.q, 4 * ns},     {&B.x, fh->S.x, 3 * ns},     {&B.xm1, fh->S.xm1, 3 * ns},
                               {&B.xm2, fh->S.xm2, 3 * ns}, {&B.u, fh->S.u, 3 * ns},     {&B.um1, fh->S.um1, 3 * ns},
                               {&B.m, fh->S.mult, nm},      {&B.m1, fh->S.mult_m1, nm},  {&B.m2, fh->S.mult_m2, nm}};
    if (fh->S.bx) {
        items.insert(items.end(), {{&B.bx, fh->S.bx, 3 * nsw},
                                   {&B.bxm1, fh->S.bxm1, 3 * nsw},
                                   {&B.bxm2, fh->S.bxm2, 3 * nsw},
                                   {&B.bu, fh->S.bu, 3 * nsw},
                                   {&B.bum1, fh->S.bum1, 3 * nsw},
                                   {&B.bq, fh->S.bq, 4 * nsw}});
    }
    if (e->rpy && nb3) {
        items.insert(items.end(), {{&B.fb, e->desc.f_blobs, nb3}, {&B.fb1, e->fbm1, nb3}, {&B.fb2, e->fbm2, nb3}});
    }
    hipError_t err = hipStreamSynchronize(fh->stream);
    for (auto& it : items) {
        if (err != hipSuccess) break;
        if (to_host) {
            it.v->resize(it.n);
            err = hipMemcpy(it.v->data(), it.d, 8 * it.n, hipMemcpyDeviceToHost);
        } else {
            err = hipMemcpy(it.d, it.v->data(), 8 * it.n, hipMemcpyHostToDevice);
        }
    }
    return check(err, "tfs_elastic backup transfer");
}

}  // namespace

extern "C" {

int tfs_elastic_backup_size(tfs_elastic* e, int64_t* count) {
    if (!e || !count) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_backup_size: null argument");
    *count = backup_len(e);
    return TFS_OK;
}

int tfs_elastic_backup(tfs_elastic* e, const double* held_bodies, double* out) {
    if (!e || !out) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_backup: null argument");
    tfs_filament* fh = e->desc.filaments;
    Backup B;
    int rc = transfer(e, B, true);
    if (rc) return rc;
    const int64_t N = fh->P.nseg, nsw = fh->desc.num_swim, fil = fh->desc.fil_per_swim, nm = 3 * (N + 1);
    int64_t p = 0;
    for (int64_t s = 0; s < nsw; ++s) {
        for (int64_t i = 0; i < fil; ++i) {
            const int64_t f = s * fil + i, s0 = f * N;
            for (int64_t n = 0; n < N; ++n) {
                const int64_t g = s0 + n;
                for (int k = 0; k < 4; ++k) out[p++] = B.q[4 * g + k];
                for (auto* v : {&B.x, &B.xm1, &B.xm2, &B.u, &B.um1})
                    for (int k = 0; k < 3; ++k) out[p++] = (*v)[3 * g + k];
            }
            // [tether, lambda_0 .. lambda_{N-2}, clamp] in the device layout
            for (int64_t slot : {(int64_t)0, 3 * N})
                for (auto* v : {&B.m, &B.m1, &B.m2})
                    for (int k = 0; k < 3; ++k) out[p++] = (*v)[f * nm + slot + k];
            for (int64_t l = 0; l < N - 1; ++l)
                for (auto* v : {&B.m, &B.m1, &B.m2})
                    for (int k = 0; k < 3; ++k) out[p++] = (*v)[f * nm + 3 * (l + 1) + k];
        }
        if (fh->S.bx) {
            for (int k = 0; k < 4; ++k) out[p++] = B.bq[4 * s + k];
            for (auto* v : {&B.bx, &B.bxm1, &B.bxm2, &B.bu, &B.bum1})
                for (int k = 0; k < 3; ++k) out[p++] = (*v)[3 * s + k];
        } else {
            const double* hb = held_bodies ? held_bodies + 7 * s : nullptr;
            const double q[4] = {hb ? hb[3] : 1.0, hb ? hb[4] : 0.0, hb ? hb[5] : 0.0, hb ? hb[6] : 0.0};
            for (int k = 0; k < 4; ++k) out[p++] = q[k];
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) out[p++] = hb ? hb[k] : 0.0;
            for (int k = 0; k < 6; ++k) out[p++] = 0.0;
        }
        for (int64_t c = 0; c < 3 * e->nb_swim; ++c) {
            const int64_t g = 3 * s * e->nb_swim + c;
            out[p++] = B.fb[g];
            out[p++] = B.fb1[g];
            out[p++] = B.fb2[g];
        }
    }
    return p == backup_len(e) ? TFS_OK : fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_backup: layout mismatch");
}

int tfs_elastic_restore(tfs_elastic* e, const double* in) {
    if (!e || !in) return fail(TFS_ERR_INVALID_ARGUMENT, "tfs_elastic_restore: null argument");
    tfs_filament* fh = e->desc.filaments;
    Backup B;
    int rc = transfer(e, B, true);  // sizes, and the fields a backup does not hold
    if (rc) return rc;
    const int64_t N = fh->P.nseg, nsw = fh->desc.num_swim, fil = fh->desc.fil_per_swim, nm = 3 * (N + 1);
    int64_t p = 0;
    for (int64_t s = 0; s < nsw; ++s) {
        for (int64_t i = 0; i < fil; ++i) {
            const int64_t f = s * fil + i, s0 = f * N;
            for (int64_t n = 0; n < N; ++n) {
                const int64_t g = s0 + n;
                for (int k = 0; k < 4; ++k) B.q[4 * g + k] = in[p++];
                for (auto* v : {&B.x, &B.xm1, &B.xm2, &B.u, &B.um1})
                    for (int k = 0; k < 3; ++k) (*v)[3 * g + k] = in[p++];
            }
            for (int64_t slot : {(int64_t)0, 3 * N})
                for (auto* v : {&B.m, &B.m1, &B.m2})
                    for (int k = 0; k < 3; ++k) (*v)[f * nm + slot + k] = in[p++];
            for (int64_t l = 0; l < N - 1; ++l)
                for (auto* v : {&B.m, &B.m1, &B.m2})
                    for (int k = 0; k < 3; ++k) (*v)[f * nm + 3 * (l + 1) + k] = in[p++];
        }
        if (fh->S.bx) {
            for (int k = 0; k < 4; ++k) B.bq[4 * s + k] = in[p++];
            for (auto* v : {&B.bx, &B.bxm1, &B.bxm2, &B.bu, &B.bum1})
                for (int k = 0; k < 3; ++k) (*v)[3 * s + k] = in[p++];
        } else {
            p += 19;  // a body held still keeps the pose it was given
        }
        for (int64_t c = 0; c < 3 * e->nb_swim; ++c) {
            const int64_t g = 3 * s * e->nb_swim + c;
            B.fb[g] = in[p++];
            B.fb1[g] = in[p++];
            B.fb2[g] = in[p++];
        }
    }
    rc = transfer(e, B, false);
    // previous orientations = current ones (filament.cpp:140-146, rigid_body.cpp:37-44)
    if (!rc) rc = check(hipMemcpy(fh->S.qm1, fh->S.q, 32 * num_segs(fh), hipMemcpyDeviceToDevice), "restore");
    if (!rc && fh->S.bx) {
        rc = check(hipMemcpy(fh->S.bqm1, fh->S.bq, 32 * nsw, hipMemcpyDeviceToDevice), "restore");
        if (!rc) {
            hipLaunchKernelGGL(body_rot_kernel, dim3(grid_for(nsw, 64)), dim3(64), 0, fh->stream, fh->S, (int)nsw);
            rc = check(hipGetLastError(), "restore");
        }
        if (!rc) rc = check(hipStreamSynchronize(fh->stream), "restore");
    }
    return rc;
}

}  // extern "C"
