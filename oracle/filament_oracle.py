"""CPU restatement of the reference's Cosserat filament model (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
filament kernels; the product never imports it.

One filament of NSEG segments, quaternions (w, x, y, z) per segment.  The
base segment is either clamped (Params.base == "clamped") -- the reference's
INSTABILITY_CILIA build (CILIA_TYPE 0, with a follower end force and a
first-step kick; the configuration of the elastic "RPY matvec + GMRES +
Broyden" case) and CONSTANT_BASE_ROTATION (CILIA_TYPE 2, no end force, the
base turned by a fixed Lie-algebra element every step, Filament.u0) -- or free
and driven (Params.base == "driven", GEOMETRIC_CILIA, CILIA_TYPE 1): no clamp,
the base segment's Lie-algebra element is an unknown, and a driving torque
Filament.drive acts on it (filament.cpp:1523-1531).

  quaternion algebra            src/general/quaternion.cpp:80-110 (product),
                                :183-195, 245-270 (conj, sqrt), :285-370
                                (frame, rot_mat), :378-420 (psi, left_mult),
                                :522-570 (lie_exp, midpoint), :610-660 (dexpinv,
                                dexpinv_transpose)
  rcross                        src/general/matrix.cpp:665-679 (rcross(v) x = x cross v)
  body_frame_moment             src/cilia/filament.cpp:1451-1465
  body_frame_moment_lie_derivative  filament.cpp:9-55
  internal_forces_and_torques   filament.cpp:1467-1575
  jacobian_lie_algebra_block    filament.cpp:1577-1759
  invert_approx_jacobian        filament.cpp:1761-1894 (the matrix before inversion)
  update, robot_arm             filament.cpp:1896-1930, 625-647; segment.cpp (update)
  initial_guess                 segment.cpp (initial_guess), filament.cpp:1275-1385

Unknowns of one filament (the Broyden vector's filament block, 6 NSEG):
[tether lambda, lambda_0 .. lambda_{N-2}] (3 NSEG), then [clamp lambda,
u_1 .. u_{N-1}] (3 NSEG); with a driven base u_0 takes the clamp's place
(filament.cpp:1918-1927).

The reference's CONSTANT_BASE_ROTATION branch does not compile as shipped
(clamp_lambda is declared only for INSTABILITY_CILIA, filament.hpp:66-70,
but used for CILIA_TYPE 2 at filament.cpp:606, 1565-1571), and its
accept_state_from_rigid_body would reset the base's Lie-algebra element to
the body's every step (filament.cpp:660-665), which undoes the rotation set
up at filament.cpp:487-492.  This restatement follows the evident intent:
the base segment turns by u0 every step, held by the clamp multiplier.  With
u0 = 0 it is the clamped build without end force.  GEOMETRIC_CILIA's driving
torque factor switches only when find_my_angle() passes CRITICAL_ANGLE, and
find_my_angle() returns 0 (filament.cpp:1433-1447), so the torque stays
base_torque_magnitude_factor * DEFAULT_BASE_TORQUE_MAGNITUDE * R(body_q)
(my_vertical x fast_beating_direction) with its initial factor: the caller
passes that vector as Filament.drive.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PI = np.pi
SMALL_ANGLE = 1e-5


# ---- quaternions ------------------------------------------------------------------
def qmul(p, q):
    p0, p1, p2, p3 = p
    q0, q1, q2, q3 = q
    return np.array([p0 * q0 - (p1 * q1 + p2 * q2 + p3 * q3),
                     p0 * q1 + q0 * p1 + p2 * q3 - p3 * q2,
                     p0 * q2 + q0 * p2 + p3 * q1 - p1 * q3,
                     p0 * q3 + q0 * p3 + p1 * q2 - p2 * q1])


def qconj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def qsqrt(q):
    if q[0] < -0.999999999:
        return np.array([0.0, 0.0, 0.0, 1.0])
    s = np.sqrt(0.5 * (1.0 + q[0]))
    return np.array([s, q[1] / (2 * s), q[2] / (2 * s), q[3] / (2 * s)])


def midpoint(q1, q2):
    """(q2 q1*)^(1/2) q1."""
    return qmul(qsqrt(qmul(q2, qconj(q1))), q1)


def lie_exp(u):
    th = np.sqrt(u @ u)
    if th < 1e-6:
        return np.array([1.0, 0.0, 0.0, 0.0])
    return np.concatenate([[np.cos(0.5 * th)], (np.sin(0.5 * th) / th) * u])


def frame(q):
    """(tangent, normal, binormal): the columns of the rotation matrix."""
    w, x, y, z = q
    t = np.array([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)])
    n = np.array([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)])
    b = np.array([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)])
    return t, n, b


def rot_mat(q):
    t, n, b = frame(q)
    return np.stack([t, n, b], axis=1)


def rcross(v):
    """M with M x = x cross v."""
    return np.array([[0.0, v[2], -v[1]], [-v[2], 0.0, v[0]], [v[1], -v[0], 0.0]])


def psi_mat(q):
    """4 x 3, d/d(delta) of exp(delta) q at delta = 0."""
    s, v = q[0], q[1:]
    P = np.zeros((4, 3))
    P[0] = -v
    P[1:] = rcross(v) + s * np.eye(3)
    return 0.5 * P


def left_mult_mat(q):
    """4 x 4 matrix of p -> q p."""
    s, v = q[0], q[1:]
    L = np.zeros((4, 4))
    L[0, 0] = s
    L[0, 1:] = -v
    L[1:, 0] = v
    L[1:, 1:] = -rcross(v) + s * np.eye(3)
    return L


def _dexpinv_fac(u):
    th2 = u @ u
    th = np.sqrt(th2)
    if th < SMALL_ANGLE:
        return -1.0 / 12.0
    return (0.5 * th * np.cos(0.5 * th) / np.sin(0.5 * th) - 1.0) / th2


def dexpinv(u, v):
    c = np.cross(u, v)
    return v - 0.5 * c - _dexpinv_fac(u) * np.cross(u, c)


def dexpinv_transpose(u, v):
    c = np.cross(u, v)
    return v + 0.5 * c - _dexpinv_fac(u) * np.cross(u, c)


# ---- the model ----------------------------------------------------------------------
@dataclass
class Params:
    nseg: int
    dl: float = 2.2
    kb: float = 1800.0
    kt: float = 1800.0
    mu: float = 1.0
    a: float = 1.0
    dt: float = 1.0
    end_force: float = 0.0               # END_FORCE_MAGNITUDE (follower force on the last segment)
    strain_twist: np.ndarray = field(default_factory=lambda: np.zeros(3))
    base: str = "clamped"                # "clamped" (CILIA_TYPE 0, 2) or "driven" (CILIA_TYPE 1)


@dataclass
class Filament:
    q: np.ndarray            # (N, 4)
    qm1: np.ndarray
    u: np.ndarray            # (N, 3)
    um1: np.ndarray
    x: np.ndarray            # (N, 3)
    xm1: np.ndarray
    xm2: np.ndarray
    lam: np.ndarray          # (N-1, 3)
    lamm1: np.ndarray
    lamm2: np.ndarray
    tether: np.ndarray       # (3,) and history
    tetherm1: np.ndarray
    tetherm2: np.ndarray
    clamp: np.ndarray
    clampm1: np.ndarray
    clampm2: np.ndarray
    u0: np.ndarray = field(default_factory=lambda: np.zeros(3))      # clamped base: its turn per step
    drive: np.ndarray = field(default_factory=lambda: np.zeros(3))   # driven base: torque on segment 0

    @staticmethod
    def at_rest(base, q, nseg, dl):
        q = np.array(q, dtype=float).reshape(nseg, 4)
        z3 = lambda: np.zeros(3)
        f = Filament(q=q.copy(), qm1=q.copy(), u=np.zeros((nseg, 3)), um1=np.zeros((nseg, 3)),
                     x=np.zeros((nseg, 3)), xm1=None, xm2=None, lam=np.zeros((nseg - 1, 3)),
                     lamm1=np.zeros((nseg - 1, 3)), lamm2=np.zeros((nseg - 1, 3)), tether=z3(), tetherm1=z3(),
                     tetherm2=z3(), clamp=z3(), clampm1=z3(), clampm2=z3(), u0=z3(), drive=z3())
        f.x[0] = base
        robot_arm(f, dl)
        f.xm1, f.xm2 = f.x.copy(), f.x.copy()
        return f

    def copy(self):
        return Filament(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})


def robot_arm(f: Filament, dl):
    t1 = frame(f.q[0])[0]
    for i in range(1, len(f.q)):
        t2 = frame(f.q[i])[0]
        f.x[i] = f.x[i - 1] + 0.5 * dl * (t1 + t2)
        t1 = t2


def body_frame_moment(q1, q2, p: Params):
    mid = qconj(midpoint(q1, q2))
    m = 2.0 * qmul(mid, (q2 - q1) / p.dl)
    return np.array([p.kt, p.kb, p.kb]) * (m[1:] - p.strain_twist)


def body_frame_moment_lie_derivative(q1, q2, upper, p: Params):
    A = psi_mat(q2) / p.dl if upper else psi_mat(q1) / -p.dl
    mid = qconj(midpoint(q1, q2))
    dM = left_mult_mat(mid) @ A
    dM += left_mult_mat((q2 - q1) / p.dl) @ (A * (0.5 * p.dl if upper else -0.5 * p.dl))
    out = 2.0 * dM[1:, :]
    return out * np.array([p.kt, p.kb, p.kb])[:, None]


def internal_forces(f: Filament, p: Params, perturbation=None):
    """(N, 6) forces and torques; `perturbation` maps a segment index to an
    extra force (the reference's first-step kick, filament.cpp:1533-1552)."""
    N = p.nseg
    out = np.zeros((N, 6))
    for i in range(N - 1):
        lam = f.lam[i]
        out[i, :3] -= lam
        out[i + 1, :3] += lam
        t = frame(f.q[i])[0]
        out[i, 3:] -= 0.5 * p.dl * np.cross(t, lam)
        t = frame(f.q[i + 1])[0]
        out[i + 1, 3:] -= 0.5 * p.dl * np.cross(t, lam)
        M = rot_mat(midpoint(f.q[i], f.q[i + 1])) @ body_frame_moment(f.q[i], f.q[i + 1], p)
        out[i, 3:] += M
        out[i + 1, 3:] -= M
    out[0, :3] += f.tether
    t = frame(f.q[N - 1])[0]
    out[N - 1, :3] -= p.end_force * t
    if perturbation:
        for k, v in perturbation.items():
            out[k, :3] += v
    if p.base == "driven":
        out[0, 3:] += f.drive                                      # filament.cpp:1523-1531
    else:
        out[0, 3:] += dexpinv_transpose(f.u[0], f.clamp)          # :1555-1571
    return out


def lie_algebra_block(f: Filament, forces, bdf2, p: Params):
    N = p.nseg
    E = np.zeros((3 * N, 3 * N))
    tsf = 2.0 * p.dt / 3.0 if bdf2 else p.dt
    Tfac = 1.0 / (8.0 * PI * p.mu * p.a ** 3)
    I3 = np.eye(3)

    def frame_cross(q, K):
        t, n, b = frame(q)
        return K[0] * rcross(t) + K[1] * rcross(n) + K[2] * rcross(b)

    def diag_block(dOmega, k):
        L = -rcross(f.u[k])
        Om = rcross(Tfac * forces[k, 3:])
        uv = f.u[k]
        return -tsf * (dOmega - 0.5 * (Om + L @ dOmega) + (rcross(Om @ uv) + L @ Om + L @ L @ dOmega) / 12.0)

    def off_block(dOmega, k):
        L = -rcross(f.u[k])
        return -tsf * (dOmega - 0.5 * L @ dOmega + L @ L @ dOmega / 12.0)

    for k in range(N):
        t = frame(f.q[k])[0]
        lamsum = np.zeros(3)
        elastic_d = np.zeros((3, 3))
        if k + 1 < N:                                   # link k (to k+1)
            Kph = body_frame_moment(f.q[k], f.q[k + 1], p)
            R = rot_mat(midpoint(f.q[k], f.q[k + 1]))
            elastic_d += 0.5 * frame_cross(f.q[k], Kph) + R @ body_frame_moment_lie_derivative(
                f.q[k], f.q[k + 1], False, p)
            lamsum += f.lam[k]
            right = 0.5 * frame_cross(f.q[k + 1], Kph) + R @ body_frame_moment_lie_derivative(
                f.q[k], f.q[k + 1], True, p)
            E[3 * k:3 * k + 3, 3 * k + 3:3 * k + 6] = off_block(Tfac * right, k)
        if k > 0:                                       # link k-1 (from k-1)
            Kmh = body_frame_moment(f.q[k - 1], f.q[k], p)
            R = rot_mat(midpoint(f.q[k - 1], f.q[k]))
            elastic_d -= 0.5 * frame_cross(f.q[k], Kmh) + R @ body_frame_moment_lie_derivative(
                f.q[k - 1], f.q[k], True, p)
            lamsum += f.lam[k - 1]
            left = -(0.5 * frame_cross(f.q[k - 1], Kmh) + R @ body_frame_moment_lie_derivative(
                f.q[k - 1], f.q[k], False, p))
            E[3 * k:3 * k + 3, 3 * k - 3:3 * k] = off_block(Tfac * left, k)
        constraint = -0.5 * p.dl * rcross(lamsum) @ rcross(t)
        E[3 * k:3 * k + 3, 3 * k:3 * k + 3] = diag_block(Tfac * (elastic_d + constraint), k)
    return E + np.eye(3 * N)


def approx_jacobian(f: Filament, forces, bdf2, p: Params):
    """The filament's J0 block (6N x 6N) before inversion."""
    N = p.nseg
    J = np.zeros((6 * N, 6 * N))
    fac = -p.dt / (9.0 * PI * p.mu * p.a) if bdf2 else -p.dt / (6.0 * PI * p.mu * p.a)
    for i in range(3 * N):
        J[i, i] = fac
        if i < 3 * N - 3:
            J[i, i + 3] = -fac
    for j in range(1, N):
        jp = 3 * (N + j)
        temp = p.dl * rcross(frame(f.q[j])[0])
        J[3 * j:3 * j + 3, jp:jp + 3] = 0.5 * temp
        for i in range(j + 1, N):
            J[3 * i:3 * i + 3, jp:jp + 3] = temp
    if p.base == "driven":
        # the robot arm moves every later segment with the base's tangent
        # (filament.cpp:1789-1800)
        t0 = 0.5 * p.dl * rcross(frame(f.q[0])[0])
        for i in range(1, N):
            J[3 * i:3 * i + 3, 3 * N:3 * N + 3] = t0
    tN = frame(f.q[N - 1])[0]
    J[3 * (N - 1):3 * N, 3 * (2 * N - 1):6 * N] -= fac * p.end_force * rcross(tN)
    fac2 = (p.dl * p.dt / (24.0 * PI * p.mu * p.a ** 3) if bdf2 else p.dl * p.dt / (16.0 * PI * p.mu * p.a ** 3))
    for j in range(1, N):
        for k, ip in ((j - 1, 3 * (N + j - 1)), (j, 3 * (N + j))):
            umat = rcross(f.u[k])
            tmat = -rcross(frame(f.q[k])[0])
            J[ip:ip + 3, 3 * j:3 * j + 3] = fac2 * (tmat + 0.5 * umat @ tmat + umat @ umat @ tmat / 12.0)
    J[3 * N:, 3 * N:] = lie_algebra_block(f, forces, bdf2, p)
    if p.base == "driven":
        return J
    fac3 = -p.dt / (12.0 * PI * p.mu * p.a ** 3) if bdf2 else -p.dt / (8.0 * PI * p.mu * p.a ** 3)
    umat = rcross(f.u[0])
    blk = np.eye(3) + 0.5 * umat + umat @ umat / 12.0
    blkT = np.eye(3) - 0.5 * umat + umat @ umat / 12.0
    J[3 * N:3 * N + 3, 3 * N:3 * N + 3] = fac3 * blk @ blkT
    J[3 * N + 3:3 * N + 6, 3 * N:3 * N + 3] = 0.0
    return J


def update(f: Filament, du, p: Params):
    """filament::update with the filament's 6N block of the Broyden update."""
    N = p.nseg
    du = np.asarray(du).reshape(2, N, 3)
    f.tether = f.tether + du[0, 0]
    f.lam = f.lam + du[0, 1:]
    if p.base == "driven":
        f.u[0] = f.u[0] + du[1, 0]
        f.q[0] = qmul(lie_exp(f.u[0]), f.qm1[0])
    else:
        f.clamp = f.clamp + du[1, 0]
    for i in range(1, N):
        f.u[i] = f.u[i] + du[1, i]
        f.q[i] = qmul(lie_exp(f.u[i]), f.qm1[i])
    robot_arm(f, p.dl)


def initial_guess(f: Filament, nt, p: Params, num_euler_steps=1):
    """Start of step nt for a filament on a body held still
    (filament::initial_guess, filament.cpp:1083-1385): segment history shift
    and u extrapolation (segment::initial_guess); a clamped base segment then
    turns by its fixed u0 (zero for the INSTABILITY_CILIA build, whose base
    takes the still body's rotation, accept_state_from_rigid_body,
    filament.cpp:649-672), a driven one keeps its extrapolated u; then the
    multipliers' extrapolation and the robot arm."""
    N = p.nseg
    f.xm2, f.xm1 = f.xm1.copy(), f.x.copy()
    if nt <= num_euler_steps:
        f.um1 = f.u.copy()
    else:
        f.u, f.um1 = 2.0 * f.u - f.um1, f.u.copy()
    f.qm1 = f.q.copy()
    for i in range(N):
        f.q[i] = qmul(lie_exp(f.u[i]), f.qm1[i])
    if p.base != "driven":
        f.u[0] = f.u0
        f.q[0] = qmul(lie_exp(f.u[0]), f.qm1[0])
    if nt == 1:
        f.lamm1, f.tetherm1, f.clampm1 = f.lam.copy(), f.tether.copy(), f.clamp.copy()
    elif nt == 2:
        for name in ("lam", "tether", "clamp"):
            cur, m1 = getattr(f, name), getattr(f, name + "m1")
            setattr(f, name + "m2", m1.copy())
            setattr(f, name + "m1", cur.copy())
            setattr(f, name, 2.0 * cur - m1)
    elif nt > 2:
        for name in ("lam", "tether", "clamp"):
            cur, m1, m2 = getattr(f, name), getattr(f, name + "m1"), getattr(f, name + "m2")
            setattr(f, name + "m2", m1.copy())
            setattr(f, name + "m1", cur.copy())
            setattr(f, name, 3.0 * (cur - m1) + m2)
    robot_arm(f, p.dl)


def self_mobility_residual(f: Filament, xm1, xm2, um1, bdf2, p: Params):
    """Residual of one filament with self-mobility only (v = F/(6 pi mu a),
    omega = T/(8 pi mu a^3)) in the order of compute_errors
    (mobility_solver.cpp:2685-2760): the approximate Jacobian differentiates it."""
    F = internal_forces(f, p)
    v = F[:, :3] / (6 * PI * p.mu * p.a)
    w = F[:, 3:] / (8 * PI * p.mu * p.a ** 3)
    N = p.nseg
    e = np.zeros((2, N, 3))
    for j in range(N):
        W = dexpinv(f.u[j], w[j])
        if bdf2:
            e[0, j] = f.x[j] - 4.0 / 3.0 * xm1[j] + 1.0 / 3.0 * xm2[j] - 2.0 / 3.0 * p.dt * v[j]
            e[1, j] = f.u[j] - 1.0 / 3.0 * um1[j] - 2.0 / 3.0 * p.dt * W
        else:
            e[0, j] = f.x[j] - xm1[j] - p.dt * v[j]
            e[1, j] = f.u[j] - p.dt * W
    return e.reshape(-1)
