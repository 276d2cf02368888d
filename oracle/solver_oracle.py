"""FP64 numpy restatement of the reference's per-timestep solve layer.

TEST INFRASTRUCTURE ONLY (the checker; the product never imports it).

* beat_geometry      filament::initial_guess with FIT_TO_DATA_BEAT and
                     DYNAMIC_SHAPE_ROTATION, src/cilia/filament.cpp:1083-1273,
                     with the fitted shape of filament.cpp:770-833.
* saddle_matrix      system_matrix_mult_new, prescribed cilia and body velocity,
                     src/cilia/mobility/mobility_solver.cpp:1806-2090 (dense).
* precondition_ref   apply_preconditioner_new, mobility_solver.cpp:1040-1116
                     (a loop restatement; pinned by P @ P^{-1} v = v).
* gmres_ref          solve_linear_system, mobility_solver.cpp:2092-2570: right
                     preconditioning, modified Gram-Schmidt, Givens rotations,
                     warm start and rubbish-guess rules, no restart.
* hermite_ref        cyclic_cubic_spline_interpolation, mobility_solver.cpp:240-280.
* rhs_ref            assemble_rhs, mobility_solver.cpp:282-440.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


def fitted_shape(s, phase, A, B):
    """(pos_xy, dir_xy) of filament.cpp:798-833 for one s."""
    sv = np.array([s, s * s, s * s * s])
    cos_vec = np.array([1.0] + [np.cos(n * phase) for n in (1, 2, 3)])
    sin_vec = np.array([np.sin(n * phase) for n in (1, 2, 3)])
    pos = np.array([(cos_vec @ A[ax] + sin_vec @ B[ax]) @ sv for ax in range(2)])
    dsin = np.array([0.0] + [-n * np.sin(n * phase) for n in (1, 2, 3)])
    dcos = np.array([n * np.cos(n * phase) for n in (1, 2, 3)])
    der = np.array([(dsin @ A[ax] + dcos @ B[ax]) @ sv for ax in range(2)])
    return pos, der


def beat_geometry(setup, num_seg, A, B, phase=None, angle=None):
    """Segment positions, vel_dir_phase and vel_dir_angle (3 per segment)."""
    nf = len(setup.omega0)
    phase = setup.phase0 if phase is None else phase
    angle = setup.angle0 if angle is None else angle
    fr = np.asarray(setup.frames).reshape(nf, 3, 3)    # rows e1, e2, e3
    base = np.asarray(setup.bases).reshape(nf, 3)
    L = setup.fil_length
    xs = np.zeros((nf, num_seg, 3))
    kp = np.zeros((nf, num_seg, 3))
    kt = np.zeros((nf, num_seg, 3))
    for m in range(nf):
        R = fr[m].T                                    # columns e1, e2, e3: local -> world
        c, s = np.cos(angle[m]), np.sin(angle[m])
        Rshape = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
        xs[m, 0] = base[m]
        for n in range(1, num_seg):
            pos, der = fitted_shape(setup.s_values[n], phase[m], A, B)
            ref = L * (Rshape @ np.array([pos[0], pos[1], 0.0]))
            xs[m, n] = base[m] + R @ ref
            kp[m, n] = L * (R @ (Rshape @ np.array([der[0], der[1], 0.0])))
            kt[m, n] = R @ np.array([-ref[1], ref[0], 0.0])     # e_z x ref
    return {"x_segs": xs.reshape(-1), "kphi": kp.reshape(-1), "ktheta": kt.reshape(-1)}


def k_matrix(kphi, ktheta, nfil, nseg, nblob, dyn_phase=True, dyn_angle=True):
    """Columns K_phi / K_theta as an (3*Ns + 3*Nb) x (#gen coords) matrix."""
    ns3 = 3 * nfil * nseg
    cols = []
    for which, kk in ((dyn_phase, kphi), (dyn_angle, ktheta)):
        if not which:
            continue
        K = np.zeros((ns3 + 3 * nblob, nfil))
        for m in range(nfil):
            K[3 * nseg * m: 3 * nseg * (m + 1), m] = kk[3 * nseg * m: 3 * nseg * (m + 1)]
        cols.append(K)
    return np.concatenate(cols, axis=1) if cols else np.zeros((ns3 + 3 * nblob, 0))


def k_body_matrix(x_segs, x_blobs, x_bodies, num_swim):
    """Body columns K_b (mobility_solver.cpp:1850-1908): for swimmer s, the
    velocity V + Omega x (x - x_body_s) of its particles; segments and blobs
    split evenly between the swimmers."""
    xs = np.asarray(x_segs, dtype=float).reshape(-1, 3)
    xb = np.asarray(x_blobs, dtype=float).reshape(-1, 3)
    xbod = np.asarray(x_bodies, dtype=float).reshape(num_swim, 3)
    ns, nb = len(xs), len(xb)
    K = np.zeros((3 * (ns + nb), 6 * num_swim))
    for rows, x, per in ((0, xs, ns // num_swim), (3 * ns, xb, nb // num_swim)):
        for i in range(len(x)):
            s = i // per
            r = x[i] - xbod[s]
            R = rows + 3 * i
            K[R:R + 3, 6 * s:6 * s + 3] = np.eye(3)
            # Omega x r = -[r]x Omega
            K[R:R + 3, 6 * s + 3:6 * s + 6] = np.array([[0.0, r[2], -r[1]], [-r[2], 0.0, r[0]], [r[1], -r[0], 0.0]])
    return K


def _with_bodies(K, x_segs, x_blobs, x_bodies, num_swim):
    if x_bodies is None:
        return K
    return np.concatenate([k_body_matrix(x_segs, x_blobs, x_bodies, num_swim), K], axis=1)


def saddle_matrix(x_segs, x_blobs, kphi, ktheta, nfil, nseg, *, dyn_phase=True, dyn_angle=True,
                  a_seg=1.0, a_blob=1.0, mu=1.0, x_bodies=None, num_swim=1):
    """Dense A = [[M, -K], [-K^T, 0]] with M the <3,3> grand mobility; with
    x_bodies, K = [K_b, K_phi, K_theta] (free bodies)."""
    nblob = len(x_blobs) // 3
    M = O.dense_mobility(x_segs, x_blobs, seg_dof=3, a_seg=a_seg, a_blob=a_blob, mu=mu)
    K = _with_bodies(k_matrix(kphi, ktheta, nfil, nseg, nblob, dyn_phase, dyn_angle), x_segs, x_blobs, x_bodies,
                     num_swim)
    g = K.shape[1]
    return np.block([[M, -K], [-K.T, np.zeros((g, g))]])


def precond_matrix(kphi, ktheta, nfil, nseg, nblob, *, dyn_phase=True, dyn_angle=True, a_seg=1.0,
                   a_blob=1.0, mu=1.0, x_segs=None, x_blobs=None, x_bodies=None, num_swim=1):
    """P = [[D, -K], [-K^T, 0]], D = diag(1/(6 pi mu a)): the approximation
    whose exact inverse apply_preconditioner_new applies (with x_bodies, K
    carries the body columns too)."""
    ns3 = 3 * nfil * nseg
    D = np.diag(np.concatenate([np.full(ns3, 1.0 / (6 * np.pi * mu * a_seg)),
                                np.full(3 * nblob, 1.0 / (6 * np.pi * mu * a_blob))]))
    K = _with_bodies(k_matrix(kphi, ktheta, nfil, nseg, nblob, dyn_phase, dyn_angle), x_segs, x_blobs, x_bodies,
                     num_swim)
    g = K.shape[1]
    return np.block([[D, -K], [-K.T, np.zeros((g, g))]])


def precondition_ref(v, kphi, ktheta, nfil, nseg, nblob, *, a_seg=1.0, a_blob=1.0, mu=1.0):
    """apply_preconditioner_new (dynamic phase and shape rotation), loop form."""
    out = np.array(v, dtype=float)
    gs = 6.0 * np.pi * mu * a_seg
    gb = 6.0 * np.pi * mu * a_blob
    ns3 = 3 * nfil * nseg
    off_p = ns3 + 3 * nblob
    off_a = off_p + nfil
    out[:ns3] *= gs
    out[ns3:off_p] *= gb
    out[off_p:off_a + nfil] *= -1.0
    for m in range(nfil):
        sl = slice(3 * nseg * m, 3 * nseg * (m + 1))
        kp, kt = kphi[sl], ktheta[sl]
        c11, c22, c12 = kp @ kp, kt @ kt, kp @ kt
        out[off_p + m] -= kp @ out[sl]
        out[off_a + m] -= kt @ out[sl]
        pm, am = out[off_p + m], out[off_a + m]
        det = c11 * c22 - c12 * c12
        out[off_p + m] = (pm * c22 / det - am * c12 / det) / gs
        out[off_a + m] = (-pm * c12 / det + am * c11 / det) / gs
        out[sl] += gs * (out[off_p + m] * kp + out[off_a + m] * kt)
    return out


def gmres_ref(apply_A, apply_P, rhs, x0, tol, max_iter):
    """Returns (x, iterations).  Restates solve_linear_system (right
    preconditioning branch) including its early exits."""
    n = len(rhs)
    bnorm = np.linalg.norm(rhs)
    if bnorm == 0.0:
        return np.zeros(n), 0
    x = np.array(x0, dtype=float)
    q0 = rhs - apply_A(x)
    beta0 = np.linalg.norm(q0)
    if beta0 <= tol * bnorm:
        return x, 0
    if beta0 > bnorm:
        x[:] = 0.0
        beta0 = bnorm
        q0 = rhs.copy()
    max_iter = min(n, max_iter)
    Q = np.zeros((n, max_iter + 1))
    H = np.zeros((max_iter + 1, max_iter + 1))
    beta = np.zeros(max_iter + 1)
    cs, sn = np.zeros(max_iter), np.zeros(max_iter)
    Q[:, 0] = q0 / beta0
    beta[0] = beta0
    for it in range(1, max_iter + 1):
        w = apply_A(apply_P(Q[:, it - 1]))
        for i in range(it):
            H[i, it - 1] = w @ Q[:, i]
            w = w - H[i, it - 1] * Q[:, i]
        H[it, it - 1] = np.linalg.norm(w)
        Q[:, it] = w / H[it, it - 1] if H[it, it - 1] != 0 else w
        for i in range(1, it):
            t = cs[i - 1] * H[i - 1, it - 1] + sn[i - 1] * H[i, it - 1]
            H[i, it - 1] = cs[i - 1] * H[i, it - 1] - sn[i - 1] * H[i - 1, it - 1]
            H[i - 1, it - 1] = t
        a, b = H[it - 1, it - 1], H[it, it - 1]
        if b == 0.0:
            c, s = 1.0, 0.0
        elif abs(b) > abs(a):
            t = a / b
            s = (1.0 if b > 0 else -1.0) / np.sqrt(1.0 + t * t)
            c = t * s
        else:
            t = b / a
            c = (1.0 if a > 0 else -1.0) / np.sqrt(1.0 + t * t)
            s = t * c
        cs[it - 1], sn[it - 1] = c, s
        H[it - 1, it - 1] = c * a + s * b
        H[it, it - 1] = 0.0
        beta[it] = -s * beta[it - 1]
        beta[it - 1] = c * beta[it - 1]
        if abs(beta[it]) / bnorm <= tol or it == max_iter:
            y = np.zeros(it)
            for i in range(it - 1, -1, -1):
                y[i] = (beta[i] - H[i, i + 1:it] @ y[i + 1:it]) / H[i, i]
            x = x + apply_P(Q[:, :it] @ y)
            return x, it
    return x, max_iter


def hermite_ref(phase, table):
    """cyclic_cubic_spline_interpolation, mobility_solver.cpp:240-280."""
    nt = len(table)
    u = 0.5 * phase / np.pi
    u -= np.floor(u)
    u *= nt
    lo = int(u)
    if lo == nt:
        return table[0]
    hi = 0 if lo == nt - 1 else lo + 1
    ql = table[lo]
    dql = (table[1] - table[0]) if lo == 0 else 0.5 * (table[hi] - table[lo - 1])
    qh = table[hi]
    dqh = (table[0] - table[nt - 1]) if hi == 0 else 0.5 * (table[0 if hi == nt - 1 else hi + 1] - table[lo])
    t = u - lo
    return ql * (2 * t**3 - 3 * t**2 + 1) + dql * (t**3 - 2 * t**2 + t) + qh * (3 * t**2 - 2 * t**3) + dqh * (t**3 - t**2)


def rhs_ref(setup, num_seg, nblob, x_segs, x_blobs, body_x, body_v, *, gen_force_factor=1.0,
            torsion_factor=0.046, phase=None, angle=None):
    """assemble_rhs (dynamic phase + shape rotation, prescribed body velocity)."""
    nf = len(setup.omega0)
    phase = setup.phase0 if phase is None else phase
    angle = setup.angle0 if angle is None else angle
    ns = nf * num_seg
    out = np.zeros(3 * ns + 3 * nblob + 2 * nf)
    V, W = np.asarray(body_v[:3]), np.asarray(body_v[3:])
    pts = np.concatenate([np.asarray(x_segs).reshape(-1, 3), np.asarray(x_blobs).reshape(-1, 3)])
    r = pts - np.asarray(body_x)
    out[:3 * (ns + nblob)] = (V + np.cross(W, r)).reshape(-1)
    L = setup.fil_length
    for m in range(nf):
        w = 0.5 * setup.omega0[m] / np.pi
        qa = hermite_ref(phase[m], setup.gen_angle) * w
        qa -= torsion_factor * (setup.omega0[m] * L**3) * angle[m]
        qp = hermite_ref(phase[m], setup.gen_phase) * w * gen_force_factor
        out[3 * (ns + nblob) + m] = -qp
        out[3 * (ns + nblob) + nf + m] = -qa
    return out
