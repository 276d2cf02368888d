"""CPU restatement of a free swimmer in the elastic implicit timestep (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
free-body elastic step; the product never imports it.

A swimmer is one rigid body (its blobs) carrying NFIL clamped filaments of the
Cosserat model (oracle/filament_oracle.py); the body is free
(PRESCRIBED_BODY_VELOCITIES false) and the blob forces come from the blob-only
mobility system with the body's velocity as extra unknowns (SOLVER_TYPE 1, the
reference solves it by GMRES; a dense solve here):

  rigid_body::initial_guess, update           src/cilia/rigid_body.cpp:395-447, 506-535
  swimmer::initial_guess, update              src/cilia/swimmer.cpp:433-467, 1072-1098
  filament::accept_state_from_rigid_body      src/cilia/filament.cpp:649-672
  swimmer::forces_and_torques (free body)     swimmer.cpp:469-533
  blob-only system with body rows             mobility_solver.cpp:544-560 (rhs),
                                              1806-2090 (operator), 2574-2659 (velocities)
  make_body_reference_matrices                mobility_solver.cpp:2950-3047
  swimmer::prepare_jacobian_inv               swimmer.cpp:535-553, 688-880 (the
                                              GMRES branch, free body)
  swimmer::jacobian_inv_mult                  swimmer.cpp:915-921, 1012-1066
  the Broyden-for-everything form, free body  swimmer.cpp:565-676 (bfe_coupling),
                                              931-1010 (BfeInverse); the loop
                                              free_bfe_loop

Broyden vector of the swimmer: NFIL filament blocks of 6 NSEG, then the body's
6 unknowns [position (3), Lie-algebra element (3)] (then, Broyden for
everything, 3 NBLOB blob forces).

What pins it (tests/test_oracle_swimmer.py): the coupled inverse against a
dense solve of the assembled block Jacobian; anchoring and quadratic
extrapolation on closed forms; the free blob solve's force and torque balance
and rigid blob motion; the body mobility of a sphere of blobs against Stokes
drag; convergence of the free swimmer, with fewer iterations than an identity
body block; the Broyden-for-everything loop reaching the trajectory of the
form that solves for the blob forces.  tests/test_oracle_body_coupling.py pins
body_coupling entry by entry: A, B and S0 against central differences of the
swimmer residual they linearise (self-mobility segments, the body's 6 x 6
mobility, the sphere term from the pinned pair tensor), Euler and BDF2, with
and without the sphere, the reference's own simplifications checked as their
exact closed-form differences; and blob_force_guess as exact on linear and
quadratic force histories.  tests/test_oracle_bfe_free.py pins the
Broyden-for-everything form with a free body the same way: BfeInverse as the
exact inverse of its block matrix, the blocks against central differences of
the residual with self-mobility segments and blobs (the reference's left-out
end-force lever and BDF2 angular-velocity terms checked as their closed
forms), and the loop reaching the GMRES form's trajectory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import filament_oracle as F

PI = np.pi


@dataclass
class Body:
    x: np.ndarray
    xm1: np.ndarray
    xm2: np.ndarray
    u: np.ndarray
    um1: np.ndarray
    q: np.ndarray
    qm1: np.ndarray
    Q_init: np.ndarray

    @staticmethod
    def at_rest(x, q=(1.0, 0.0, 0.0, 0.0)):
        x = np.array(x, dtype=float)
        q = np.array(q, dtype=float)
        return Body(x=x.copy(), xm1=x.copy(), xm2=x.copy(), u=np.zeros(3), um1=np.zeros(3), q=q.copy(),
                    qm1=q.copy(), Q_init=F.rot_mat(q))

    def copy(self):
        return Body(**{k: v.copy() for k, v in self.__dict__.items()})


def body_initial_guess(b: Body, nt, num_euler_steps=1):
    """rigid_body::initial_guess (rigid_body.cpp:395-447)."""
    if nt <= num_euler_steps:
        gx, gu = 2.0 * b.x - b.xm1, b.u.copy()
    else:
        gx, gu = 3.0 * (b.x - b.xm1) + b.xm2, 2.0 * b.u - b.um1
    b.xm2, b.xm1, b.x = b.xm1.copy(), b.x.copy(), gx
    b.um1, b.u = b.u.copy(), gu
    b.qm1 = b.q.copy()
    b.q = F.qmul(F.lie_exp(b.u), b.qm1)
    b.Q_init = F.rot_mat(b.q)


def body_update(b: Body, du):
    """rigid_body::update (rigid_body.cpp:506-535)."""
    b.x = b.x + du[:3]
    b.u = b.u + du[3:6]
    b.q = F.qmul(F.lie_exp(b.u), b.qm1)


def accept_state(f: F.Filament, base, u_body, p: F.Params):
    """filament::accept_state_from_rigid_body: the base segment sits at the
    body's anchor and turns with the body (segment::update to u_body)."""
    f.x[0] = base
    f.u[0] = u_body
    f.q[0] = F.qmul(F.lie_exp(f.u[0]), f.qm1[0])
    F.robot_arm(f, p.dl)


def swimmer_initial_guess(b: Body, fils, refs, nt, p: F.Params):
    body_initial_guess(b, nt)
    R = F.rot_mat(b.q)
    for f, ref in zip(fils, refs):
        F.initial_guess(f, nt, p)
        accept_state(f, b.x + R @ ref, b.u, p)


def swimmer_update(b: Body, fils, refs, du, p: F.Params):
    """swimmer::update: the body first, then each filament from its new anchor."""
    n6 = 6 * p.nseg
    body_update(b, du[n6 * len(fils):n6 * len(fils) + 6])
    R = F.rot_mat(b.q)
    for i, (f, ref) in enumerate(zip(fils, refs)):
        accept_state(f, b.x + R @ ref, b.u, p)
        F.update(f, du[n6 * i:n6 * (i + 1)], p)


def blob_positions(b: Body, blob_refs):
    """rigid_body::blob_positions (rigid_body.cpp:482-500)."""
    return (b.x + np.asarray(blob_refs).reshape(-1, 3) @ F.rot_mat(b.q).T).reshape(-1)


def body_force(b: Body, fils, forces):
    """swimmer.f of the free body: minus the total force and torque (about the
    body's centre) of its filaments' segments."""
    out = np.zeros(6)
    for fl, Fi in zip(fils, forces):
        Fi = np.asarray(Fi).reshape(-1, 6)
        out[:3] -= Fi[:, :3].sum(0)
        out[3:] -= (Fi[:, 3:] + np.cross(fl.x - b.x, Fi[:, :3])).sum(0)
    return out


def k_blobs(x_blobs, x_body):
    """K: 3 NBLOB x 6, V + Omega x (x_blob - x_body) per blob."""
    xb = np.asarray(x_blobs).reshape(-1, 3)
    K = np.zeros((3 * len(xb), 6))
    for m, x in enumerate(xb):
        r = x - x_body
        K[3 * m:3 * m + 3, :3] = np.eye(3)
        K[3 * m:3 * m + 3, 3:] = np.array([[0.0, r[2], -r[1]], [-r[2], 0.0, r[0]], [r[1], -r[0], 0.0]])
    return K


def free_blob_solve(M_bb, x_blobs, x_body, v_blobs_from_segs, swimmer_f):
    """The blob-only system with the body's velocity as unknowns:
    M_bb lam - K V = -v_blobs,  -K^T lam = -swimmer.f (mobility_solver.cpp:544-560
    and its operator); returns (lam, V)."""
    nb3 = M_bb.shape[0]
    K = k_blobs(x_blobs, x_body)
    A = np.zeros((nb3 + 6, nb3 + 6))
    A[:nb3, :nb3] = M_bb
    A[:nb3, nb3:] = -K
    A[nb3:, :nb3] = -K.T
    rhs = np.concatenate([-np.asarray(v_blobs_from_segs), -np.asarray(swimmer_f)])
    sol = np.linalg.solve(A, rhs)
    return sol[:nb3], sol[nb3:]


def body_mobility_reference(M_bb_ref, blob_refs):
    """make_body_reference_matrices: the 6 x 6 body mobility (V, Omega from a
    unit force or torque on the body, no filaments) at the reference pose."""
    K = k_blobs(blob_refs, np.zeros(3))
    return np.linalg.inv(K.T @ np.linalg.solve(M_bb_ref, K))


def clamping_blocks(f: F.Filament, forces, bdf2, p: F.Params):
    """elastic_clamping_block1, 2: the rows of segments 0 and 1 in the Lie-algebra
    block's first column, before invert_approx_jacobian overwrites them
    (filament.cpp:1883-1884)."""
    E = F.lie_algebra_block(f, forces, bdf2, p)
    return E[0:3, 0:3].copy(), E[3:6, 0:3].copy()


def rotate_mobility(N_ref, Q):
    out = np.zeros((6, 6))
    for r in (0, 3):
        for c in (0, 3):
            out[r:r + 3, c:c + 3] = Q @ N_ref[r:r + 3, c:c + 3] @ Q.T
    return out


def body_coupling(b: Body, fils, refs, forces, nt, N_ref, p: F.Params, sphere_radius=0.0,
                  num_euler_steps=1):
    """swimmer::prepare_jacobian_inv, GMRES branch with a free body
    (swimmer.cpp:688-880), before the Schur complement is formed: the body block
    S0, and per filament A_i (its rows, the body's columns: 6 NSEG x 6) and B_i
    (the body's rows, its columns: 6 x 6 NSEG); J0 = [[diag J_i, A], [B, S0]]."""
    N, dl, EF = p.nseg, p.dl, p.end_force
    dt_fac = -p.dt if nt < num_euler_steps else -2.0 * p.dt / 3.0
    u = b.u
    L = -F.rcross(u)                       # body_lie_cross
    cbo = -L / dt_fac                      # cross_body_omega
    S0 = np.eye(6)
    S0[3:, 3:] += dt_fac * (L @ cbo / 12.0 - 0.5 * cbo)
    Q = b.Q_init
    old = np.zeros((3, 3))
    for f, ref in zip(fils, refs):
        d = Q @ ref
        old -= F.rcross(f.tether) @ F.rcross(d)
        uxl = np.cross(u, f.clamp)
        old += 0.5 * F.rcross(f.clamp) - (F.rcross(uxl) + L @ F.rcross(f.clamp)) / 12.0
        t = F.frame(f.q[N - 1])[0]
        nd = np.linalg.norm(d)
        old += F.rcross(EF * t) @ F.rcross((0.5 * dl + nd) * d / nd)
    Nb = rotate_mobility(N_ref, Q)
    old = Nb[3:, 3:] @ old
    S0[3:, 3:] += dt_fac * (old - 0.5 * L @ old + L @ L @ old / 12.0)
    Dinv = np.eye(3) + L @ L / 12.0 - 0.5 * L
    DinvT = np.eye(3) + L @ L / 12.0 + 0.5 * L
    As, Bs = [], []
    a, mu = p.a, p.mu
    for i, (f, ref) in enumerate(zip(fils, refs)):
        d = Q @ ref
        B = np.zeros((6, 6 * N))
        B[0:3, 0:3] = -np.eye(3)
        B[3:6, 0:3] = F.rcross(d)
        B[3:6, 3 * N:3 * N + 3] = -DinvT
        t = F.frame(f.q[N - 1])[0]
        B[0:3, 6 * N - 3:] = F.rcross(EF * t)
        diff = f.x[N - 1] - b.x
        B[3:6, 6 * N - 3:] = F.rcross(0.5 * dl * t - diff) @ B[0:3, 6 * N - 3:]
        for m in range(1, N - 1):
            tm = F.frame(f.q[m])[0]
            B[3:6, 3 * (N + m):3 * (N + m) + 3] = B[0:3, 6 * N - 3:] @ F.rcross(dl * tm)
        B = Nb @ B
        if sphere_radius > 0.0:
            R = sphere_radius
            for n in range(N):
                rv = f.x[n] - b.x
                r = np.linalg.norm(rv)
                rh = rv / r
                rr = np.outer(rh, rh)
                Mtt = ((1.0 + (R * R + a * a) / (3.0 * r * r)) * np.eye(3) +
                       (1.0 - (R * R + a * a) / (r * r)) * rr) / (8.0 * PI * mu * r)
                Mtr = F.rcross(rh) / (8.0 * PI * mu * r * r)
                Mrr = (3.0 * rr - np.eye(3)) / (16.0 * PI * mu * r ** 3)
                tmp = F.rcross(F.frame(f.q[n])[0])
                if n == 0:
                    B[0:3, 0:3] += Mtt
                    B[3:6, 0:3] += Mtr
                    B[0:3, 3 * N:3 * N + 3] += Mtr @ DinvT
                    B[3:6, 3 * N:3 * N + 3] += Mrr @ DinvT
                    B[0:3, 3:6] += 0.5 * dl * Mtr @ tmp - Mtt
                    B[3:6, 3:6] += 0.5 * dl * Mrr @ tmp - Mtr
                elif n == N - 1:
                    B[0:3, 6 * N - 3:] += -EF * Mtt @ tmp
                    B[3:6, 6 * N - 3:] += -EF * Mtr @ tmp
                    B[0:3, 3 * N - 3:3 * N] += 0.5 * dl * Mtr @ tmp + Mtt
                    B[3:6, 3 * N - 3:3 * N] += 0.5 * dl * Mrr @ tmp + Mtr
                else:
                    B[0:3, 3 * n:3 * n + 3] += 0.5 * dl * Mtr @ tmp + Mtt
                    B[3:6, 3 * n:3 * n + 3] += 0.5 * dl * Mrr @ tmp + Mtr
                    B[0:3, 3 * (n + 1):3 * (n + 1) + 3] += 0.5 * dl * Mtr @ tmp - Mtt
                    B[3:6, 3 * (n + 1):3 * (n + 1) + 3] += 0.5 * dl * Mrr @ tmp - Mtr
        for n in range(N):
            B[3:6, 3 * n:3 * n + 3] = Dinv @ B[3:6, 3 * n:3 * n + 3]
            B[3:6, 3 * (n + N):3 * (n + N) + 3] = Dinv @ B[3:6, 3 * (n + N):3 * (n + N) + 3]
        B *= dt_fac
        c1, c2 = clamping_blocks(f, np.asarray(forces[i]).reshape(N, 6), nt >= num_euler_steps, p)
        A = np.zeros((6 * N, 6))
        nd = np.linalg.norm(d)
        Dm = F.rcross((0.5 * dl + nd) * d / nd)
        for n in range(N):
            A[3 * n:3 * n + 3, 0:3] = np.eye(3)
            A[3 * n:3 * n + 3, 3:6] = Dm
        A[3 * N:3 * N + 3, 3:6] = c1
        A[3 * N + 3:3 * N + 6, 3:6] = c2
        As.append(A)
        Bs.append(B)
    return S0, As, Bs


class CoupledInverse:
    """J0^{-1} of one swimmer through the body's Schur complement
    (prepare_jacobian_inv's schur_mat_inv and jacobian_B_blocks, applied as
    jacobian_inv_mult)."""

    def __init__(self, J_blocks, S0, As, Bs):
        self.Jinv = [np.linalg.inv(J) for J in J_blocks]
        self.BJ = [B @ Ji for B, Ji in zip(Bs, self.Jinv)]
        S = S0.copy()
        for BJ, A in zip(self.BJ, As):
            S -= BJ @ A
        self.Sinv = np.linalg.inv(S)
        self.As = As
        self.n6 = J_blocks[0].shape[0]

    def apply(self, x):
        n6, nf = self.n6, len(self.Jinv)
        xb = x[n6 * nf:n6 * nf + 6].copy()
        for i in range(nf):
            xb -= self.BJ[i] @ x[n6 * i:n6 * (i + 1)]
        out = np.zeros_like(x)
        ob = self.Sinv @ xb
        out[n6 * nf:n6 * nf + 6] = ob
        for i in range(nf):
            out[n6 * i:n6 * (i + 1)] = self.Jinv[i] @ (x[n6 * i:n6 * (i + 1)] - self.As[i] @ ob)
        return out


def assembled_jacobian(J_blocks, S0, As, Bs):
    """The whole J0 of one swimmer, [[diag J_i, A], [B, S0]]."""
    n6, nf = J_blocks[0].shape[0], len(J_blocks)
    n = n6 * nf + 6
    J = np.zeros((n, n))
    for i in range(nf):
        s = slice(n6 * i, n6 * (i + 1))
        J[s, s] = J_blocks[i]
        J[s, n - 6:] = As[i]
        J[n - 6:, s] = Bs[i]
    J[n - 6:, n - 6:] = S0
    return J


class BroydenOp:
    """broyden_solver.cpp:26-129 with J0^{-1} given as an operator (the
    swimmer's coupled inverse) instead of diagonal blocks."""

    def __init__(self, n, max_iter, confidence, jinv):
        self.n, self.max_iter, self.c, self.jinv = n, max_iter, confidence, jinv
        self.C = np.zeros((n, max_iter))
        self.D = np.zeros((n, max_iter))
        self.error = np.zeros(n)
        self.iter = 0

    def start(self, error):
        self.error = np.array(error, dtype=np.float64)
        self.iter = 0

    def _inverse(self, e):
        u = -self.c * self.jinv(e)
        for i in range(self.iter):
            u -= (e @ self.D[:, i]) * self.C[:, i]
        return u

    def find_update(self):
        return self._inverse(self.error)

    def end_of_iter(self, new_error, error_is_too_large):
        new_error = np.asarray(new_error, dtype=np.float64)
        if error_is_too_large and self.iter < self.max_iter - 1:
            k = self.iter
            ck = self._inverse(new_error)
            dk = new_error - self.error
            nrm = np.linalg.norm(dk)
            self.C[:, k] = ck / nrm
            self.D[:, k] = dk / nrm
            self.error = new_error.copy()
        self.iter += 1


def free_swimmer_loop(b: Body, fils, refs, blob_refs, p: F.Params, steps, tol, perturbation, N_ref,
                      sphere_radius=0.0, coupled=True, max_iter=400, confidence=0.1):
    """The elastic time loop (cilia_sim_main.cu:353-502) for one free swimmer:
    initial guess, forces, velocities, errors, J0 from prepare_jacobian_inv
    (or, coupled=False, the filament blocks with an identity body block), then
    Broyden until the residual meets TOL.  `perturbation` is (NFIL, 6): the
    first-step kick on segments 4 and 12.  Returns per step the Broyden
    iterations and copies of the state."""
    from oracle import broyden_oracle as BO
    from oracle import oracle as O

    N, nfil = p.nseg, len(fils)
    n = 6 * (nfil * N + 1)
    cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
    out = []
    for nt in range(steps):
        swimmer_initial_guess(b, fils, refs, nt, p)

        def evaluate():
            kicks = [{4: perturbation[i, :3], 12: perturbation[i, 3:]} if nt == 0 else None for i in range(nfil)]
            forces = [F.internal_forces(fl, p, kicks[i]) for i, fl in enumerate(fils)]
            x, f = cat("x"), np.concatenate([fc.reshape(-1) for fc in forces])
            xb = blob_positions(b, blob_refs)
            nb3 = xb.size
            vb = O.block("bs", x, xb, f, np.zeros(nb3), seg_dof=6, bs_accumulate=False)
            lam, V = free_blob_solve(O.dense_mobility(np.zeros(0), xb), xb, b.x, vb, body_force(b, fils, forces))
            vs = O.block("ss", x, xb, f, lam, seg_dof=6)
            vs = O.block("sb", x, xb, f, lam, vs, seg_dof=6)
            st = {"seg_x": x, "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"),
                  "seg_um1": cat("um1"), "v_segs": vs, "body_x": b.x, "body_xm1": b.xm1, "body_xm2": b.xm2,
                  "body_u": b.u, "body_um1": b.um1, "v_bodies": V}
            e, big = BO.compute_errors(st, 1, nfil, N, nt >= 1, p.dt, p.dl, tol)
            return forces, e, big, lam, V

        forces, e, big, lam, V = evaluate()
        Js = [F.approx_jacobian(fl, forces[i], nt >= 1, p) for i, fl in enumerate(fils)]
        if coupled:
            S0, As, Bs = body_coupling(b, fils, refs, forces, nt, N_ref, p, sphere_radius)
            jinv = CoupledInverse(Js, S0, As, Bs).apply
        else:
            blocks = [np.linalg.inv(J) for J in Js] + [np.eye(6)]
            jinv = lambda v, blocks=blocks: BO.block_diag_apply(blocks, v)
        ref = BroydenOp(n, max_iter, confidence, jinv)
        ref.start(e)
        while big and ref.iter < max_iter:
            swimmer_update(b, fils, refs, ref.find_update(), p)
            forces, e, big, lam, V = evaluate()
            ref.end_of_iter(e, big)
        out.append({"iters": ref.iter, "converged": not big, "x": cat("x"), "body": b.copy(),
                    "forces": [fc.copy() for fc in forces], "f_blobs": lam, "v_body": V, "error": e})
    return out


def bfe_coupling(b: Body, fils, refs, blob_refs, f_blobs, nt, p: F.Params, blob_radius=1.0,
                 num_euler_steps=1):
    """swimmer::prepare_jacobian_inv, Broyden-for-everything branch with a free
    body (swimmer.cpp:565-676), before the Schur matrix is inverted: S (6 x 6,
    with the blob terms of the diagonal blob approximation), per filament D_n
    (the body rows' dependence on its unknowns, 6 x 6 NSEG: force and torque
    balance) and B_n (its rows' dependence on the body's unknowns, 6 NSEG x 6),
    and the blob displacements d_m = Q_init ref_m; jacobian_B_blocks are
    D_n J_n^{-1} (bfe_inverse)."""
    N, dl, EF = p.nseg, p.dl, p.end_force
    Q, u = b.Q_init, b.u
    bmf = (-6.0 if nt < num_euler_steps else -9.0) * PI * p.mu * blob_radius / p.dt
    refs_b = np.asarray(blob_refs, dtype=float).reshape(-1, 3)
    fb = np.asarray(f_blobs, dtype=float).reshape(-1, 3)
    nb = len(refs_b)
    Smat = np.zeros((6, 6))
    Smat[:3, :3] = -nb * bmf * np.eye(3)
    disp = refs_b @ Q.T
    for m in range(nb):
        rd = F.rcross(disp[m])
        Smat[3:, 3:] += F.rcross(fb[m]) @ rd
        Smat[:3, 3:] -= bmf * rd
        Smat[3:, :3] += bmf * rd
        Smat[3:, 3:] += bmf * rd @ rd
    U = F.rcross(u)
    Ds, Bs = [], []
    for f, ref in zip(fils, refs):
        d = Q @ ref
        Smat[3:, 3:] += F.rcross(f.tether) @ F.rcross(d)
        Smat[3:, 3:] += 0.5 * F.rcross(f.clamp) + (F.rcross(np.cross(u, f.clamp)) - U @ F.rcross(f.clamp)) / 12.0
        t_base, t_end = F.frame(f.q[0])[0], F.frame(f.q[N - 1])[0]
        Smat[3:, 3:] += F.rcross(-EF * t_end) @ F.rcross(0.5 * dl * t_base)
        D = np.zeros((6, 6 * N))
        D[:3, :3] = np.eye(3)
        D[3:, :3] = F.rcross(-d)
        D[3:, 3 * N:3 * N + 3] = U @ U / 12.0 - 0.5 * U + np.eye(3)
        D[:3, 6 * N - 3:] = F.rcross(-EF * t_end)
        D[3:, 6 * N - 3:] = F.rcross(0.5 * dl * t_end - (f.x[N - 1] - b.x)) @ D[:3, 6 * N - 3:]
        for m in range(1, N - 1):
            D[3:, 3 * (N + m):3 * (N + m) + 3] = D[:3, 6 * N - 3:] @ F.rcross(dl * F.frame(f.q[m])[0])
        B = np.zeros((6 * N, 6))
        Du = F.rcross((0.5 * dl + np.linalg.norm(d)) * d / np.linalg.norm(d))
        for m in range(N):
            B[3 * m:3 * m + 3, :3] = np.eye(3)
            B[3 * m:3 * m + 3, 3:] = Du
        Ds.append(D)
        Bs.append(B)
    return Smat, Ds, Bs, disp


def bfe_clamping(fils, forces, nt, p: F.Params, Bs, num_euler_steps=1):
    """B_n's Lie-algebra rows of segments 0 and 1: the elastic clamping blocks
    (swimmer.cpp:645-646)."""
    for f, Fi, B in zip(fils, forces, Bs):
        c1, c2 = clamping_blocks(f, Fi, nt >= num_euler_steps, p)
        N = p.nseg
        B[3 * N:3 * N + 3, 3:6] = c1
        B[3 * N + 3:3 * N + 6, 3:6] = c2


class BfeInverse:
    """J0^{-1} of one swimmer in the Broyden-for-everything form with a free
    body (swimmer::jacobian_inv_mult, swimmer.cpp:931-1010): the blob rows
    eliminated with the diagonal blob mobility, the body's Schur matrix,
    then the filaments."""

    def __init__(self, J_blocks, Smat, Ds, Bs, disp, nt, p: F.Params, blob_radius=1.0, num_euler_steps=1):
        self.Jinv = [np.linalg.inv(J) for J in J_blocks]
        self.JB = [D @ Ji for D, Ji in zip(Ds, self.Jinv)]
        Sf = Smat.copy()
        for JB, B in zip(self.JB, Bs):
            Sf -= JB @ B
        self.Sinv = np.linalg.inv(Sf)
        self.Bs, self.disp = Bs, np.asarray(disp)
        self.fac = -6.0 * PI * p.mu * blob_radius
        self.dt_fac = 1.0 / p.dt if nt < num_euler_steps else 1.5 / p.dt
        self.n6 = J_blocks[0].shape[0]

    def apply(self, x):
        n6, nf, fac = self.n6, len(self.Jinv), self.fac
        b1 = x[:n6 * nf]
        b3 = x[n6 * nf:n6 * nf + 6].copy()
        b2 = x[n6 * nf + 6:].reshape(-1, 3).copy()
        for n in range(nf):
            b3 -= self.JB[n] @ b1[n6 * n:n6 * (n + 1)]
        for m, d in enumerate(self.disp):
            b3[:3] -= fac * b2[m]
            b3[3:] -= fac * np.cross(d, b2[m])
        out = np.zeros_like(x)
        oB = self.Sinv @ b3
        out[n6 * nf:n6 * nf + 6] = oB
        for m, d in enumerate(self.disp):
            b2[m] -= self.dt_fac * (oB[:3] + np.cross(oB[3:], d))
        out[n6 * nf + 6:] = fac * b2.reshape(-1)
        for n in range(nf):
            out[n6 * n:n6 * (n + 1)] = self.Jinv[n] @ (b1[n6 * n:n6 * (n + 1)] - self.Bs[n] @ oB)
        return out


def free_bfe_loop(b: Body, fils, refs, blob_refs, p: F.Params, steps, tol, perturbation, blob_radius=1.0,
                  max_iter=400, confidence=0.1):
    """The elastic time loop for one free swimmer in the Broyden-for-everything
    form (USE_BROYDEN_FOR_EVERYTHING, PRESCRIBED_BODY_VELOCITIES false): the
    blob forces are Broyden unknowns after the body's six, extrapolated at each
    step's start (rigid_body.cpp:447-478) and updated with the body
    (rigid_body.cpp:506-520); one plain matvec per evaluation; the residual's
    body rows are the force and torque balance and its blob rows the rigid
    motion the body's positions imply; J0^{-1} through bfe_coupling /
    BfeInverse.  Returns per step the Broyden iterations and the state."""
    from oracle import broyden_oracle as BO
    from oracle import oracle as O

    N, nfil = p.nseg, len(fils)
    nb3 = np.asarray(blob_refs).size
    n6 = 6 * N
    n = n6 * nfil + 6 + nb3
    cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
    fb, fbm1, fbm2 = np.zeros(nb3), np.zeros(nb3), np.zeros(nb3)
    kw = dict(seg_dof=6, a_blob=blob_radius)
    out = []
    for nt in range(steps):
        swimmer_initial_guess(b, fils, refs, nt, p)
        fb, fbm1, fbm2 = blob_force_guess(fb, fbm1, fbm2, nt)

        def evaluate(fb):
            kicks = [{4: perturbation[i, :3], 12: perturbation[i, 3:]} if nt == 0 else None for i in range(nfil)]
            forces = [F.internal_forces(fl, p, kicks[i]) for i, fl in enumerate(fils)]
            x, f = cat("x"), np.concatenate([fc.reshape(-1) for fc in forces])
            xb = blob_positions(b, blob_refs)
            vb = O.block("bb", x, xb, f, fb, **kw)
            vb = O.block("bs", x, xb, f, fb, vb, bs_accumulate=True, **kw)
            vs = O.block("sb", x, xb, f, fb, O.block("ss", x, xb, f, fb, **kw), **kw)
            st = {"seg_x": x, "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"),
                  "seg_um1": cat("um1"), "v_segs": vs, "body_x": b.x, "body_xm1": b.xm1, "body_xm2": b.xm2,
                  "body_u": b.u, "body_um1": b.um1, "v_bodies": np.zeros(6), "x_blobs": xb, "v_blobs": vb,
                  "f_segs": f, "f_blobs": fb}
            e, big = BO.compute_errors(st, 1, nfil, N, nt >= 1, p.dt, p.dl, tol, blobs_per_swim=nb3 // 3,
                                       free_bodies=True)
            return forces, e, big

        forces, e, big = evaluate(fb)
        Js = [F.approx_jacobian(fl, forces[i], nt >= 1, p) for i, fl in enumerate(fils)]
        Smat, Ds, Bs, disp = bfe_coupling(b, fils, refs, blob_refs, fb, nt, p, blob_radius)
        bfe_clamping(fils, forces, nt, p, Bs)
        jinv = BfeInverse(Js, Smat, Ds, Bs, disp, nt, p, blob_radius).apply
        ref = BroydenOp(n, max_iter, confidence, jinv)
        ref.start(e)
        while big and ref.iter < max_iter:
            du = ref.find_update()
            swimmer_update(b, fils, refs, du[:n6 * nfil + 6], p)
            fb = fb + du[n6 * nfil + 6:]
            forces, e, big = evaluate(fb)
            ref.end_of_iter(e, big)
        out.append({"iters": ref.iter, "converged": not big, "x": cat("x"), "body": b.copy(), "f_blobs": fb.copy(),
                    "error": e})
    return out


def blob_force_guess(fb, fbm1, fbm2, nt):
    """The blob forces' extrapolation of rigid_body::initial_guess in the
    Broyden-for-everything form (rigid_body.cpp:447-478)."""
    if nt == 1:
        fbm1 = fb.copy()
    elif nt == 2:
        fbm2, fbm1 = fbm1.copy(), fb.copy()
        fb = 2.0 * fbm1 - fbm2
    elif nt > 2:
        t = 3.0 * (fb - fbm1) + fbm2
        fbm2, fbm1, fb = fbm1.copy(), fb.copy(), t
    return fb, fbm1, fbm2


def held_still_loop(fils, x_blobs, p: F.Params, steps, tol, perturbation, broyden_blobs, blob_radius=1.0,
                    max_iter=400, confidence=0.1):
    """The elastic time loop (cilia_sim_main.cu:353-502) for one body held
    still, in either solver form: broyden_blobs=False solves the blob forces
    each evaluation (SOLVER_TYPE 1; a dense solve here), True makes them
    Broyden unknowns (USE_BROYDEN_FOR_EVERYTHING): the velocities are then one
    plain matvec (mobility_solver.cpp:2603-2640), the residual has the blob
    rows, the blob forces are extrapolated at each step's start and updated by
    the Broyden update (rigid_body.cpp:506-520), and J0^{-1} scales the blob
    rows by -6 pi mu a (swimmer.cpp:931-945)."""
    import scipy.linalg as sla

    from oracle import broyden_oracle as BO
    from oracle import oracle as O

    N, nfil = p.nseg, len(fils)
    xb = np.asarray(x_blobs, dtype=float).reshape(-1)
    nb3 = xb.size
    n = 6 * (nfil * N + 1) + (nb3 if broyden_blobs else 0)
    cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
    lu = None if broyden_blobs else sla.lu_factor(O.dense_mobility(np.zeros(0), xb, a_blob=blob_radius))
    fb, fbm1, fbm2 = np.zeros(nb3), np.zeros(nb3), np.zeros(nb3)
    z3, z6 = np.zeros(3), np.zeros(6)
    out = []
    for nt in range(steps):
        for fl in fils:
            F.initial_guess(fl, nt, p)
        if broyden_blobs:
            fb, fbm1, fbm2 = blob_force_guess(fb, fbm1, fbm2, nt)

        def evaluate(fb):
            kicks = [{4: perturbation[i, :3], 12: perturbation[i, 3:]} if nt == 0 else None for i in range(nfil)]
            f = np.concatenate([F.internal_forces(fl, p, kicks[i]).reshape(-1) for i, fl in enumerate(fils)])
            x = cat("x")
            kw = dict(seg_dof=6, a_blob=blob_radius)
            if broyden_blobs:
                vb = O.block("bb", x, xb, f, fb, **kw)
                vb = O.block("bs", x, xb, f, fb, vb, bs_accumulate=True, **kw)
            else:
                fb = sla.lu_solve(lu, -O.block("bs", x, xb, f, np.zeros(nb3), bs_accumulate=False, **kw))
                vb = None
            vs = O.block("sb", x, xb, f, fb, O.block("ss", x, xb, f, fb, **kw), **kw)
            st = {"seg_x": x, "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"),
                  "seg_um1": cat("um1"), "v_segs": vs, "body_x": z3, "body_xm1": z3, "body_xm2": z3,
                  "body_u": z3, "body_um1": z3, "v_bodies": z6, "x_blobs": xb, "v_blobs": vb}
            e, big = BO.compute_errors(st, 1, nfil, N, nt >= 1, p.dt, p.dl, tol,
                                       blobs_per_swim=nb3 // 3 if broyden_blobs else 0)
            return f, e, big, fb

        f, e, big, fb_eval = evaluate(fb)
        blocks = [np.linalg.inv(F.approx_jacobian(fl, f.reshape(nfil, N, 6)[i], nt >= 1, p))
                  for i, fl in enumerate(fils)] + [np.eye(6)]
        if broyden_blobs:
            blocks += [-6.0 * PI * p.mu * blob_radius * np.eye(3)] * (nb3 // 3)
        ref = BO.BroydenRef(n, max_iter, confidence, blocks)
        ref.start(e)
        while big and ref.iter < max_iter:
            du = ref.find_update()
            for i, fl in enumerate(fils):
                F.update(fl, du[6 * N * i:6 * N * (i + 1)], p)
            if broyden_blobs:
                fb = fb + du[6 * (nfil * N + 1):]
            f, e, big, fb_eval = evaluate(fb)
            ref.end_of_iter(e, big)
        out.append({"iters": ref.iter, "converged": not big, "x": cat("x"), "f_blobs": fb_eval.copy()})
    return out
