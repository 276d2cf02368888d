"""Pins the oracle's free swimmer in the Broyden-for-everything form
(oracle/swimmer_oracle.py: bfe_coupling, BfeInverse, free_bfe_loop, restating
swimmer.cpp:565-676, 931-1010; oracle/broyden_oracle.py: compute_errors with
free bodies, mobility_solver.cpp:2674-2684, 2774-2850) on the CPU.

The residual's unknowns are the filament blocks, the body's position and
Lie-algebra element and the blob forces; its rows the filament equations, the
body's force and torque balance and the blobs' rigid motion.  J0 is the block
matrix [[diag J_n, B, 0], [D, S_body, E], [0, F, G]] whose inverse
jacobian_inv_mult applies by eliminating the blobs with their diagonal
mobility (G = -I / (6 pi mu a)) and the body through its Schur matrix.  Against
central differences of that residual with every segment and blob moving with
its self-mobility, at states where the blocks are exact derivatives:

* D (force and torque balance against the filament unknowns), E, G and, in
  backward-Euler steps, F are exact;
* in BDF2 steps F leaves out the terms of the body's angular velocity that
  the reference's own comment names (swimmer.cpp:581-582): checked as that
  closed-form difference;
* S_body: the end force's lever arm about the body's centre is taken as
  0.5 DL t_base (swimmer.cpp:620-627), leaving out the anchor's displacement:
  S_body differs by rcross(-EF t_end) rcross(d_anchor), checked as such;
* B is the GMRES form's A (pinned in test_oracle_body_coupling.py), with its
  segment-0 lever and clamp-column simplifications.

Planted sign flips fail these checks.  The loop reaches the trajectory of the
GMRES form of the same free swimmer (tests/test_oracle_swimmer.py)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import broyden_oracle as BO
from oracle import filament_oracle as F
from oracle import oracle as O
from oracle import swimmer_oracle as S
import test_oracle_body_coupling as BC

PI = np.pi


def state(seed, nb=6, end_force=30.0):
    p, b, fils, refs, _ = BC.swimmer(seed, end_force=end_force)
    rng = np.random.default_rng(seed + 100)
    blob_refs = 9.0 * rng.standard_normal((nb, 3))
    fb = rng.standard_normal(3 * nb) * 5
    return p, b, fils, refs, blob_refs, fb


def residual(b0, fils0, refs, blob_refs, fb0, p, bdf2, dz, a_b=1.0):
    """The swimmer's residual after the Broyden update dz, with self-mobility
    segments and blobs."""
    b, fils = b0.copy(), [f.copy() for f in fils0]
    n6, nf = 6 * p.nseg, len(fils)
    S.swimmer_update(b, fils, refs, dz[:n6 * nf + 6], p)
    fb = fb0 + dz[n6 * nf + 6:]
    forces = [F.internal_forces(fl, p) for fl in fils]
    e = [F.self_mobility_residual(fl, f0.xm1, f0.xm2, f0.um1, bdf2, p) for fl, f0 in zip(fils, fils0)]
    xb = S.blob_positions(b, blob_refs)
    st = {"seg_x": np.concatenate([f.x.reshape(-1) for f in fils]), "v_segs": np.zeros(0),
          "body_x": b.x, "body_xm1": b0.xm1, "body_xm2": b0.xm2, "body_u": b.u, "body_um1": b0.um1,
          "v_bodies": np.zeros(6), "x_blobs": xb, "v_blobs": fb / (6 * PI * p.mu * a_b),
          "f_segs": np.concatenate([fc.reshape(-1) for fc in forces]), "f_blobs": fb}
    # the body and blob rows as compute_errors forms them (no filament rows: num_fil 0)
    per_seg = {k: st[k] for k in st}
    eb, _ = BO.compute_errors(dict(per_seg, seg_x=np.zeros(0)), 1, 0, p.nseg, bdf2, p.dt, p.dl, 1.0,
                              blobs_per_swim=len(blob_refs), free_bodies=True)
    # compute_errors' body rows with num_fil 0 miss the segments: add them
    for fl, Fi in zip(fils, forces):
        eb[:3] += Fi[:, :3].sum(0)
        eb[3:6] += (Fi[:, 3:] + np.cross(fl.x - b.x, Fi[:, :3])).sum(0)
    return np.concatenate(e + [eb])


def central_differences(b, fils, refs, blob_refs, fb, p, bdf2, h=1e-6):
    n = 6 * p.nseg * len(fils) + 6 + fb.size
    cols = []
    for c in range(n):
        dz = np.zeros(n)
        dz[c] = h
        cols.append((residual(b, fils, refs, blob_refs, fb, p, bdf2, dz) -
                     residual(b, fils, refs, blob_refs, fb, p, bdf2, -dz)) / (2 * h))
    return np.array(cols).T


def blocks(nt, seed=5, **kw):
    p, b, fils, refs, blob_refs, fb = state(seed + nt, **kw)
    bdf2 = nt >= 1
    forces = [F.internal_forces(fl, p) for fl in fils]
    Smat, Ds, Bs, disp = S.bfe_coupling(b, fils, refs, blob_refs, fb, nt, p)
    S.bfe_clamping(fils, forces, nt, p, Bs)
    nb = len(blob_refs)
    fac = -6.0 * PI * p.mu
    dt_fac = 1.0 / p.dt if nt < 1 else 1.5 / p.dt
    E = np.zeros((6, 3 * nb))
    Fm = np.zeros((3 * nb, 6))
    for m, d in enumerate(disp):
        E[:3, 3 * m:3 * m + 3] = np.eye(3)
        E[3:, 3 * m:3 * m + 3] = -F.rcross(d)
        Fm[3 * m:3 * m + 3, :3] = dt_fac * np.eye(3)
        Fm[3 * m:3 * m + 3, 3:] = dt_fac * F.rcross(d)
    G = np.eye(3 * nb) / fac
    Sbody = Smat + fac * E @ Fm          # the Schur matrix's blob terms are -fac E F
    Js = [F.approx_jacobian(fl, forces[i], bdf2, p) for i, fl in enumerate(fils)]
    return dict(p=p, b=b, fils=fils, refs=refs, blob_refs=blob_refs, fb=fb, forces=forces, Smat=Smat, Ds=Ds,
                Bs=Bs, disp=disp, E=E, F=Fm, G=G, Sbody=Sbody, Js=Js, nt=nt, bdf2=bdf2)


def assembled(c):
    n6, nf, nb3 = 6 * c["p"].nseg, len(c["fils"]), c["fb"].size
    o = n6 * nf
    K = np.zeros((o + 6 + nb3, o + 6 + nb3))
    for i in range(nf):
        s = slice(n6 * i, n6 * (i + 1))
        K[s, s] = c["Js"][i]
        K[s, o:o + 6] = c["Bs"][i]
        K[o:o + 6, s] = c["Ds"][i]
    K[o:o + 6, o:o + 6] = c["Sbody"]
    K[o:o + 6, o + 6:] = c["E"]
    K[o + 6:, o:o + 6] = c["F"]
    K[o + 6:, o + 6:] = c["G"]
    return K


@pytest.mark.parametrize("nt", [0, 3])
def test_bfe_inverse_is_the_inverse_of_the_block_matrix(nt):
    c = blocks(nt)
    K = assembled(c)
    inv = S.BfeInverse(c["Js"], c["Smat"], c["Ds"], c["Bs"], c["disp"], nt, c["p"])
    x = np.random.default_rng(3).standard_normal(K.shape[0])
    assert np.abs(K @ inv.apply(x) - x).max() <= 1e-9 * np.abs(x).max()


def f_deviation(c):
    """The BDF2 terms of the body's angular velocity left out of F: at u = 0,
    Omega = w = (um1 / 3) / (-2 dt / 3) and d Omega / du = dt_fac I + w x / 2."""
    b, p = c["b"], c["p"]
    dev = np.zeros_like(c["F"])
    if not c["bdf2"]:
        return dev
    w = (b.um1 / 3.0 - b.u) / (-2.0 / 3.0 * p.dt)
    for m, d in enumerate(c["disp"]):
        for k in range(3):
            e = np.eye(3)[k]
            dev[3 * m:3 * m + 3, 3 + k] = np.cross(0.5 * np.cross(e, w), d) + np.cross(w, np.cross(e, d))
    return dev


def s_deviation(c):
    dev = np.zeros((6, 6))
    for f, ref in zip(c["fils"], c["refs"]):
        tN = F.frame(f.q[-1])[0]
        dev[3:, 3:] += F.rcross(-c["p"].end_force * tN) @ F.rcross(c["b"].Q_init @ ref)
    return dev


def b_deviation(c, i):
    p, f = c["p"], c["fils"][i]
    N = p.nseg
    d = c["b"].Q_init @ c["refs"][i]
    dev = np.zeros((6 * N, 6))
    dev[0:3, 3:6] = F.rcross(0.5 * p.dl * d / np.linalg.norm(d))
    tsf = 2.0 * p.dt / 3.0 if c["bdf2"] else p.dt
    dev[3 * N:3 * N + 3, 3:6] = tsf / (8 * PI * p.mu * p.a ** 3) * 0.5 * F.rcross(f.clamp)
    return dev


@pytest.mark.parametrize("nt", [0, 3])
def test_bfe_coupling_is_the_jacobian_of_the_swimmer_residual(nt):
    c = blocks(nt)
    J = central_differences(c["b"], c["fils"], c["refs"], c["blob_refs"], c["fb"], c["p"], c["bdf2"])
    n6, nf = 6 * c["p"].nseg, len(c["fils"])
    o = n6 * nf
    tol = lambda A: 1e-7 * max(np.abs(A).max(), 1.0)
    for i in range(nf):
        assert np.abs(c["Ds"][i] - J[o:o + 6, n6 * i:n6 * (i + 1)]).max() <= tol(c["Ds"][i])
        B_fd = J[n6 * i:n6 * (i + 1), o:o + 6]
        assert np.abs(c["Bs"][i] - B_fd - b_deviation(c, i)).max() <= tol(c["Bs"][i])
        assert np.abs(J[n6 * i:n6 * (i + 1), o + 6:]).max() == 0.0      # filaments do not see blob forces
    assert np.abs(c["E"] - J[o:o + 6, o + 6:]).max() <= tol(c["E"])
    assert np.abs(c["G"] - J[o + 6:, o + 6:]).max() <= tol(c["G"])
    assert np.abs(J[o + 6:, :o]).max() == 0.0
    assert np.abs(c["F"] + f_deviation(c) - J[o + 6:, o:o + 6]).max() <= tol(c["F"])
    assert np.abs(c["Sbody"] + s_deviation(c) - J[o:o + 6, o:o + 6]).max() <= tol(c["Sbody"])
    # the left-out terms are real here
    assert np.abs(s_deviation(c)).max() > 0.1 * np.abs(c["Sbody"]).max()
    if c["bdf2"]:
        assert np.abs(f_deviation(c)).max() > 1e-4 * np.abs(c["F"]).max()
    # planted sign flips are seen
    for name, blk, ref in (("D", c["Ds"][0][3:, :3], J[o + 3:o + 6, :3]),
                           ("S", c["Sbody"][3:, 3:] + s_deviation(c)[3:, 3:], J[o + 3:o + 6, o + 3:o + 6]),
                           ("F", c["F"][:, 3:] + f_deviation(c)[:, 3:], J[o + 6:, o + 3:o + 6])):
        assert np.abs(-blk - ref).max() > 1e-3 * np.abs(ref).max(), name


def test_without_end_force_the_body_block_is_exact():
    c = blocks(0, end_force=0.0)
    J = central_differences(c["b"], c["fils"], c["refs"], c["blob_refs"], c["fb"], c["p"], c["bdf2"])
    o = 6 * c["p"].nseg * len(c["fils"])
    assert np.abs(c["Sbody"] - J[o:o + 6, o:o + 6]).max() <= 1e-7 * np.abs(c["Sbody"]).max()


def test_free_bfe_loop_reaches_the_gmres_form_trajectory():
    """The two forms solve the same equations for a free swimmer (blob forces
    from Broyden against the blob-only system with body rows each
    evaluation): after two steps the positions and the body agree to the
    tolerance."""
    import test_oracle_swimmer as TS
    c = TS.free_case(nfil=2, nseg=20, nb=200, R=8.0)
    b1, f1 = c["fresh"]()
    gm = S.free_swimmer_loop(b1, f1, c["refs"], c["body"].blob_refs, c["p"], 2, c["tol"], c["pert"], c["N_ref"],
                             sphere_radius=c["R"])
    b2, f2 = c["fresh"]()
    bf = S.free_bfe_loop(b2, f2, c["refs"], c["body"].blob_refs, c["p"], 2, c["tol"], c["pert"])
    print("Broyden iterations, GMRES form", [r["iters"] for r in gm], "Broyden for everything",
          [r["iters"] for r in bf])
    assert all(r["converged"] for r in gm + bf)
    for a, bb in zip(gm, bf):
        assert np.abs(a["x"] - bb["x"]).max() <= 1e-6
        assert np.abs(a["body"].x - bb["body"].x).max() <= 1e-7
        # blob forces to what their rows' velocity tolerance fixes through the blob mobility
        assert np.abs(a["f_blobs"] - bb["f_blobs"]).max() <= 1e-3 * np.abs(a["f_blobs"]).max()
    # the body swam
    assert np.abs(bf[-1]["body"].x - c["body"].centre).max() > 1e-6
