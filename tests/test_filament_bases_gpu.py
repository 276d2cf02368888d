"""The base segment's other builds on the GPU (tfs_filament_set_base) against the
oracle (oracle/filament_oracle.py, pinned on the CPU in
tests/test_filament_oracle.py):

* GEOMETRIC_CILIA (CILIA_TYPE 1): a free base segment whose Lie-algebra
  element is the unknown in the clamp's slot, driven by a torque on segment 0
  (filament.cpp:1523-1531, 1789-1800, 1918-1927);
* CONSTANT_BASE_ROTATION (CILIA_TYPE 2): the clamped base turned by a fixed
  element every step (filament.cpp:487-492, 1565-1571).

State, internal forces and J0 blocks over several start-of-step / update
cycles <= 1e-12 relative; one and two elastic implicit steps through
tfs_elastic_step (4 filaments x 12 segments on 300 blobs) against the same
loop on the oracle: the same Broyden iteration counts (within 10%) and
positions within 1e-6."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch
from scipy.spatial.transform import Rotation

from oracle import broyden_oracle as BO
from oracle import filament_oracle as F
from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import filament as FM
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.broyden import BroydenSolver
from timfilsimmod_amd.solver import SOLVE_FP64, ElasticVelocities

pytestmark = pytest.mark.gpu

NFIL, NSEG = 4, 8
PARAMS = dict(dl=2.2, kb=1800.0, kt=1300.0, mu=1.0, radius=1.0, dt=0.2, end_force=0.0)


def oracle_params(base, nseg=NSEG):
    return F.Params(nseg=nseg, dl=PARAMS["dl"], kb=PARAMS["kb"], kt=PARAMS["kt"], mu=PARAMS["mu"],
                    a=PARAMS["radius"], dt=PARAMS["dt"], end_force=0.0, base=base)


def setup(seed, nfil=NFIL, nseg=NSEG):
    rng = np.random.default_rng(seed)
    bases = rng.standard_normal((nfil, 3)) * 10
    qs = []
    for f in range(nfil):
        q0 = np.roll(Rotation.random(random_state=seed + f).as_quat(), 1)
        row = [q0]
        for _ in range(nseg - 1):
            row.append(F.qmul(F.lie_exp(0.15 * rng.standard_normal(3)), row[-1]))
        qs.append(row)
    return bases, np.array(qs), rng


def base_values(mode, qs, rng):
    """Per filament: the turn per step about the base tangent (clamped) or a
    driving torque perpendicular to it (driven)."""
    t0 = np.array([F.frame(q[0])[0] for q in qs])
    if mode == "rotating":
        return 0.04 * t0
    side = np.cross(t0, rng.standard_normal((len(qs), 3)))
    return 60.0 * side / np.linalg.norm(side, axis=1, keepdims=True)


def oracle_fils(bases, qs, p, mode, vals):
    fils = [F.Filament.at_rest(bases[f], qs[f], NSEG, p.dl) for f in range(len(bases))]
    for f, fil in enumerate(fils):
        if mode == "rotating":
            fil.u0 = vals[f].copy()
        else:
            fil.drive = vals[f].copy()
    return fils


def compare_state(model, fils):
    cat = lambda name: np.concatenate([getattr(f, name).reshape(-1) for f in fils])
    mult = np.concatenate([np.concatenate([f.tether, f.lam.reshape(-1), f.clamp]) for f in fils])
    for which, want in [(FM.X, cat("x")), (FM.XM1, cat("xm1")), (FM.U, cat("u")), (FM.UM1, cat("um1")),
                        (FM.Q, cat("q")), (FM.MULT, mult)]:
        got = model.read(which)
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), which


@pytest.mark.parametrize("mode", ["driven", "rotating"])
def test_base_builds_match_oracle(mode):
    print(f"case filament base {mode}", flush=True)
    base = "driven" if mode == "driven" else "clamped"
    p = oracle_params(base)
    bases, qs, rng = setup(31)
    vals = base_values(mode, qs, rng)
    model = FM.FilamentModel(1, NFIL, NSEG, **PARAMS)
    n6 = 6 * NSEG
    try:
        model.set_state(bases.reshape(-1), qs.reshape(-1))
        model.set_base(FM.BASE_DRIVEN if mode == "driven" else FM.BASE_CLAMPED, vals.reshape(-1))
        fils = oracle_fils(bases, qs, p, mode, vals)
        for nt in range(4):
            model.initial_guess(nt)
            for fil in fils:
                F.initial_guess(fil, nt, p)
            compare_state(model, fils)
            du = np.zeros(NFIL * n6 + 6)
            for f in range(NFIL):
                o = f * n6
                du[o:o + 3 * NSEG] = rng.standard_normal(3 * NSEG) * 20.0                   # multipliers
                du[o + 3 * NSEG:o + 3 * NSEG + 3] = rng.standard_normal(3) * (0.05 if base == "driven" else 20.0)
                du[o + 3 * NSEG + 3:o + n6] = rng.standard_normal(3 * NSEG - 3) * 0.05      # u_1..
            model.update(du)
            for f, fil in enumerate(fils):
                F.update(fil, du[f * n6:(f + 1) * n6], p)
            compare_state(model, fils)
            fs = model.forces(nt)
            want_f = np.concatenate([F.internal_forces(fil, p).reshape(-1) for fil in fils])
            assert np.linalg.norm(fs - want_f) <= 1e-13 * np.linalg.norm(want_f)
            blocks = np.full(NFIL * n6 * n6 + 36, 7.0)
            model.jacobian(nt, fs, blocks)
            for f, fil in enumerate(fils):
                J = F.approx_jacobian(fil, want_f.reshape(NFIL, NSEG, 6)[f], nt >= 1, p)
                Jg = blocks[f * n6 * n6:(f + 1) * n6 * n6].reshape(n6, n6)
                assert np.abs(Jg - J).max() <= 1e-12 * np.abs(J).max(), (nt, f)
    finally:
        model.close()


def test_driven_base_is_rejected_with_a_free_body():
    model = FM.FilamentModel(1, 2, 4, **PARAMS)
    try:
        with pytest.raises(L.TfsError):
            model.set_base(7)
        model.set_body(np.zeros(3), np.array([1.0, 0, 0, 0]), np.zeros(6))
        with pytest.raises(L.TfsError, match="held still"):
            model.set_base(FM.BASE_DRIVEN, np.zeros(6))
    finally:
        model.close()


STEPS = 2
ENFIL, ENSEG, ENB = 4, 12, 300


def elastic_case(mode):
    body = G.make_body(ENFIL, ENSEG, ENB, 9.0)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    qs = np.repeat(q0[:, None, :], ENSEG, axis=1)
    if mode == "rotating":
        vals = 0.2 * body.fil_normal                      # BASE_ROTATION_RATE DT about each base tangent
    else:
        fast = np.cross(body.fil_normal, [0.0, 0.0, 1.0])  # my_vertical x fast_beating_direction, scaled
        vals = -150.0 * fast / np.linalg.norm(fast, axis=1, keepdims=True)
    return body, qs, vals


def oracle_loop(mode, dt, tol):
    body, qs, vals = elastic_case(mode)
    base = "driven" if mode == "driven" else "clamped"
    p = F.Params(nseg=ENSEG, dl=2.2, kb=1800.0, kt=1800.0, dt=dt, base=base)
    fils = [F.Filament.at_rest(body.fil_base[i], qs[i], ENSEG, p.dl) for i in range(ENFIL)]
    for i, fl in enumerate(fils):
        if mode == "rotating":
            fl.u0 = vals[i].copy()
        else:
            fl.drive = vals[i].copy()
    lu = sla.lu_factor(O.dense_mobility(np.zeros(0), body.x_blobs))

    def vel(x, f):
        vb = O.block("bs", x, body.x_blobs, f, np.zeros(3 * ENB), seg_dof=6, bs_accumulate=False)
        lam = sla.lu_solve(lu, -vb)
        vs = O.block("ss", x, body.x_blobs, f, lam, seg_dof=6)
        return O.block("sb", x, body.x_blobs, f, lam, vs, seg_dof=6)

    cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
    z3, z6 = np.zeros(3), np.zeros(6)
    iters, xs = [], []
    x_start = cat("x")
    for nt in range(STEPS):
        for fl in fils:
            F.initial_guess(fl, nt, p)

        def evaluate():
            f = np.concatenate([F.internal_forces(fl, p).reshape(-1) for fl in fils])
            v = vel(cat("x"), f)
            st = {"seg_x": cat("x"), "seg_xm1": cat("xm1"), "seg_xm2": cat("xm2"), "seg_u": cat("u"),
                  "seg_um1": cat("um1"), "v_segs": v, "body_x": z3, "body_xm1": z3, "body_xm2": z3,
                  "body_u": z3, "body_um1": z3, "v_bodies": z6}
            e, big = BO.compute_errors(st, 1, ENFIL, ENSEG, nt >= 1, dt, 2.2, tol)
            return f, e, big

        f, e, big = evaluate()
        inv = [np.linalg.inv(F.approx_jacobian(fl, f.reshape(ENFIL, ENSEG, 6)[i], nt >= 1, p))
               for i, fl in enumerate(fils)] + [np.eye(6)]
        ref = BO.BroydenRef(6 * (ENFIL * ENSEG + 1), 400, 0.1, inv)
        ref.start(e)
        while big and ref.iter < 400:
            du = ref.find_update()
            for i, fl in enumerate(fils):
                F.update(fl, du[6 * ENSEG * i:6 * ENSEG * (i + 1)], p)
            f, e, big = evaluate()
            ref.end_of_iter(e, big)
        iters.append(ref.iter)
        xs.append(cat("x"))
    return iters, xs, x_start, fils


@pytest.mark.parametrize("mode", ["driven", "rotating"])
def test_elastic_step_with_base_build_matches_oracle_loop(mode):
    print(f"case elastic step base {mode}, {ENFIL}x{ENSEG} on {ENB} blobs", flush=True)
    dt, tol = 0.1, 1e-8
    ref_iters, ref_x, x_start, ref_fils = oracle_loop(mode, dt, tol)
    body, qs, vals = elastic_case(mode)
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    model = FM.FilamentModel(1, ENFIL, ENSEG, dl=2.2, kb=1800.0, kt=1800.0, dt=dt, memory=L.MEMORY_DEVICE)
    ev = ElasticVelocities(1, ENFIL * ENSEG, ENB, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                           memory=L.MEMORY_DEVICE)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
    keep = [dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * ENB))]
    step = None
    iters, xs = [], []
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(qs.reshape(-1)))
        model.set_base(FM.BASE_DRIVEN if mode == "driven" else FM.BASE_CLAMPED, dev(vals.reshape(-1)))
        step = FM.ElasticStep(model, ev, eng, *keep, tol=tol, max_iter=400)
        for nt in range(STEPS):
            st = step.step(nt)
            assert st["converged"] == 1
            iters.append(st["broyden_iterations"])
            xs.append(model.read(FM.X))
        u = model.read(FM.U).reshape(ENFIL, ENSEG, 3)
    finally:
        if step is not None:
            step.close()
        model.close()
        ev.close()
        eng.close()
    print("broyden iterations", iters, "oracle", ref_iters,
          "max |dx|", [float(np.abs(a - b).max()) for a, b in zip(xs, ref_x)], flush=True)
    for s in range(STEPS):
        assert ref_iters[s] >= 3 and abs(iters[s] - ref_iters[s]) <= max(2, ref_iters[s] // 10)
        assert np.abs(xs[s] - ref_x[s]).max() <= 1e-6
    # the base does what the build says: it turned by the fixed element, or the
    # driving torque moved the filaments
    if mode == "rotating":
        assert np.allclose(u[:, 0], vals, atol=1e-14)
    else:
        assert np.abs(xs[-1] - x_start).max() > 1e-3
        assert np.allclose(u[:, 0], np.array([fl.u[0] for fl in ref_fils]), atol=1e-8)
