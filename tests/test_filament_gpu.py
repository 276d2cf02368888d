"""The Cosserat filament model on the GPU (include/tfs_filament.h) against the
oracle's restatement (oracle/filament_oracle.py, pinned on the CPU):

* state, internal forces and approximate-Jacobian blocks of random bent
  filaments with random multipliers, in backward-Euler and BDF2 steps, with
  the first-step kick, host and device memory (<= 1e-12 relative);
* the start-of-step extrapolation and update over several steps;
* one elastic implicit timestep at the small-body size with the real model
  (16 filaments x 20 segments, 2,000 blobs; follower end force of
  DIMENSIONLESS_FORCE 220): the device loop -- filament kernels, device J0
  inversion, tfs_velocities_compute, compute_errors and the Broyden engine --
  against the same loop on the oracle."""
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
from timfilsimmod_amd.solver import BLOB_DIRECT, PRECOND_BLOCKS, PRECOND_REFERENCE, SOLVE_FP64, ElasticVelocities

pytestmark = pytest.mark.gpu

NSWIM, NFIL, NSEG = 2, 3, 8
PARAMS = dict(dl=2.2, kb=1800.0, kt=1300.0, mu=1.0, radius=1.0, dt=0.2, end_force=150.0)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def oracle_params(nseg=NSEG, **kw):
    d = dict(PARAMS, **kw)
    return F.Params(nseg=nseg, dl=d["dl"], kb=d["kb"], kt=d["kt"], mu=d["mu"], a=d["radius"], dt=d["dt"],
                    end_force=d["end_force"])


def random_setup(seed):
    rng = np.random.default_rng(seed)
    nf = NSWIM * NFIL
    bases = rng.standard_normal((nf, 3)) * 10
    qs = []
    for f in range(nf):
        q0 = np.roll(Rotation.random(random_state=seed + f).as_quat(), 1)
        row = [q0]
        for _ in range(NSEG - 1):
            row.append(F.qmul(F.lie_exp(0.2 * rng.standard_normal(3)), row[-1]))
        qs.append(row)
    return bases, np.array(qs)


def random_update(rng):
    du = np.zeros(NSWIM * (NFIL * 6 * NSEG + 6))
    for s in range(NSWIM):
        for f in range(NFIL):
            o = s * (NFIL * 6 * NSEG + 6) + f * 6 * NSEG
            du[o:o + 3 * NSEG + 3] = rng.standard_normal(3 * NSEG + 3) * 20.0      # multipliers
            du[o + 3 * NSEG + 3:o + 6 * NSEG] = rng.standard_normal(3 * NSEG - 3) * 0.05   # u_1..
    return du


def oracle_fils(bases, qs, p):
    return [F.Filament.at_rest(bases[f], qs[f], NSEG, p.dl) for f in range(len(bases))]


def oracle_update(fils, du, p):
    for f, fil in enumerate(fils):
        s, k = divmod(f, NFIL)
        o = s * (NFIL * 6 * NSEG + 6) + k * 6 * NSEG
        F.update(fil, du[o:o + 6 * NSEG], p)


def compare_state(model, fils):
    cat = lambda name: np.concatenate([getattr(f, name).reshape(-1) for f in fils])
    mult = np.concatenate([np.concatenate([f.tether, f.lam.reshape(-1), f.clamp]) for f in fils])
    for which, want in [(FM.X, cat("x")), (FM.XM1, cat("xm1")), (FM.XM2, cat("xm2")), (FM.U, cat("u")),
                        (FM.UM1, cat("um1")), (FM.Q, cat("q")), (FM.MULT, mult)]:
        got = model.read(which)
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), which


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_forces_and_jacobian_match_oracle(memory):
    print(f"case filament kernels memory={memory}", flush=True)
    p = oracle_params()
    bases, qs = random_setup(11)
    rng = np.random.default_rng(12)
    du = random_update(rng)
    pert = rng.standard_normal((NSWIM * NFIL, 6)) * 30
    model = FM.FilamentModel(NSWIM, NFIL, NSEG, memory=memory, **PARAMS)
    dev = (lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")) if memory == L.MEMORY_DEVICE else \
        (lambda a: np.ascontiguousarray(a, dtype=np.float64))
    def host(a):
        if memory != L.MEMORY_DEVICE:
            return a
        model.sync()          # the model's stream, then torch's copy
        return a.cpu().numpy()
    try:
        model.set_state(dev(bases.reshape(-1)), dev(qs.reshape(-1)))
        model.set_perturbation(4, 6, dev(pert.reshape(-1)))
        model.update(dev(du))
        fils = oracle_fils(bases, qs, p)
        oracle_update(fils, du, p)
        compare_state(model, fils)
        n6 = 6 * NSEG
        nblk = NSWIM * (NFIL * n6 * n6 + 36)
        for nt in (0, 3):
            fs = dev(np.zeros(6 * NSWIM * NFIL * NSEG))
            model.forces(nt, fs)
            kick = lambda f: {4: pert[f, :3], 6: pert[f, 3:]} if nt == 0 else None
            want_f = np.concatenate([F.internal_forces(fil, p, kick(f)).reshape(-1) for f, fil in enumerate(fils)])
            got_f = host(fs)
            assert rel(got_f, want_f) <= 1e-13
            blocks = dev(np.full(nblk, 7.0))
            model.jacobian(nt, fs, blocks)
            got = host(blocks)
            for s in range(NSWIM):
                o = s * (NFIL * n6 * n6 + 36)
                assert np.all(got[o + NFIL * n6 * n6:o + NFIL * n6 * n6 + 36] == 7.0)   # body block untouched
                for k in range(NFIL):
                    f = s * NFIL + k
                    J = F.approx_jacobian(fils[f], want_f.reshape(-1, NSEG, 6)[f], nt >= 1, p)
                    Jg = got[o + k * n6 * n6:o + (k + 1) * n6 * n6].reshape(n6, n6)
                    assert np.abs(Jg - J).max() <= 1e-12 * np.abs(J).max(), (nt, f)
    finally:
        model.close()


def test_initial_guess_sequence_matches_oracle():
    p = oracle_params()
    bases, qs = random_setup(21)
    rng = np.random.default_rng(22)
    model = FM.FilamentModel(NSWIM, NFIL, NSEG, **PARAMS)
    try:
        model.set_state(bases.reshape(-1), qs.reshape(-1))
        fils = oracle_fils(bases, qs, p)
        for nt in range(5):
            model.initial_guess(nt)
            for fil in fils:
                F.initial_guess(fil, nt, p)
            compare_state(model, fils)
            du = random_update(rng) * 0.3
            model.update(du)
            oracle_update(fils, du, p)
            compare_state(model, fils)
    finally:
        model.close()


STEPS = 2


def elastic_case():
    body = G.small_body()
    nfil, nseg = body.num_fil, body.num_seg
    dl, kb, dt, tol = 2.2, 1800.0, 0.1, 1e-8
    efm = 220.0 * kb / (dl * dl * nseg * nseg)
    ex = np.array([1.0, 0.0, 0.0])
    q0 = np.array([np.roll(Rotation.align_vectors([nrm], [ex])[0].as_quat(), 1) for nrm in body.fil_normal])
    qs = np.repeat(q0[:, None, :], nseg, axis=1)
    pert = np.random.default_rng(0).standard_normal((nfil, 6)) * 0.1 * efm
    return dict(body=body, nfil=nfil, nseg=nseg, nb=body.num_blobs, dl=dl, kb=kb, dt=dt, tol=tol, efm=efm, qs=qs,
                pert=pert)


def residual_state(x, xm1, xm2, u, um1, v):
    z3, z6 = np.zeros(3), np.zeros(6)
    return {"seg_x": x, "seg_xm1": xm1, "seg_xm2": xm2, "seg_u": u, "seg_um1": um1, "v_segs": v,
            "body_x": z3, "body_xm1": z3, "body_xm2": z3, "body_u": z3, "body_um1": z3, "v_bodies": z6}


@pytest.fixture(scope="module")
def oracle_run():
    """The elastic time loop (cilia_sim_main.cu:353-502, body held still) on the oracle."""
    c = elastic_case()
    body, nfil, nseg, nb, dl, dt, tol = c["body"], c["nfil"], c["nseg"], c["nb"], c["dl"], c["dt"], c["tol"]
    ns = nfil * nseg
    p = F.Params(nseg=nseg, dl=dl, kb=c["kb"], kt=c["kb"], dt=dt, end_force=c["efm"])
    fils = [F.Filament.at_rest(body.fil_base[i], c["qs"][i], nseg, dl) for i in range(nfil)]
    x_start = np.concatenate([f.x.reshape(-1) for f in fils])
    lu = sla.lu_factor(O.dense_mobility(np.zeros(0), body.x_blobs))

    def vel(x, f):
        vb = O.block("bs", x, body.x_blobs, f, np.zeros(3 * nb), seg_dof=6, bs_accumulate=False)
        lam = sla.lu_solve(lu, -vb)
        vs = O.block("ss", x, body.x_blobs, f, lam, seg_dof=6)
        return O.block("sb", x, body.x_blobs, f, lam, vs, seg_dof=6)

    cat = lambda name: np.concatenate([getattr(fl, name).reshape(-1) for fl in fils])
    iters, xs = [], []
    for nt in range(STEPS):
        for fl in fils:
            F.initial_guess(fl, nt, p)

        def evaluate():
            f = np.concatenate([F.internal_forces(fl, p, {4: c["pert"][i, :3], 12: c["pert"][i, 3:]}
                                                  if nt == 0 else None).reshape(-1) for i, fl in enumerate(fils)])
            v = vel(cat("x"), f)
            e, big = BO.compute_errors(residual_state(cat("x"), cat("xm1"), cat("xm2"), cat("u"), cat("um1"), v),
                                       1, nfil, nseg, nt >= 1, dt, dl, tol)
            return f, e, big

        f, e, big = evaluate()
        inv = [np.linalg.inv(F.approx_jacobian(fl, f.reshape(nfil, nseg, 6)[i], nt >= 1, p))
               for i, fl in enumerate(fils)] + [np.eye(6)]
        ref = BO.BroydenRef(6 * (ns + 1), 400, 0.1, inv)
        ref.start(e)
        while big and ref.iter < 400:
            du = ref.find_update()
            for i, fl in enumerate(fils):
                F.update(fl, du[6 * nseg * i:6 * nseg * (i + 1)], p)
            f, e, big = evaluate()
            ref.end_of_iter(e, big)
        iters.append(ref.iter)
        xs.append(cat("x"))
    return c, iters, xs, x_start


def check_against_oracle(oracle_run, iters, xs):
    _, ref_iters, ref_x, x_start = oracle_run
    print("broyden iterations", iters, "oracle", ref_iters,
          "max |dx|", [float(np.abs(a - b).max()) for a, b in zip(xs, ref_x)], flush=True)
    for s in range(STEPS):
        assert ref_iters[s] >= 10 and abs(iters[s] - ref_iters[s]) <= max(2, ref_iters[s] // 10)
        assert np.abs(xs[s] - ref_x[s]).max() <= 1e-6
        # the filaments moved (the end force and the kick act)
        assert np.abs(ref_x[s] - (ref_x[s - 1] if s else x_start)).max() > 1e-3


def test_elastic_implicit_step_matches_oracle_loop(oracle_run):
    print("case elastic implicit step 16x20 + 2000 blobs, host loop", flush=True)
    c = oracle_run[0]
    body, nfil, nseg, nb, dl, dt, tol = c["body"], c["nfil"], c["nseg"], c["nb"], c["dl"], c["dt"], c["tol"]
    ns = nfil * nseg
    zero3, zero6 = np.zeros(3), np.zeros(6)
    model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=c["kb"], kt=c["kb"], dt=dt, end_force=c["efm"])
    ev = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10)
    eng = BroydenSolver(model.block_sizes())
    iters, xs = [], []
    try:
        model.set_state(body.fil_base.reshape(-1), c["qs"].reshape(-1))
        model.set_perturbation(4, 12, c["pert"].reshape(-1))
        n6 = 6 * nseg
        blocks = np.zeros(nfil * n6 * n6 + 36)
        blocks[nfil * n6 * n6:] = np.eye(6).reshape(-1)
        for nt in range(STEPS):
            model.initial_guess(nt)

            def evaluate():
                f = model.forces(nt)
                v, _, _ = ev.compute(model.read(FM.X), f, body.x_blobs, zero3, zero6)
                big = eng.compute_errors(residual_state(*[model.read(w) for w in (FM.X, FM.XM1, FM.XM2, FM.U,
                                                                                   FM.UM1)], v),
                                         1, num_swim=1, num_fil=nfil, num_seg=nseg, bdf2=nt >= 1, dt=dt, dl=dl,
                                         tol=tol)
                return f, big

            f, big = evaluate()
            model.jacobian(nt, f, blocks)
            eng.set_jacobian(blocks)
            eng.start(eng.read_error(1))
            while big and eng.iter < 400:
                model.update(eng.find_update())
                _, big = evaluate()
                eng.end_of_iter(eng.read_error(1), big)
            iters.append(eng.iter)
            xs.append(model.read(FM.X))
    finally:
        model.close()
        ev.close()
        eng.close()
    check_against_oracle(oracle_run, iters, xs)


@pytest.mark.parametrize("precond", [PRECOND_REFERENCE, PRECOND_BLOCKS, BLOB_DIRECT])
def test_device_resident_elastic_step(oracle_run, precond):
    print(f"case elastic implicit step 16x20 + 2000 blobs, tfs_elastic_step, precond {precond}", flush=True)
    c = oracle_run[0]
    body, nfil, nseg, nb, dl, dt, tol = c["body"], c["nfil"], c["nseg"], c["nb"], c["dl"], c["dt"], c["tol"]
    ns = nfil * nseg
    dev = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), device="cuda")
    model = FM.FilamentModel(1, nfil, nseg, dl=dl, kb=c["kb"], kt=c["kb"], dt=dt, end_force=c["efm"],
                             memory=L.MEMORY_DEVICE)
    ev = ElasticVelocities(1, ns, nb, mode=SOLVE_FP64, restart=30, max_cycles=200, tol=1e-10,
                           memory=L.MEMORY_DEVICE)
    ev.set_preconditioner(precond)
    eng = BroydenSolver(model.block_sizes(), memory=L.MEMORY_DEVICE)
    x_blobs, x_bodies, v_bodies, f_blobs = dev(body.x_blobs), dev(np.zeros(3)), dev(np.zeros(6)), dev(np.zeros(3 * nb))
    step = None
    iters, xs = [], []
    try:
        model.set_state(dev(body.fil_base.reshape(-1)), dev(c["qs"].reshape(-1)))
        model.set_perturbation(4, 12, dev(c["pert"].reshape(-1)))
        step = FM.ElasticStep(model, ev, eng, x_blobs, x_bodies, v_bodies, f_blobs, tol=tol, max_iter=400)
        for nt in range(STEPS):
            st = step.step(nt)
            print("step", nt, st, flush=True)
            assert st["converged"] == 1 and st["velocity_solves"] == st["broyden_iterations"] + 1
            iters.append(st["broyden_iterations"])
            xs.append(model.read(FM.X))
    finally:
        if step is not None:
            step.close()
        model.close()
        ev.close()
        eng.close()
    check_against_oracle(oracle_run, iters, xs)
