"""Independent bodies solved in lock-step (tfs_solver_set_independent_bodies):
one GMRES per body inside every batched launch.

The bar is the one-body solve: an ensemble timestep gives each body the state
its own one-body BeatSimulation reaches with the same setup and settings (the
one-body path is pinned against the oracle in test_solver_gpu.py and
test_configs_gpu.py).  Rounding differs only in the order of the Arnoldi sums
(per-body slices vs one vector), so FP64 trajectories agree to far below the
solve tolerance and iteration counts to within one step per solve.  Mixed
mode evaluates the FP32 tiles in a fixed-point frame scaled over all bodies,
so its inner iterations may differ; every body still meets the FP64 outer
tolerance and its solution agrees with the one-body solution to the
tolerance times a modest condition factor.  Graph replay and eager launches
run the same kernels in the same order: bitwise equal."""
import numpy as np
import pytest

from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import (PRECOND_BLOCKS, SOLVE_FP64, SOLVE_MIXED, BeatSimulation,
                                     EnsembleBeat, Solver, beat_setup, solver_desc)

pytestmark = pytest.mark.gpu

# (num_fil, num_blobs, radius): uneven bodies, one without blobs
SIZES = [(4, 120, 8.0), (9, 300, 10.0), (2, 0, 6.0), (6, 60, 7.0)]
NSEG = 10


def make_bodies(sizes=SIZES, nseg=NSEG):
    return [G.make_body(nf, nseg, nb, r, force_seed=10 + i, perturb_seed=20 + i, centre=(60.0 * i, 0.0, 0.0))
            for i, (nf, nb, r) in enumerate(sizes)]


def run_steps(sim, steps):
    stats = [sim.step() for _ in range(steps)]
    return stats, sim.state()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("precond", [0, PRECOND_BLOCKS])
def test_ensemble_fp64_steps_match_separate_bodies(precond):
    print(f"case fp64 precond={precond}", flush=True)
    bodies = make_bodies()
    setups = [beat_setup(b, phase_seed=3 + i) for i, b in enumerate(bodies)]
    kw = dict(mode=SOLVE_FP64, restart=30, tol=1e-10, max_cycles=100, precond=precond)
    ens = EnsembleBeat(bodies, setups=setups, **kw)
    try:
        iters_e = []
        for _ in range(3):
            st = ens.step()
            assert st["converged"] == 1
            bs = ens.body_stats()
            assert bs["converged"].all() and (bs["relres"] <= 1e-10).all()
            iters_e.append(bs["iterations"].copy())
        parts = ens.split(ens.state())
    finally:
        ens.close()
    for i, (b, su) in enumerate(zip(bodies, setups)):
        one = BeatSimulation(b, setup=su, **kw)
        try:
            stats, state = run_steps(one, 3)
        finally:
            one.close()
        for k, st in enumerate(stats):
            assert abs(int(iters_e[k][i]) - st["iterations"]) <= 1, (i, k, iters_e[k][i], st["iterations"])
        for key in ("phase", "angle", "x_segs"):
            assert rel(parts[i][key], state[key]) <= 1e-12, (i, key)
        assert rel(parts[i]["solution"], state["solution"]) <= 1e-8, i


def test_ensemble_mixed_steps_meet_the_fp64_tolerance():
    print("case mixed", flush=True)
    bodies = make_bodies()
    setups = [beat_setup(b, phase_seed=3 + i) for i, b in enumerate(bodies)]
    kw = dict(mode=SOLVE_MIXED, restart=30, tol=1e-10, inner_tol=1e-4, max_refinements=20, max_cycles=100,
              precond=PRECOND_BLOCKS)
    ens = EnsembleBeat(bodies, setups=setups, **kw)
    try:
        for _ in range(3):
            st = ens.step()
            bs = ens.body_stats()
            assert st["converged"] == 1 and bs["converged"].all()
            assert (bs["relres"] <= 1e-10).all()
            assert st["fp32_matvecs"] > st["fp64_matvecs"]
            assert (bs["solve_ms"] >= 0).all() and bs["solve_ms"].max() > 0
        parts = ens.split(ens.state())
    finally:
        ens.close()
    for i, (b, su) in enumerate(zip(bodies, setups)):
        one = BeatSimulation(b, setup=su, **kw)
        try:
            _, state = run_steps(one, 3)
        finally:
            one.close()
        assert rel(parts[i]["phase"], state["phase"]) <= 1e-9, i
        assert rel(parts[i]["solution"], state["solution"]) <= 1e-6, i


@pytest.mark.parametrize("mode", [SOLVE_FP64, SOLVE_MIXED])
def test_graph_replay_equals_eager_launches(mode):
    print(f"case graphs mode={mode}", flush=True)
    bodies = make_bodies()
    kw = dict(mode=mode, restart=20, tol=1e-10, inner_tol=1e-4, max_refinements=20, max_cycles=100,
              precond=PRECOND_BLOCKS)
    out = {}
    for graphs in (True, False):
        ens = EnsembleBeat(bodies, graphs=graphs, **kw)
        try:
            _, state = run_steps(ens, 3)
            out[graphs] = (state, ens.body_stats())
        finally:
            ens.close()
    (sg, bg), (se, be) = out[True], out[False]
    assert bg["graph_cycles"] > 0 and bg["eager_cycles"] == 0
    assert be["eager_cycles"] > 0 and be["graph_cycles"] == 0
    for key in ("phase", "angle", "x_segs", "solution"):
        assert np.array_equal(sg[key], se[key]), key
    assert np.array_equal(bg["iterations"], be["iterations"])


def split_solution(x, fil, blobs, nseg):
    """Operator layout [lambda_s | lambda_b | phase | angle] -> per-body systems."""
    fo, bo = np.concatenate([[0], np.cumsum(fil)]), np.concatenate([[0], np.cumsum(blobs)])
    ns3, nf = 3 * fo[-1] * nseg, fo[-1]
    op, oa = ns3 + 3 * bo[-1], ns3 + 3 * bo[-1] + nf
    return [np.concatenate([x[3 * fo[i] * nseg:3 * fo[i + 1] * nseg], x[ns3 + 3 * bo[i]:ns3 + 3 * bo[i + 1]],
                            x[op + fo[i]:op + fo[i + 1]], x[oa + fo[i]:oa + fo[i + 1]]])
            for i in range(len(fil))]


def join_solution(parts, fil, blobs, nseg):
    segs, bl, ph, an = [], [], [], []
    for p, nf, nb in zip(parts, fil, blobs):
        segs.append(p[:3 * nf * nseg])
        bl.append(p[3 * nf * nseg:3 * nf * nseg + 3 * nb])
        ph.append(p[3 * nf * nseg + 3 * nb:3 * nf * nseg + 3 * nb + nf])
        an.append(p[3 * nf * nseg + 3 * nb + nf:])
    return np.concatenate(segs + bl + ph + an)


@pytest.mark.parametrize("mode", [SOLVE_FP64, SOLVE_MIXED])
def test_lockstep_solve_applies_each_bodys_own_rules(mode):
    """A zero right-hand side (zero solution), an exact warm start (no
    iterations) and a rubbish guess (dropped) in three bodies of one solve,
    against each body's one-body solve and the oracle's saddle matrix."""
    print(f"case rules mode={mode}", flush=True)
    bodies = make_bodies(SIZES[:3])
    fil = [b.num_fil for b in bodies]
    blobs = [b.num_blobs for b in bodies]
    geos = [SO.beat_geometry(beat_setup(b, phase_seed=3 + i), NSEG, B.BEAT_A, B.BEAT_B)
            for i, b in enumerate(bodies)]
    kw = dict(mode=mode, restart=25, tol=1e-10, inner_tol=1e-4, max_refinements=20, max_cycles=100)
    mats = [SO.saddle_matrix(g["x_segs"], b.x_blobs, g["kphi"], g["ktheta"], b.num_fil, NSEG)
            for g, b in zip(geos, bodies)]
    rng = np.random.default_rng(11)
    rhs = [rng.normal(size=A.shape[0]) for A in mats]
    rhs[1][:] = 0.0
    exact = [np.linalg.solve(A, r) for A, r in zip(mats, rhs)]
    x0 = [exact[0].copy(), rng.normal(size=mats[1].shape[0]), 1e6 * rng.normal(size=mats[2].shape[0])]
    s = Solver(solver_desc(sum(fil), NSEG, sum(blobs), **kw))
    try:
        s.set_geometry(np.concatenate([g["x_segs"] for g in geos]), np.concatenate([b.x_blobs for b in bodies]),
                       np.concatenate([g["kphi"] for g in geos]), np.concatenate([g["ktheta"] for g in geos]))
        s.set_independent_bodies(fil, blobs)
        x, st = s.solve(join_solution(rhs, fil, blobs, NSEG), join_solution(x0, fil, blobs, NSEG))
        bs = s.body_stats()
    finally:
        s.close()
    print(st, bs)
    parts = split_solution(x, fil, blobs, NSEG)
    assert st["converged"] == 1 and bs["converged"].all()
    assert bs["iterations"][0] == 0 and bs["iterations"][1] == 0 and bs["iterations"][2] > 0
    assert np.all(parts[1] == 0.0)
    for i in (0, 2):
        A = mats[i]
        assert np.linalg.norm(rhs[i] - A @ parts[i]) <= 1.05e-10 * np.linalg.norm(rhs[i]), i
        assert rel(parts[i], exact[i]) <= 1e-6, i
    # the same body alone
    b, g = bodies[2], geos[2]
    one = Solver(solver_desc(b.num_fil, NSEG, b.num_blobs, **kw))
    try:
        one.set_geometry(g["x_segs"], b.x_blobs, g["kphi"], g["ktheta"])
        x1, st1 = one.solve(rhs[2], x0[2])
    finally:
        one.close()
    assert rel(parts[2], x1) <= (1e-9 if mode == SOLVE_FP64 else 1e-6)
    if mode == SOLVE_FP64:
        assert abs(int(bs["iterations"][2]) - st1["iterations"]) <= 1


def test_independent_bodies_argument_checks():
    s = Solver(solver_desc(6, NSEG, 100, restart=30))
    try:
        with pytest.raises(L.TfsError, match="add up"):
            s.set_independent_bodies([3, 2], [50, 50])
        with pytest.raises(L.TfsError, match="particles"):
            s.set_independent_bodies([6, 0], [100, 0])
        with pytest.raises(L.TfsError, match="no independent bodies"):
            s.set_graphs(False)
        s.set_independent_bodies([3, 3], [40, 60])
        s.set_graphs(False)
        s.set_independent_bodies([6], [100])  # back to one system
    finally:
        s.close()
    big = Solver(solver_desc(6, NSEG, 100, restart=150))
    try:
        with pytest.raises(L.TfsError, match="restart"):
            big.set_independent_bodies([3, 3], [40, 60])
    finally:
        big.close()


def test_many_uneven_bodies_in_one_solve():
    """48 bodies of 3..24 filaments and 40..400 blobs, two mixed steps; eight
    of them checked against their one-body runs."""
    print("case 48 bodies", flush=True)
    u = G.uniform_open(5, 96)
    sizes = [(3 + int(u[2 * i] * 22), 40 + int(u[2 * i + 1] * 361), 6.0 + 10.0 * u[2 * i + 1]) for i in range(48)]
    bodies = make_bodies(sizes)
    setups = [beat_setup(b, phase_seed=3 + i) for i, b in enumerate(bodies)]
    kw = dict(mode=SOLVE_MIXED, restart=30, tol=1e-9, inner_tol=1e-4, max_refinements=20, max_cycles=100,
              precond=PRECOND_BLOCKS)
    ens = EnsembleBeat(bodies, setups=setups, **kw)
    try:
        for _ in range(2):
            ens.step()
            bs = ens.body_stats()
            assert bs["converged"].all() and (bs["relres"] <= 1e-9).all()
        parts = ens.split(ens.state())
    finally:
        ens.close()
    for i in (0, 5, 11, 17, 23, 31, 40, 47):
        one = BeatSimulation(bodies[i], setup=setups[i], **kw)
        try:
            _, state = run_steps(one, 2)
        finally:
            one.close()
        assert rel(parts[i]["phase"], state["phase"]) <= 1e-9, i
        assert rel(parts[i]["solution"], state["solution"]) <= 1e-5, i
