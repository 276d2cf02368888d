"""GPU parity of the saddle-point operator, preconditioner, GMRES solve and
prescribed-beat step against the FP64 oracle (oracle/solver_oracle.py).

Bars: operator and preconditioner <= 1e-12 relative (FP64 tiles); solves meet
their stated residual tolerance and agree with a dense direct solve to the
tolerance times a modest condition factor; mixed mode (FP32-tile inner
matvecs, FP64 residuals) meets the same FP64 tolerance."""
import numpy as np
import pytest

from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import (SOLVE_FP64, SOLVE_MIXED, BeatSimulation, Solver, beat_setup,
                                     solver_desc)

pytestmark = pytest.mark.gpu


def setup_case(nfil, nseg, nblob, radius, **kw):
    body = G.make_body(nfil, nseg, nblob, radius)
    su = beat_setup(body)
    geo = SO.beat_geometry(su, nseg, B.BEAT_A, B.BEAT_B)
    s = Solver(solver_desc(nfil, nseg, nblob, **kw))
    s.set_geometry(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"])
    return body, su, geo, s


@pytest.mark.parametrize("nfil,nseg,nblob", [(3, 6, 40), (8, 20, 300), (1, 2, 0)])
def test_system_mult_and_preconditioner_match_oracle(nfil, nseg, nblob):
    print(f"case nfil={nfil} nseg={nseg} nblob={nblob}", flush=True)
    body, su, geo, s = setup_case(nfil, nseg, nblob, 8.0)
    try:
        A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], nfil, nseg)
        v = np.random.default_rng(3).normal(size=s.n)
        av = s.system_mult(v)
        assert np.linalg.norm(av - A @ v) <= 1e-12 * np.linalg.norm(A @ v)
        av32 = s.system_mult(v, L.PRECISION_FP32_ACC64)
        assert np.linalg.norm(av32 - A @ v) <= 1e-6 * np.linalg.norm(A @ v)
        pv = s.precondition(v)
        ref = SO.precondition_ref(v, geo["kphi"], geo["ktheta"], nfil, nseg, nblob)
        assert np.linalg.norm(pv - ref) <= 1e-12 * np.linalg.norm(ref)
    finally:
        s.close()


@pytest.mark.parametrize("mode", [SOLVE_FP64, SOLVE_MIXED])
@pytest.mark.parametrize("restart", [50, 7])
def test_solve_matches_dense_direct_solve(mode, restart):
    print(f"case mode={mode} restart={restart}", flush=True)
    body, su, geo, s = setup_case(8, 20, 300, 8.0, mode=mode, restart=restart, tol=1e-10,
                                  max_cycles=200, inner_tol=1e-3, max_refinements=20)
    try:
        A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], 8, 20)
        rhs = np.random.default_rng(5).normal(size=s.n)
        x, st = s.solve(rhs)
        exact = np.linalg.solve(A, rhs)
        relres = np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs)
        print(st, relres)
        assert st["converged"] == 1
        assert relres <= 2e-10
        assert np.linalg.norm(x - exact) <= 1e-6 * np.linalg.norm(exact)
        if mode == SOLVE_MIXED:
            assert st["fp32_matvecs"] > st["fp64_matvecs"] >= 2
    finally:
        s.close()


def test_reference_settings_reproduce_reference_gmres():
    """restart = min(n, 350), one cycle, tol 1e-4 (config.hpp:312-321): same
    iteration count as the oracle's restatement of solve_linear_system."""
    body, su, geo, s = setup_case(8, 20, 300, 8.0, restart=350, max_cycles=1, tol=1e-4)
    try:
        A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], 8, 20)
        apply_P = lambda v: SO.precondition_ref(v, geo["kphi"], geo["ktheta"], 8, 20, 300)
        rhs = np.random.default_rng(7).normal(size=s.n)
        x_ref, it_ref = SO.gmres_ref(lambda v: A @ v, apply_P, rhs, np.zeros(s.n), 1e-4, 350)
        x, st = s.solve(rhs)
        print(st, it_ref)
        assert abs(st["iterations"] - it_ref) <= 1
        assert np.linalg.norm(x - x_ref) <= 1e-3 * np.linalg.norm(x_ref)
        # warm start at the answer: no iterations
        _, st2 = s.solve(rhs, x)
        assert st2["iterations"] == 0
    finally:
        s.close()


def test_error_paths():
    body, su, geo, s = setup_case(3, 6, 40, 6.0, restart=2, max_cycles=1, tol=1e-14)
    try:
        rhs = np.random.default_rng(1).normal(size=s.n)
        x, st = s.solve(rhs, allow_unconverged=True)
        assert st["converged"] == 0
        with pytest.raises(L.TfsError) as e:
            s.solve(rhs)
        assert e.value.code == L.TFS_ERR_NOT_CONVERGED
        bad = rhs.copy()
        bad[5] = np.nan
        with pytest.raises(L.TfsError) as e:
            s.solve(bad)
        assert e.value.code == L.TFS_ERR_NON_FINITE
        z, st = s.solve(np.zeros(s.n), np.ones(s.n))   # zero rhs -> zero solution
        assert np.all(z == 0) and st["converged"] == 1
    finally:
        s.close()


def test_beat_step_matches_oracle_geometry_and_rhs():
    body = G.make_body(6, 20, 200, 8.0)
    sim = BeatSimulation(body, tol=1e-11, restart=60, max_cycles=50)
    try:
        st = sim.step()
        state = sim.state()
        su = sim.setup
        geo = SO.beat_geometry(su, 20, B.BEAT_A, B.BEAT_B)
        np.testing.assert_allclose(state["x_segs"], geo["x_segs"], rtol=0, atol=1e-12)
        A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], 6, 20)
        rhs = SO.rhs_ref(su, 20, 200, geo["x_segs"], body.x_blobs, body.centre, np.zeros(6))
        exact = np.linalg.solve(A, rhs)
        x = state["solution"]
        # the solve meets its tolerance on the dense system, and its error obeys
        # ||x - x*|| <= ||A^-1||_2 ||b - A x|| (slack for the dense solve's rounding)
        res = np.linalg.norm(rhs - A @ x)
        assert res <= 2e-11 * np.linalg.norm(rhs), st
        inv_norm = 1.0 / np.linalg.svd(A, compute_uv=False)[-1]
        assert np.linalg.norm(x - exact) <= 2.0 * inv_norm * res + 1e-12 * np.linalg.norm(exact), st
        # step 2: explicit phase update phase += phase_dot*dt (filament.cpp:1087-1097)
        phase_dot = state["solution"][3 * 120 + 3 * 200: 3 * 120 + 3 * 200 + 6]
        sim.step()
        st2 = sim.state()
        np.testing.assert_allclose(st2["phase"], su.phase0 + phase_dot / 300.0, rtol=0, atol=1e-12)
    finally:
        sim.close()


def test_mixed_trajectory_is_as_accurate_as_fp64():
    """20 beat steps: the mixed solve (FP32-tile inner matvecs) deviates from a
    tight FP64 trajectory no more than the FP64 solve at the same tolerance."""
    body = G.make_body(16, 20, 600, 12.0)
    runs = {}
    for name, mode, tol in (("truth", SOLVE_FP64, 1e-12), ("fp64", SOLVE_FP64, 1e-8),
                            ("mixed", SOLVE_MIXED, 1e-8)):
        sim = BeatSimulation(body, mode=mode, tol=tol, inner_tol=1e-3, max_cycles=100,
                             max_refinements=30)
        try:
            its = [sim.step()["iterations"] for _ in range(20)]
            runs[name] = (sim.state(), its)
        finally:
            sim.close()
    truth = runs["truth"][0]
    dev = {k: np.abs(runs[k][0]["x_segs"] - truth["x_segs"]).max() / 2.2 for k in ("fp64", "mixed")}
    sol_err = {k: np.linalg.norm(runs[k][0]["solution"] - truth["solution"]) /
               np.linalg.norm(truth["solution"]) for k in ("fp64", "mixed")}
    print(f"20 steps: deviation/DL {dev}, solution rel err {sol_err}, "
          f"iterations fp64 {runs['fp64'][1]} mixed {runs['mixed'][1]}")
    assert dev["mixed"] <= 3 * dev["fp64"] + 1e-9
    assert sol_err["mixed"] <= 3 * sol_err["fp64"] + 1e-9


@pytest.mark.parametrize("order", [1, 2])
def test_extrapolated_warm_start_follows_the_same_trajectory(order):
    """tfs_beat_set_warm_start: the linear / quadratic extrapolation of the last
    solutions starts each solve closer (smaller initial residual, fewer
    iterations once the history exists) and the trajectory stays that of the
    reference's warm start (the previous solution) to the solve tolerance."""
    print(f"case warm start order {order}", flush=True)
    body = G.make_body(6, 20, 200, 8.0)
    runs = {}
    for o in (0, order):
        sim = BeatSimulation(body, tol=1e-10, restart=60, max_cycles=50, warm_start=o)
        try:
            stats = [sim.step() for _ in range(6)]
            runs[o] = (stats, sim.state())
        finally:
            sim.close()
    (st0, s0), (st1, s1) = runs[0], runs[order]
    # identical until the history holds order + 1 solutions
    assert st1[0]["iterations"] == st0[0]["iterations"]
    for k in range(order + 1, 6):
        assert st1[k]["initial_relres"] < 0.3 * st0[k]["initial_relres"], (k, st1[k], st0[k])
    assert sum(s["iterations"] for s in st1[order + 1:]) < sum(s["iterations"] for s in st0[order + 1:])
    assert all(s["converged"] for s in st0 + st1)
    scale = np.abs(s0["x_segs"]).max()
    assert np.abs(s1["x_segs"] - s0["x_segs"]).max() <= 1e-7 * scale
    np.testing.assert_allclose(s1["phase"], s0["phase"], rtol=0, atol=1e-7)


def test_warm_start_order_is_checked():
    body = G.make_body(2, 20, 100, 6.0)
    sim = BeatSimulation(body)
    try:
        with pytest.raises(L.TfsError):
            sim.set_warm_start(3)
        with pytest.raises(L.TfsError):
            sim.set_warm_start(-1)
    finally:
        sim.close()
