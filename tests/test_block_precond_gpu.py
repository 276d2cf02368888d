"""GPU parity of the block preconditioner (tfs_solver_set_preconditioner,
TFS_PRECOND_BLOCKS): P = [D -K; -K^T 0] with D the block-diagonal of every
filament's own RPY mobility and of the blobs' mobility within patches of 16
blobs consecutive in Morton order.

The oracle side builds P densely from the oracle's grand mobility (its
filament and patch sub-blocks) and the oracle's K; the GPU's P^{-1} must match
a dense solve with it to 1e-10 relative.  Solves with the block form meet the
same tolerance as with the reference's diagonal form, agree with a dense
direct solve, and take fewer iterations; a prescribed-beat run follows the
same trajectory."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import solver_oracle as SO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import (PRECOND_BLOCKS, PRECOND_REFERENCE, SOLVE_FP64, SOLVE_MIXED, BeatSimulation,
                                     Solver, beat_setup, solver_desc)

pytestmark = pytest.mark.gpu
PATCH = 16


def morton_order(x):
    """Blob order of the patches: 10-bit Z-order code over the bounding box
    (bit 3b + d holds bit b of axis d), stable."""
    p = x.reshape(-1, 3)
    lo, hi = p.min(0), p.max(0)
    span = hi - lo
    t = np.where(span > 0, (p - lo) / np.where(span > 0, span, 1.0), 0.0)
    q = np.minimum(1023, np.maximum(0, np.floor(t * 1023.0))).astype(np.int64)
    code = np.zeros(len(p), dtype=np.int64)
    for b in range(10):
        for d in range(3):
            code |= ((q[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind="stable")


def dense_block_precond(x_segs, x_blobs, kphi, ktheta, nfil, nseg):
    nb = x_blobs.size // 3
    ns3 = 3 * nfil * nseg
    M = O.dense_mobility(x_segs, x_blobs, seg_dof=3)
    D = np.zeros_like(M)
    for m in range(nfil):
        sl = slice(3 * nseg * m, 3 * nseg * (m + 1))
        D[sl, sl] = M[sl, sl]
    order = morton_order(x_blobs) if nb else np.zeros(0, dtype=int)
    for c in range(0, nb, PATCH):
        idx = order[c:c + PATCH]
        ii = ns3 + (3 * idx[:, None] + np.arange(3)[None, :]).reshape(-1)
        D[np.ix_(ii, ii)] = M[np.ix_(ii, ii)]
    K = SO.k_matrix(kphi, ktheta, nfil, nseg, nb)
    g = K.shape[1]
    return np.block([[D, -K], [-K.T, np.zeros((g, g))]])


def setup_case(nfil, nseg, nblob, radius, **kw):
    body = G.make_body(nfil, nseg, nblob, radius)
    su = beat_setup(body)
    geo = SO.beat_geometry(su, nseg, B.BEAT_A, B.BEAT_B)
    s = Solver(solver_desc(nfil, nseg, nblob, **kw))
    s.set_geometry(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"])
    return body, geo, s


# (2, 24, 20): 72-row filament blocks take the general (pivoting, multi-register) paths
@pytest.mark.parametrize("nfil,nseg,nblob", [(3, 6, 40), (8, 20, 300), (4, 5, 0), (2, 20, 37), (2, 24, 20)])
def test_block_preconditioner_inverts_the_block_saddle_matrix(nfil, nseg, nblob):
    print(f"case block precond nfil={nfil} nseg={nseg} nblob={nblob}", flush=True)
    body, geo, s = setup_case(nfil, nseg, nblob, 8.0)
    try:
        s.set_preconditioner(PRECOND_BLOCKS)
        P = dense_block_precond(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], nfil, nseg)
        for seed in (1, 2):
            v = np.random.default_rng(seed).normal(size=s.n)
            ref = np.linalg.solve(P, v)
            got = s.precondition(v)
            assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)
        # back to the reference form: bit for bit what it was
        s.set_preconditioner(PRECOND_REFERENCE)
        v = np.random.default_rng(3).normal(size=s.n)
        ref = SO.precondition_ref(v, geo["kphi"], geo["ktheta"], nfil, nseg, nblob)
        assert np.linalg.norm(s.precondition(v) - ref) <= 1e-12 * np.linalg.norm(ref)
    finally:
        s.close()


@pytest.mark.parametrize("mode", [SOLVE_FP64, SOLVE_MIXED])
def test_block_preconditioned_solve_matches_dense_and_saves_iterations(mode):
    print(f"case block precond solve mode={mode}", flush=True)
    body, geo, s = setup_case(8, 20, 300, 8.0, mode=mode, restart=50, tol=1e-10, max_cycles=200, inner_tol=1e-3,
                              max_refinements=20)
    try:
        A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], 8, 20)
        rhs = np.random.default_rng(5).normal(size=s.n)
        x_ref, st_ref = s.solve(rhs)
        s.set_preconditioner(PRECOND_BLOCKS)
        x, st = s.solve(rhs)
        print(st_ref, st)
        exact = np.linalg.solve(A, rhs)
        assert st["converged"] == 1
        assert np.linalg.norm(rhs - A @ x) <= 2e-10 * np.linalg.norm(rhs)
        assert np.linalg.norm(x - exact) <= 1e-6 * np.linalg.norm(exact)
        assert st["iterations"] < st_ref["iterations"]
    finally:
        s.close()


def test_beat_with_block_preconditioner_follows_the_same_trajectory():
    print("case block precond beat", flush=True)
    body = G.make_body(6, 20, 200, 8.0)
    runs = {}
    for kind in (PRECOND_REFERENCE, PRECOND_BLOCKS):
        sim = BeatSimulation(body, tol=1e-10, restart=60, max_cycles=50, precond=kind)
        try:
            stats = [sim.step() for _ in range(4)]
            runs[kind] = (stats, sim.state())
        finally:
            sim.close()
    (st0, s0), (st1, s1) = runs[PRECOND_REFERENCE], runs[PRECOND_BLOCKS]
    assert all(s["converged"] for s in st0 + st1)
    assert sum(s["iterations"] for s in st1) < sum(s["iterations"] for s in st0)
    assert np.abs(s1["x_segs"] - s0["x_segs"]).max() <= 1e-7 * np.abs(s0["x_segs"]).max()
    np.testing.assert_allclose(s1["phase"], s0["phase"], rtol=0, atol=1e-7)


def test_block_preconditioner_limits():
    s = Solver(solver_desc(1, 130, 10))
    try:
        with pytest.raises(L.TfsError):
            s.set_preconditioner(PRECOND_BLOCKS)
        with pytest.raises(L.TfsError):
            s.set_preconditioner(2)
    finally:
        s.close()
