"""Pins the numpy restatement of the reference's solve layer
(oracle/solver_oracle.py) against facts it does not share code with:
GMRES agrees with a direct dense solve; the preconditioner is the exact
inverse of [D -K; -K^T 0]; the saddle matrix is symmetric; the cyclic Hermite
spline reproduces its nodes and is third-order accurate; the beat's
vel_dir_phase / vel_dir_angle are the phase / angle derivatives of the
segment positions (finite differences)."""
import numpy as np
import pytest

from oracle import solver_oracle as SO
from timfilsimmod_amd import beat as B
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.solver import beat_setup


@pytest.fixture(scope="module")
def small():
    body = G.make_body(3, 6, 40, 6.0)
    su = beat_setup(body)
    geo = SO.beat_geometry(su, body.num_seg, B.BEAT_A, B.BEAT_B)
    A = SO.saddle_matrix(geo["x_segs"], body.x_blobs, geo["kphi"], geo["ktheta"], 3, 6)
    return body, su, geo, A


def test_saddle_matrix_symmetric(small):
    _, _, _, A = small
    assert np.abs(A - A.T).max() < 1e-14 * np.abs(A).max()


def test_preconditioner_is_exact_inverse(small):
    body, _, geo, _ = small
    P = SO.precond_matrix(geo["kphi"], geo["ktheta"], 3, 6, body.num_blobs)
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.normal(size=P.shape[0])
        w = SO.precondition_ref(v, geo["kphi"], geo["ktheta"], 3, 6, body.num_blobs)
        np.testing.assert_allclose(P @ w, v, rtol=0, atol=1e-11 * np.abs(v).max())


def test_gmres_ref_matches_direct_solve(small):
    body, _, geo, A = small
    apply_P = lambda v: SO.precondition_ref(v, geo["kphi"], geo["ktheta"], 3, 6, body.num_blobs)
    rhs = np.random.default_rng(1).normal(size=A.shape[0])
    x, it = SO.gmres_ref(lambda v: A @ v, apply_P, rhs, np.zeros_like(rhs), 1e-13, 350)
    np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=0, atol=1e-9 * np.abs(x).max())
    assert 0 < it <= A.shape[0]
    # warm start at the solution: no iterations (mobility_solver.cpp:2304-2308)
    x2, it2 = SO.gmres_ref(lambda v: A @ v, apply_P, rhs, x, 1e-10, 350)
    assert it2 == 0 and np.array_equal(x2, x)
    # a rubbish warm start is dropped (mobility_solver.cpp:2310-2317)
    x3, _ = SO.gmres_ref(lambda v: A @ v, apply_P, rhs, 1e6 * np.ones_like(rhs), 1e-13, 350)
    np.testing.assert_allclose(x3, x, atol=1e-9 * np.abs(x).max())


def test_hermite_nodes_and_order():
    tab = np.cos(2 * np.pi * np.arange(40) / 40) + 0.3
    for k in (0, 7, 39):
        assert SO.hermite_ref(2 * np.pi * k / 40, tab) == pytest.approx(tab[k], abs=1e-14)
    errs = []
    for n in (40, 80):
        t = np.cos(2 * np.pi * np.arange(n) / n)
        ph = np.linspace(0.1, 6.2, 57)
        errs.append(max(abs(SO.hermite_ref(p, t) - np.cos(p)) for p in ph))
    assert errs[1] < errs[0] / 6   # at least ~third order


@pytest.mark.parametrize("which", ["phase", "angle"])
def test_beat_directions_are_derivatives_of_positions(small, which):
    body, su, geo, _ = small
    h = 1e-6
    ph, an = su.phase0.copy(), su.angle0.copy() + 0.2
    base = SO.beat_geometry(su, 6, B.BEAT_A, B.BEAT_B, phase=ph, angle=an)
    if which == "phase":
        up = SO.beat_geometry(su, 6, B.BEAT_A, B.BEAT_B, phase=ph + h, angle=an)
        dn = SO.beat_geometry(su, 6, B.BEAT_A, B.BEAT_B, phase=ph - h, angle=an)
        key = "kphi"
    else:
        up = SO.beat_geometry(su, 6, B.BEAT_A, B.BEAT_B, phase=ph, angle=an + h)
        dn = SO.beat_geometry(su, 6, B.BEAT_A, B.BEAT_B, phase=ph, angle=an - h)
        key = "ktheta"
    fd = (up["x_segs"] - dn["x_segs"]) / (2 * h)
    np.testing.assert_allclose(base[key], fd, atol=1e-6 * np.abs(fd).max())


def test_beat_table_matches_package_shape():
    # the package's vectorised shape() and the oracle's fitted_shape agree
    for ph in (0.0, 1.3, 4.0):
        for s in (0.1, 0.5, 1.0):
            pos, der = SO.fitted_shape(s, ph, B.BEAT_A, B.BEAT_B)
            p2, d2 = B.shape(np.array(s), ph)
            np.testing.assert_allclose(pos, p2, atol=1e-14)
            np.testing.assert_allclose(der, d2, atol=1e-14)


def test_beat_has_unit_mean_length():
    # the fit is normalised so the period-averaged curve length is 1 (filament.cpp:684-690)
    s = np.linspace(0, 1, 401)
    lens = []
    for ph in np.linspace(0, 2 * np.pi, 60, endpoint=False):
        pos, _ = B.shape(s, ph)
        lens.append(np.sum(np.linalg.norm(np.diff(pos, axis=0), axis=1)))
    assert abs(np.mean(lens) - 1.0) < 0.02
