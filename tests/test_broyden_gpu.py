"""GPU parity of the Broyden engine (include/tfs_broyden.h) against the oracle
(oracle/broyden_oracle.py).

Bars: every Broyden update <= 1e-11 relative to the oracle's when both see the
same residual sequence (FP64 throughout; the only differences are the summation
orders of the dot products, amplified at most by the 1/||Delta e|| scaling);
device Gauss-Jordan inverses <= 1e-12 relative to numpy's; residuals of
compute_errors within 1e-15 of the scale of their terms (the kernel evaluates
them uncontracted in the oracle's order; only dexpinv's trigonometric factor
can round differently)."""
import numpy as np
import pytest

from oracle import broyden_oracle as BO
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.broyden import BroydenSolver

pytestmark = pytest.mark.gpu


def _problem(seed, sizes, coupling=0.1, eps=0.2):
    rng = np.random.default_rng(seed)
    n = sum(sizes)
    J0 = [rng.normal(size=(s, s)) * 0.3 + 4.0 * np.eye(s) for s in sizes]
    A = np.zeros((n, n))
    r = 0
    for b in J0:
        A[r:r + b.shape[0], r:r + b.shape[0]] = b
        r += b.shape[0]
    A += coupling * rng.normal(size=(n, n)) / np.sqrt(n)
    rhs = rng.normal(size=n)
    return J0, (lambda x: A @ x + eps * np.sin(x) - rhs), n


def _engine(sizes, memory, **kw):
    """Engine plus the stream context its device vectors are made in: in device
    mode the engine works on a torch stream, so that torch's copies of its
    inputs and outputs are ordered with its kernels."""
    import contextlib
    if memory != L.MEMORY_DEVICE:
        return BroydenSolver(sizes, memory=memory, **kw), contextlib.nullcontext()
    import torch
    st = torch.cuda.Stream()
    return BroydenSolver(sizes, memory=memory, stream=st.cuda_stream, **kw), torch.cuda.stream(st)


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
def test_updates_match_oracle_step_by_step(memory):
    import torch
    print(f"case broyden lockstep memory={memory}", flush=True)
    sizes = [120] * 6 + [6, 7]
    J0, F, n = _problem(1, sizes)
    Jinv = [np.linalg.inv(b) for b in J0]
    flat = np.concatenate([b.reshape(-1) for b in Jinv])
    eng, ctx = _engine(sizes, memory, max_iter=30)
    ref = BO.BroydenRef(n, 30, 0.1, Jinv)
    device = memory == L.MEMORY_DEVICE
    dev = (lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()) if device else (lambda a: a)
    host = (lambda t: t.cpu().numpy()) if device else (lambda a: a)
    try:
        with ctx:
            eng.set_jacobian_inverse(dev(flat))
            x = np.zeros(n)
            e = F(x)
            eng.start(dev(e))
            ref.start(e)
            worst = 0.0
            big = True
            while big and ref.iter < 30:
                u = host(eng.find_update())
                ur = ref.find_update()
                worst = max(worst, _rel(u, ur))
                assert _rel(u, ur) <= 1e-11, (ref.iter, _rel(u, ur))
                x = x + ur
                e = F(x)
                big = np.linalg.norm(e) > 1e-10
                eng.end_of_iter(dev(e), big)
                ref.end_of_iter(e, big)
                assert eng.iter == ref.iter
            assert ref.iter >= 5
            err = np.linalg.norm(ref.error)
            assert abs(eng.error_norm() - err) <= 1e-12 * max(1.0, err)
        print("worst update deviation", worst, "iterations", ref.iter)
    finally:
        eng.close()


def test_gpu_broyden_loop_converges_like_oracle():
    print("case broyden loop", flush=True)
    sizes = [60] * 10 + [6]
    J0, F, n = _problem(2, sizes, coupling=0.2)
    eng = BroydenSolver(sizes)
    try:
        eng.set_jacobian(J0)     # inverted on the device
        x = np.zeros(n)
        e = F(x)
        eng.start(e)
        tol = 1e-10
        while np.linalg.norm(e) > tol and eng.iter < 400:
            x = x + eng.find_update()
            e = F(x)
            eng.end_of_iter(e, np.linalg.norm(e) > tol)
        xr, it_ref, norms = BO.solve(F, np.zeros(n), [np.linalg.inv(b) for b in J0], tol=tol)
        assert np.linalg.norm(F(x)) <= tol
        assert abs(eng.iter - it_ref) <= 2, (eng.iter, it_ref)
        assert _rel(x, xr) <= 1e-9
    finally:
        eng.close()


@pytest.mark.parametrize("sizes", [[120] * 8 + [6], [1, 7, 33, 200, 64], [512]])
def test_device_gauss_jordan_inverse(sizes):
    print(f"case gauss-jordan sizes={sizes}", flush=True)
    rng = np.random.default_rng(9)
    # general (non-symmetric) blocks that need row pivoting: a permuted, well-conditioned matrix
    blocks = []
    for s in sizes:
        M = rng.normal(size=(s, s)) + 3.0 * np.eye(s) * np.sqrt(s)
        blocks.append(M[rng.permutation(s)])
    n = sum(sizes)
    eng = BroydenSolver(sizes, confidence=1.0)
    try:
        eng.set_jacobian(blocks)
        e = rng.normal(size=n)
        eng.start(e)
        got = eng.find_update()
        want = -BO.block_diag_apply([np.linalg.inv(b) for b in blocks], e)
        assert _rel(got, want) <= 1e-12, _rel(got, want)
    finally:
        eng.close()


def test_singular_block_is_reported():
    print("case singular block", flush=True)
    blocks = [np.eye(4), np.ones((3, 3))]
    eng = BroydenSolver([4, 3])
    try:
        with pytest.raises(L.TfsError) as ei:
            eng.set_jacobian(blocks)
        assert ei.value.code == L.TFS_ERR_NON_FINITE
    finally:
        eng.close()


def _random_state(rng, nswim, nfil, nseg, scale_u=0.3):
    ns = nswim * nfil * nseg
    st = {}
    for pre, m in (("seg_", ns), ("body_", nswim)):
        for k in ("x", "xm1", "xm2"):
            st[pre + k] = rng.normal(size=3 * m) * 5.0
        for k in ("u", "um1"):
            st[pre + k] = rng.normal(size=3 * m) * scale_u
    st["v_segs"] = rng.normal(size=6 * ns)
    st["v_bodies"] = rng.normal(size=6 * nswim)
    # a few segments with tiny or zero rotation (small-angle branch of dexpinv)
    st["seg_u"][:3] = 0.0
    st["seg_u"][3:6] = 3e-6
    return st


def _scale(st, nswim, nfil, nseg, dt):
    """Per-entry magnitude of the terms a residual sums (for a rounding bar)."""
    n = 6 * (nswim * nfil * nseg + nswim)
    sc = np.zeros(n)
    for s in range(nswim):
        base = s * 6 * (nfil * nseg + 1)
        for i in range(nfil):
            for j in range(nseg):
                g = j + nseg * (i + s * nfil)
                p = base + 6 * i * nseg + 3 * j
                t = (np.abs(st["seg_x"][3 * g:3 * g + 3]) + np.abs(st["seg_xm1"][3 * g:3 * g + 3])
                     + np.abs(st["seg_xm2"][3 * g:3 * g + 3]) + dt * np.abs(st["v_segs"][6 * g:6 * g + 3]))
                sc[p:p + 3] = t
                om = np.linalg.norm(st["v_segs"][6 * g + 3:6 * g + 6])
                ul = np.abs(st["seg_u"][3 * g:3 * g + 3]) + np.abs(st["seg_um1"][3 * g:3 * g + 3])
                sc[p + 3 * nseg:p + 3 * nseg + 3] = ul + dt * om * (1 + np.linalg.norm(st["seg_u"][3 * g:3 * g + 3])) ** 2
        p = base + 6 * nfil * nseg
        sc[p:p + 3] = (np.abs(st["body_x"][3 * s:3 * s + 3]) + np.abs(st["body_xm1"][3 * s:3 * s + 3])
                       + np.abs(st["body_xm2"][3 * s:3 * s + 3]) + dt * np.abs(st["v_bodies"][6 * s:6 * s + 3]))
        om = np.linalg.norm(st["v_bodies"][6 * s + 3:6 * s + 6])
        sc[p + 3:p + 6] = (np.abs(st["body_u"][3 * s:3 * s + 3]) + np.abs(st["body_um1"][3 * s:3 * s + 3])
                           + dt * om * (1 + np.linalg.norm(st["body_u"][3 * s:3 * s + 3])) ** 2)
    return sc


@pytest.mark.parametrize("memory", [L.MEMORY_HOST, L.MEMORY_DEVICE])
@pytest.mark.parametrize("bdf2", [False, True])
@pytest.mark.parametrize("shape", [(1, 16, 20), (2, 3, 5), (1, 0, 0)])
def test_compute_errors_match_oracle(memory, bdf2, shape):
    import torch
    nswim, nfil, nseg = shape
    print(f"case compute_errors memory={memory} bdf2={bdf2} shape={shape}", flush=True)
    rng = np.random.default_rng(31)
    st = _random_state(rng, nswim, nfil, nseg)
    dt, dl, tol = 0.01, 2.2, 1e-4
    n = 6 * (nswim * nfil * nseg + nswim)
    ref, big_ref = BO.compute_errors(st, nswim, nfil, nseg, bdf2, dt, dl, tol)
    eng, ctx = _engine([n], memory)
    try:
        with ctx:
            arg = st if memory == L.MEMORY_HOST else {k: torch.from_numpy(v).cuda() for k, v in st.items()}
            big = eng.compute_errors(arg, 1, num_swim=nswim, num_fil=nfil, num_seg=nseg, bdf2=bdf2, dt=dt, dl=dl,
                                     tol=tol)
            got = eng.read_error(1)
            got = got.cpu().numpy() if memory == L.MEMORY_DEVICE else got
            sc = _scale(st, nswim, nfil, nseg, dt)
            assert np.all(np.abs(got - ref) <= 1e-15 * sc + 1e-300), np.max(np.abs(got - ref) / sc)
            assert big == big_ref
            # a state that satisfies the step equations is "small enough"; slot 0 then start()
            ex = {k: v.copy() for k, v in st.items()}
            c1, c2, c3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
            for pre, vk, m in (("seg_", "v_segs", nswim * nfil * nseg), ("body_", "v_bodies", nswim)):
                v = ex[vk].reshape(m, 6)
                xm1, xm2 = ex[pre + "xm1"].reshape(m, 3), ex[pre + "xm2"].reshape(m, 3)
                ex[pre + "x"] = (-(c1 * xm1 + c2 * xm2 + c3 * dt * v[:, :3]) if bdf2 else xm1 + dt * v[:, :3]).reshape(-1)
                u, um1 = ex[pre + "u"].reshape(m, 3), ex[pre + "um1"].reshape(m, 3)
                W = (u - c2 * um1) / (-c3 * dt) if bdf2 else u / dt
                for k in range(m):
                    v[k, 3:] = BO.dexp(u[k], W[k])
                ex[vk] = v.reshape(-1)
            arg = ex if memory == L.MEMORY_HOST else {k: torch.from_numpy(v).cuda() for k, v in ex.items()}
            assert not eng.compute_errors(arg, 0, num_swim=nswim, num_fil=nfil, num_seg=nseg, bdf2=bdf2, dt=dt,
                                          dl=dl, tol=tol)
            eng.start()
            assert eng.error_norm() < 1e-9
    finally:
        eng.close()


def test_compute_errors_reports_non_finite_and_bad_sizes():
    print("case compute_errors errors", flush=True)
    rng = np.random.default_rng(32)
    st = _random_state(rng, 1, 2, 3)
    st["v_segs"][7] = np.nan
    eng = BroydenSolver([6 * (6 + 1)])
    try:
        with pytest.raises(L.TfsError) as ei:
            eng.compute_errors(st, 1, num_swim=1, num_fil=2, num_seg=3, bdf2=True, dt=0.1, dl=1.0)
        assert ei.value.code == L.TFS_ERR_NON_FINITE
        with pytest.raises(L.TfsError) as ei:
            eng.compute_errors(st, 1, num_swim=1, num_fil=2, num_seg=4, bdf2=True, dt=0.1, dl=1.0)
        assert ei.value.code == L.TFS_ERR_INVALID_ARGUMENT
    finally:
        eng.close()
