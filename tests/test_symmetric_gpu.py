"""The symmetric form of the pair kernel (tfs_rpy_set_symmetric: each unordered
pair evaluated once, its term added to both particles' sums, the second
particle's sums rotating one lane per step) against the FP64 oracle
(oracle/rpy_oracle.c full_mobility) on the same seeded inputs:

* FP32 tiles, 3-DOF segments, <= 1e-6 relative L2, over sizes that leave the
  last tile and the last group of four tiles partly empty, one particle class
  only, overlapping pairs, unequal radii, and both BS write rules (the
  overwriting BS drops the blob-blob pairs);
* repeated evaluations bitwise identical (a fixed summation order);
* agreement with the default source-split shape;
* evaluations of a block subset take the default shape (bitwise equal to a
  handle without the symmetric form);
* the argument rules: FP32 tiles, 3-DOF segments and one body only, no
  persistent form; enable = 0 restores the default shape;
* a mixed-precision GMRES solve of the saddle-point system with the symmetric
  form on the solver's handle is as close to an FP64 solve at tolerance 1e-12
  as the same solve with the default shape, in as many iterations (+-2)."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.mobility import RpyHandle

pytestmark = pytest.mark.gpu

TOL = 1e-6
P32 = L.PRECISION_FP32_ACC64


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cluster(ns, nb, box, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-box, box, 3 * ns), rng.uniform(-box, box, 3 * nb),
            rng.normal(size=6 * ns), rng.normal(size=3 * nb))


def handle(ns, nb, xs, xb, fs, fb, bs_acc=True, a_blob=1.0):
    h = RpyHandle(ns, nb, seg_dof=3, precision=P32, bs_accumulate=bs_acc, seg_radius=1.0, blob_radius=a_blob)
    if ns:
        h.upload(L.SEG_POSITIONS, xs)
        h.upload(L.SEG_FORCES, fs)
    if nb:
        h.upload(L.BLOB_POSITIONS, xb)
        h.upload(L.BLOB_FORCES, fb)
    return h


def read(h, ns, nb):
    vs, vb = np.zeros(6 * ns), np.zeros(3 * nb)
    if ns:
        h.download(L.SEG_VELOCITIES, vs)
    if nb:
        h.download(L.BLOB_VELOCITIES, vb)
    h.sync()
    return np.concatenate([vs, vb])


@pytest.mark.parametrize("bs_acc", [True, False])
@pytest.mark.parametrize("ns,nb", [(700, 530), (100, 37), (129, 0), (0, 200), (257, 511), (3000, 2100)])
def test_symmetric_matches_oracle(ns, nb, bs_acc):
    print(f"case symmetric ns={ns} nb={nb} bs_accumulate={bs_acc}", flush=True)
    box = 1.1 * (ns + nb) ** (1 / 3) + 2.0  # dense enough for overlapping pairs
    xs, xb, fs, fb = cluster(ns, nb, box, ns * 3 + nb)
    rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=3, bs_accumulate=bs_acc, a_seg=1.0, a_blob=1.3)
    ref = np.concatenate([rs, rb])
    h = handle(ns, nb, xs, xb, fs, fb, bs_acc, a_blob=1.3)
    try:
        h.set_symmetric(P32)
        assert h.symmetric(P32)
        h.evaluate()
        got = read(h, ns, nb)
        if not np.any(ref):
            # blobs only under an overwriting BS: every velocity is M_bs f_s = 0
            assert not np.any(got)
        else:
            err = rel(got, ref)
            print(f"  rel err {err:.2e}", flush=True)
            assert err <= TOL
        if ns:
            assert np.all(got[:6 * ns].reshape(ns, 6)[:, 3:] == 0.0)
        h.evaluate()
        assert np.array_equal(read(h, ns, nb), got)
    finally:
        h.close()


def test_symmetric_agrees_with_split_and_falls_back_for_block_subsets():
    ns, nb = 1500, 1100
    xs, xb, fs, fb = cluster(ns, nb, 14.0, 5)
    h = handle(ns, nb, xs, xb, fs, fb)
    plain = handle(ns, nb, xs, xb, fs, fb)
    try:
        h.evaluate()
        split = read(h, ns, nb)
        h.set_symmetric(P32)
        h.evaluate()
        sym = read(h, ns, nb)
        print(f"case symmetric against split: rel diff {rel(sym, split):.2e}", flush=True)
        assert rel(sym, split) <= TOL
        # a block subset: the default shape, bit for bit on the velocities it
        # writes (segments for SS, SB; blobs for BB, BS)
        for blocks in (L.BLOCK_SS | L.BLOCK_SB, L.BLOCK_BB | L.BLOCK_BS, L.BLOCK_SS):
            h.evaluate(blocks)
            plain.evaluate(blocks)
            part = slice(0, 6 * ns) if blocks & L.BLOCK_SS else slice(6 * ns, None)
            assert np.array_equal(read(h, ns, nb)[part], read(plain, ns, nb)[part])
        h.set_symmetric(P32, False)
        assert not h.symmetric(P32)
        assert h.tile_shape(P32) == plain.tile_shape(P32)
        h.evaluate()
        assert np.array_equal(read(h, ns, nb), split)
    finally:
        h.close()
        plain.close()


def test_symmetric_argument_rules():
    ns, nb = 64, 64
    xs, xb, fs, fb = cluster(ns, nb, 5.0, 3)
    h = handle(ns, nb, xs, xb, fs, fb)
    h6 = RpyHandle(ns, nb, seg_dof=6, precision=P32, bs_accumulate=True)
    hb = handle(ns, nb, xs, xb, fs, fb)
    try:
        with pytest.raises(RuntimeError):
            h.set_symmetric(L.PRECISION_FP64)           # FP32 tiles only
        with pytest.raises(RuntimeError):
            h6.set_symmetric(P32)                       # 3-DOF segments only
        hb.set_bodies([32, 32], [40, 24])
        with pytest.raises(RuntimeError):
            hb.set_symmetric(P32)                       # one body only
        h.set_symmetric(P32)
        with pytest.raises(RuntimeError):
            h.set_persistent(P32, 2)                    # no persistent form
        h.set_symmetric(P32, False)
        h.set_tile_shape(P32, 128, 1, 128)
        h.set_persistent(P32, 2)
        with pytest.raises(RuntimeError):
            h.set_symmetric(P32)                        # nor from a persistent shape
    finally:
        h.close()
        h6.close()
        hb.close()


def test_gmres_with_symmetric_form_meets_tolerance():
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.solver import SOLVE_FP64, SOLVE_MIXED, BeatSimulation

    body = G.small_body()
    print("case GMRES, 16 x 20 segments + 2,000 blobs, mixed, symmetric form on the solver's handle", flush=True)
    out = {}
    for name, mode, tol, sym in (("truth", SOLVE_FP64, 1e-12, False), ("mixed", SOLVE_MIXED, 1e-10, False),
                                 ("mixed_sym", SOLVE_MIXED, 1e-10, True)):
        sim = BeatSimulation(body, mode=mode, tol=tol)
        try:
            if sym:
                L.check(L.lib().tfs_rpy_set_symmetric(sim.rpy_handle(), P32, 1))
            st = sim.step()
            assert st["converged"] == 1
            out[name] = (sim.state()["solution"], st)
        finally:
            sim.close()
    x_t = out["truth"][0]
    e_mixed, e_sym = rel(out["mixed"][0], x_t), rel(out["mixed_sym"][0], x_t)
    print("iterations", out["mixed"][1]["iterations"], out["mixed_sym"][1]["iterations"],
          "errors against the FP64 truth", e_mixed, e_sym, flush=True)
    assert e_sym <= max(2 * e_mixed, 1e-8)
    assert abs(out["mixed_sym"][1]["iterations"] - out["mixed"][1]["iterations"]) <= 2
