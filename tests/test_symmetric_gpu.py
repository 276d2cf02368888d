"""The symmetric form of the pair kernel (tfs_rpy_set_symmetric: each unordered
pair evaluated once, its term added to both particles' sums, the second
particle's sums rotating one lane per step) against the FP64 oracle
(oracle/rpy_oracle.c full_mobility) on the same seeded inputs:

* FP32 tiles <= 1e-6 and FP64 tiles <= 1e-12 relative L2, 3-DOF segments,
  over sizes that leave the
  last tile and the last group of four tiles partly empty, one particle class
  only, overlapping pairs, unequal radii, and both BS write rules (the
  overwriting BS drops the blob-blob pairs);
* repeated evaluations bitwise identical (a fixed summation order), and hipGraph
  replays bitwise equal to the eager launch with eager work between replays;
* agreement with the default source-split shape;
* evaluations of a block subset take the default shape (bitwise equal to a
  handle without the symmetric form);
* independent bodies (block-diagonal mobility, pairs within a body only) of
  uneven sizes, whose groups of four tiles straddle bodies, against the
  per-body oracle, with per-body completion times from the same launch;
* the argument rules: 3-DOF segments only, no persistent form; enable = 0
  restores the default shape;
* a mixed-precision GMRES solve of the saddle-point system with the symmetric
  form on the solver's handle is as close to an FP64 solve at tolerance 1e-12
  as the same solve with the default shape, in as many iterations (+-2);
* the lock-step solve of independent bodies (converged bodies' items exit at
  once) with the symmetric form meets every body's FP64 tolerance and agrees
  with the same ensemble solved with the default shape."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import _lib as L
from timfilsimmod_amd.mobility import RpyHandle

pytestmark = pytest.mark.gpu

TOL = 1e-6
P32 = L.PRECISION_FP32_ACC64
P64 = L.PRECISION_FP64
TOLS = {P32: 1e-6, P64: 1e-12}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cluster(ns, nb, box, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-box, box, 3 * ns), rng.uniform(-box, box, 3 * nb),
            rng.normal(size=6 * ns), rng.normal(size=3 * nb))


def handle(ns, nb, xs, xb, fs, fb, bs_acc=True, a_blob=1.0, precision=P32):
    h = RpyHandle(ns, nb, seg_dof=3, precision=precision, bs_accumulate=bs_acc, seg_radius=1.0,
                  blob_radius=a_blob)
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


@pytest.mark.parametrize("precision", [P32, P64])
@pytest.mark.parametrize("bs_acc", [True, False])
@pytest.mark.parametrize("ns,nb", [(700, 530), (100, 37), (129, 0), (0, 200), (257, 511), (3000, 2100)])
def test_symmetric_matches_oracle(ns, nb, bs_acc, precision):
    print(f"case symmetric ns={ns} nb={nb} bs_accumulate={bs_acc} precision={precision}", flush=True)
    box = 1.1 * (ns + nb) ** (1 / 3) + 2.0  # dense enough for overlapping pairs
    xs, xb, fs, fb = cluster(ns, nb, box, ns * 3 + nb)
    rs, rb = O.full_mobility(xs, xb, fs, fb, seg_dof=3, bs_accumulate=bs_acc, a_seg=1.0, a_blob=1.3)
    ref = np.concatenate([rs, rb])
    h = handle(ns, nb, xs, xb, fs, fb, bs_acc, a_blob=1.3, precision=precision)
    try:
        h.set_symmetric(precision)
        assert h.symmetric(precision)
        h.evaluate()
        got = read(h, ns, nb)
        if not np.any(ref):
            # blobs only under an overwriting BS: every velocity is M_bs f_s = 0
            assert not np.any(got)
        else:
            err = rel(got, ref)
            print(f"  rel err {err:.2e}", flush=True)
            assert err <= TOLS[precision]
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


@pytest.mark.parametrize("precision", [P32, P64])
def test_symmetric_graph_replays_match_eager(precision):
    ns, nb = 2000, 1500
    xs, xb, fs, fb = cluster(ns, nb, 16.0, 9)
    h = handle(ns, nb, xs, xb, fs, fb, precision=precision)
    try:
        h.set_symmetric(precision)
        h.evaluate()
        eager = read(h, ns, nb)
        g = h.graph(L.BLOCK_ALL, 2)
        rng = np.random.default_rng(0)
        for k in range(4):
            # eager work between replays: overwrite the outputs with noise
            h.upload(L.SEG_VELOCITIES, rng.normal(size=6 * ns))
            h.upload(L.BLOB_VELOCITIES, rng.normal(size=3 * nb))
            h.sync()
            g.launch()
            assert np.array_equal(read(h, ns, nb), eager), k
        g.close()
    finally:
        h.close()


def test_symmetric_argument_rules():
    ns, nb = 64, 64
    xs, xb, fs, fb = cluster(ns, nb, 5.0, 3)
    h = handle(ns, nb, xs, xb, fs, fb)
    h6 = RpyHandle(ns, nb, seg_dof=6, precision=P32, bs_accumulate=True)
    try:
        with pytest.raises(RuntimeError):
            h6.set_symmetric(P32)                       # 3-DOF segments only
        with pytest.raises(RuntimeError):
            h6.set_symmetric(P64)
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


# (num_fil, num_blobs, radius): bodies of one or two tiles next to bodies of
# several groups, one without blobs, one without filaments
BODIES = [(3, 40, 5.0), (1, 0, 3.0), (0, 70, 6.0), (30, 400, 14.0), (2, 33, 4.0), (45, 900, 18.0), (5, 120, 7.0)]


@pytest.mark.parametrize("precision", [P32, P64])
@pytest.mark.parametrize("bs_acc", [True, False])
def test_symmetric_independent_bodies_match_per_body_oracle(bs_acc, precision):
    from timfilsimmod_amd.ensemble import make_ensemble

    e = make_ensemble(sizes=BODIES, num_seg=10)
    print(f"case symmetric, {len(e.bodies)} independent bodies, bs_accumulate={bs_acc}", flush=True)
    vs, vb = [], []
    for b in e.bodies:
        s_, b_ = O.full_mobility(b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=3, bs_accumulate=bs_acc)
        vs.append(s_)
        vb.append(b_)
    ref = np.concatenate(vs + vb)
    h = RpyHandle(e.num_segs, e.num_blobs, seg_dof=3, precision=precision, bs_accumulate=bs_acc)
    try:
        h.set_bodies(e.segs_per_body, e.blobs_per_body)
        h.upload(L.SEG_POSITIONS, e.x_segs)
        h.upload(L.BLOB_POSITIONS, e.x_blobs)
        h.upload(L.SEG_FORCES, e.f_segs)
        h.upload(L.BLOB_FORCES, e.f_blobs)
        h.set_symmetric(precision)
        h.body_timing(True)
        h.evaluate()
        got = read(h, e.num_segs, e.num_blobs)
        err = rel(got, ref)
        print(f"  rel err {err:.2e}", flush=True)
        assert err <= TOLS[precision]
        ms = h.body_times(len(e.bodies))
        print("  per-body completion ms", ms, flush=True)
        assert np.all(np.isfinite(ms)) and np.all(ms >= 0) and ms.max() > 0
        h.body_timing(False)
        h.evaluate()
        assert np.array_equal(read(h, e.num_segs, e.num_blobs), got)
    finally:
        h.close()


def test_lockstep_ensemble_with_symmetric_form():
    from timfilsimmod_amd import geometry as G
    from timfilsimmod_amd.solver import PRECOND_BLOCKS, SOLVE_MIXED, EnsembleBeat, beat_setup

    sizes = [(4, 120, 8.0), (9, 300, 10.0), (2, 0, 6.0), (6, 60, 7.0)]
    bodies = [G.make_body(nf, 10, nb, r, force_seed=10 + i, perturb_seed=20 + i, centre=(60.0 * i, 0.0, 0.0))
              for i, (nf, nb, r) in enumerate(sizes)]
    setups = [beat_setup(b, phase_seed=3 + i) for i, b in enumerate(bodies)]
    kw = dict(mode=SOLVE_MIXED, restart=30, tol=1e-10, inner_tol=1e-4, max_refinements=20, max_cycles=100,
              precond=PRECOND_BLOCKS)
    print("case lock-step mixed ensemble, 4 bodies, symmetric form on the solver's handle", flush=True)
    out = {}
    for sym in (False, True):
        ens = EnsembleBeat(bodies, setups=setups, **kw)
        try:
            if sym:
                L.check(L.lib().tfs_rpy_set_symmetric(ens.rpy_handle(), P32, 1))
            for _ in range(3):
                st = ens.step()
                bs = ens.body_stats()
                assert st["converged"] == 1 and bs["converged"].all()
                assert (bs["relres"] <= 1e-10).all()
            out[sym] = ens.split(ens.state())
        finally:
            ens.close()
    for i in range(len(bodies)):
        assert rel(out[True][i]["phase"], out[False][i]["phase"]) <= 1e-9, i
        assert rel(out[True][i]["solution"], out[False][i]["solution"]) <= 1e-6, i


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
