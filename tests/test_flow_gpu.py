"""GPU parity of the flow-field evaluator (include/tfs_flow.h) against the FP64
oracle: Stokeslet+rotlet mode (with and without net-force removal, with the
Blakelet images when there are no blobs) and RPY tracer mode, FP64 <= 1e-12
and FP32-tile/FP64-accumulate <= 1e-6 relative L2; plus a size-independent
check at full plane size (linearity in the sources)."""
import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import geometry as G
from timfilsimmod_amd.flow import RPY, STOKESLET_ROTLET, FlowFieldEvaluator

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def sources(nbod, nfil, nseg, nblob, seed=0):
    bodies = [G.make_body(nfil, nseg, nblob, 10.0, centre=(30.0 * b, 5.0 * b, -3.0 * b), force_seed=seed + b)
              for b in range(nbod)]
    cat = lambda k: np.concatenate([getattr(b, k) for b in bodies]) if bodies else np.zeros(0)
    return cat("x_segs"), cat("f_segs"), cat("x_blobs"), cat("f_blobs"), \
        np.concatenate([b.centre for b in bodies])


def targets(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(-40, 60, 3 * n)


@pytest.mark.parametrize("nbod,nblob,remove", [(1, 150, False), (2, 150, True), (1, 0, False), (3, 0, True)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stokeslet_rotlet_mode_matches_oracle(nbod, nblob, remove, precision):
    print(f"case nbod={nbod} nblob={nblob} remove={remove} {precision}", flush=True)
    xs, fs, xb, fb, xbody = sources(nbod, 5, 8, nblob)
    if nblob == 0:   # half-space case: sources above the wall z = 0
        xs = xs.reshape(-1, 3)
        xs[:, 2] = np.abs(xs[:, 2]) + 2.0
        xs = xs.reshape(-1)
    pts = targets(1000)
    if nblob == 0:
        pts = pts.reshape(-1, 3)
        pts[:, 2] = np.abs(pts[:, 2])
        pts = pts.reshape(-1)
    ref = O.flow_velocities(pts, xs, fs, xb, fb, xbody, num_bodies=nbod, mu=0.8, remove_net_force=remove)
    ev = FlowFieldEvaluator(nbod, 40, nblob, mode=STOKESLET_ROTLET, precision=precision,
                            remove_net_force=remove, mu=0.8)
    try:
        ev.set_sources(xs, fs, xb if nblob else None, fb if nblob else None, xbody)
        got = ev.velocities(pts)
    finally:
        ev.close()
    assert rel(got, ref) <= (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_rpy_mode_matches_oracle(precision, block):
    xs, fs, xb, fb, xbody = sources(1, 6, 10, 300)
    pts = np.concatenate([targets(777), xs[:30] + 0.3, xb[:30] * 1.01])   # includes near/inside points
    ref = O.flow_rpy(pts, xs, fs, xb, fb, a_seg=1.0, a_blob=1.2, mu=1.1)
    ev = FlowFieldEvaluator(1, 60, 300, mode=RPY, precision=precision, seg_radius=1.0, blob_radius=1.2, mu=1.1)
    try:
        ev.set_block(block)
        ev.set_sources(xs, fs, xb, fb)
        got = ev.velocities(pts)
    finally:
        ev.close()
    assert rel(got, ref) <= (1e-12 if precision == "fp64" else 1e-6)


def test_flow_source_body_plane_sample_and_linearity():
    """The flow-field workload's sources (1,280 x 20 segments + 9,000 blobs) on
    a 64 x 64 sample of the plane vs the oracle; linearity on 2048^2 points."""
    body = G.flow_source_body()
    half = body.radius + 1.2 * body.dl * (body.num_seg - 1)
    pts = G.plane_targets(64, half)
    ref = O.flow_velocities(pts, body.x_segs, body.f_segs, body.x_blobs, body.f_blobs, body.centre)
    ev = FlowFieldEvaluator(1, body.num_segs, body.num_blobs, precision="fp64")
    ev32 = FlowFieldEvaluator(1, body.num_segs, body.num_blobs, precision="fp32")
    try:
        for e in (ev, ev32):
            e.set_sources(body.x_segs, body.f_segs, body.x_blobs, body.f_blobs, body.centre)
        e64, e32 = rel(ev.velocities(pts), ref), rel(ev32.velocities(pts), ref)
        print(f"plane sample: fp64 {e64:.2e} fp32 {e32:.2e}")
        assert e64 <= 1e-12 and e32 <= 1e-6
        big = G.plane_targets(2048, half)
        v1 = ev.velocities(big)
        ev.set_sources(body.x_segs, 2.0 * body.f_segs, body.x_blobs, 2.0 * body.f_blobs, body.centre)
        v2 = ev.velocities(big)
        assert np.array_equal(2.0 * v1, v2)   # exact: scaling by 2 is exact in binary
    finally:
        ev.close()
        ev32.close()
