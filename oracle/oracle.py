"""ctypes front-end for the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker, and by bench.py's CPU-baseline leg.  The product never imports it.
Array arguments are float64 numpy arrays in the reference's AoS layouts.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        d, i = ctypes.c_double, ctypes.c_int
        lib.orc_rpy_pair.argtypes = [i, i, _dp, _dp, d, d, d, d, d, d, d, d, d]
        lib.orc_Mss.argtypes = [_dp, _dp, _dp, i, i, i, i, d, d]
        lib.orc_Msb.argtypes = [_dp, _dp, _dp, _dp, i, i, i, i, d, d, d]
        lib.orc_Mbb.argtypes = [_dp, _dp, _dp, i, i, i, d, d]
        lib.orc_Mbs.argtypes = [_dp, _dp, _dp, _dp, i, i, i, i, i, d, d, d]
        lib.orc_full_mobility.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i, i, i, i, d, d, d]
        lib.orc_dense_mobility.argtypes = [_dp, _dp, _dp, i, i, i, d, d, d]
        lib.orc_flow_velocities.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, i, i, i, i, d, i]
        lib.orc_flow_rpy.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, i, i, i, d, d, d]
        lib.orc_rpy_pair_wall.argtypes = [i, i, _dp, _dp, i, d, d, d, d, d, d]
        lib.orc_barrier_forces.argtypes = [_dp, _dp, _dp, _dp, i, i, i, d, d, d, i, d]
        lib.orc_set_domain.argtypes = [i]
        lib.orc_get_domain.restype = ctypes.c_int
        lib.orc_num_threads.restype = ctypes.c_int
        lib.orc_set_num_threads.argtypes = [i]
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _arr(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    if n is not None:
        assert a.size >= n, (a.size, n)
    return a


def num_threads() -> int:
    return _load().orc_num_threads()


def set_num_threads(n: int) -> None:
    _load().orc_set_num_threads(int(n))


def rpy_pair(vel_dof, force_dof, f, xi, ai, xj, aj, mu=1.0):
    """Velocity (vel_dof values) of particle i from force/torque f on j."""
    lib = _load()
    v = np.zeros(6)
    ff = np.zeros(6)
    ff[: len(f)] = f
    lib.orc_rpy_pair(vel_dof, force_dof, _p(v), _p(ff), float(xi[0]), float(xi[1]), float(xi[2]),
                     float(ai), float(xj[0]), float(xj[1]), float(xj[2]), float(aj), float(mu))
    return v[:vel_dof]


UNBOUNDED, HALF_SPACE = 0, 1


def set_domain(d: int) -> None:
    """Domain of the block sums and the dense mobility: UNBOUNDED or HALF_SPACE
    (no-slip wall at z = 0, radius 1 and viscosity 1)."""
    _load().orc_set_domain(int(d))


class domain:
    """Context manager: `with domain(HALF_SPACE): ...` restores the previous domain."""

    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.prev = _load().orc_get_domain()
        set_domain(self.d)

    def __exit__(self, *exc):
        set_domain(self.prev)


def rpy_pair_wall(vel_dof, force_dof, f, xi, xj, same=False):
    """Half-space pair tensor applied to f (cuda_functions.cu:12-203)."""
    lib = _load()
    ff = np.zeros(6)
    ff[:len(f)] = f
    v = np.zeros(6)
    lib.orc_rpy_pair_wall(vel_dof, force_dof, _p(v), _p(ff), int(bool(same)), float(xi[0]), float(xi[1]),
                          float(xi[2]), float(xj[0]), float(xj[1]), float(xj[2]))
    return v[:vel_dof]


def full_mobility(x_segs, x_blobs, f_segs, f_blobs, *, seg_dof=6, bs_accumulate=False,
                  a_seg=1.0, a_blob=1.0, mu=1.0, v_segs=None, v_blobs=None):
    """evaluate_full_mobility: returns (v_segs[6*Ns], v_blobs[3*Nb])."""
    lib = _load()
    xs, xb = _arr(x_segs), _arr(x_blobs)
    ns, nb = xs.size // 3, xb.size // 3
    fs, fb = _arr(f_segs, 6 * ns), _arr(f_blobs, 3 * nb)
    vs = np.zeros(6 * ns) if v_segs is None else _arr(v_segs).copy()
    vb = np.zeros(3 * nb) if v_blobs is None else _arr(v_blobs).copy()
    if ns == 0:
        xs = np.zeros(3); fs = np.zeros(6); vs_buf = np.zeros(6)
    else:
        vs_buf = vs
    if nb == 0:
        xb = np.zeros(3); fb = np.zeros(3); vb_buf = np.zeros(3)
    else:
        vb_buf = vb
    lib.orc_full_mobility(_p(vs_buf), _p(vb_buf), _p(fs), _p(fb), _p(xs), _p(xb), ns, nb,
                          int(seg_dof), int(bool(bs_accumulate)), float(a_seg), float(a_blob),
                          float(mu))
    return vs, vb


def block(name, x_segs, x_blobs, f_segs, f_blobs, v_init=None, *, seg_dof=6,
          bs_accumulate=False, a_seg=1.0, a_blob=1.0, mu=1.0, start=0, count=None):
    """One reference kernel: name in {"ss","sb","bb","bs"} with its write rule."""
    lib = _load()
    xs, xb = _arr(x_segs), _arr(x_blobs)
    ns, nb = xs.size // 3, xb.size // 3
    fs, fb = _arr(f_segs), _arr(f_blobs)
    if name in ("ss", "sb"):
        count = ns - start if count is None else count
        v = np.zeros(6 * count) if v_init is None else _arr(v_init).copy()
        if name == "ss":
            lib.orc_Mss(_p(v), _p(fs), _p(xs), start, count, ns, seg_dof, a_seg, mu)
        else:
            lib.orc_Msb(_p(v), _p(fb), _p(xs), _p(xb), start, count, nb, seg_dof, a_seg, a_blob, mu)
    else:
        count = nb - start if count is None else count
        v = np.zeros(3 * count) if v_init is None else _arr(v_init).copy()
        if name == "bb":
            lib.orc_Mbb(_p(v), _p(fb), _p(xb), start, count, nb, a_blob, mu)
        else:
            lib.orc_Mbs(_p(v), _p(fs), _p(xb), _p(xs), start, count, ns, seg_dof,
                        int(bool(bs_accumulate)), a_seg, a_blob, mu)
    return v


def dense_mobility(x_segs, x_blobs, *, seg_dof=6, a_seg=1.0, a_blob=1.0, mu=1.0):
    """Dense grand mobility, unknowns ordered seg (seg_dof each) then blob (3)."""
    lib = _load()
    xs, xb = _arr(x_segs), _arr(x_blobs)
    ns, nb = xs.size // 3, xb.size // 3
    n = seg_dof * ns + 3 * nb
    m = np.zeros((n, n))
    lib.orc_dense_mobility(_p(m), _p(xs if ns else np.zeros(3)), _p(xb if nb else np.zeros(3)),
                           ns, nb, seg_dof, a_seg, a_blob, mu)
    return m


def flow_velocities(x_flow, x_segs, f_segs, x_blobs, f_blobs, x_bodies, *, num_bodies=1,
                    prefac=None, mu=1.0, remove_net_force=False):
    """flow_velocities kernel (Stokeslet + rotlet [+ Blakelet if no blobs])."""
    lib = _load()
    xf = _arr(x_flow)
    nf = xf.size // 3
    xs, fs = _arr(x_segs), _arr(f_segs)
    xb, fb = _arr(x_blobs), _arr(f_blobs)
    ns_total, nb_total = xs.size // 3, xb.size // 3
    xbody = _arr(x_bodies)
    if prefac is None:
        prefac = 1.0 / (8.0 * np.pi * mu)
    v = np.zeros(3 * nf)
    lib.orc_flow_velocities(_p(v), _p(xf), _p(fs if ns_total else np.zeros(6)),
                            _p(fb if nb_total else np.zeros(3)), _p(xs if ns_total else np.zeros(3)),
                            _p(xbody), _p(xb if nb_total else np.zeros(3)), nf,
                            ns_total // num_bodies, num_bodies, nb_total // num_bodies,
                            float(prefac), int(bool(remove_net_force)))
    return v


def barrier_forces(x_segs, x_blobs, f_segs, *, nseg, strength, a_seg=1.0, a_blob=1.0, wall=False, dl=2.2):
    """barrier_forces kernel: (f_segs + barrier on slots 0-2, blob repulsion)."""
    lib = _load()
    xs, xb = _arr(x_segs), _arr(x_blobs)
    ns, nb = xs.size // 3, xb.size // 3
    fs = _arr(f_segs, 6 * ns).copy()
    rep = np.zeros(3 * nb)
    lib.orc_barrier_forces(_p(fs if ns else np.zeros(6)), _p(rep if nb else np.zeros(3)),
                           _p(xs if ns else np.zeros(3)), _p(xb if nb else np.zeros(3)), ns, nb, int(nseg),
                           float(a_seg), float(a_blob), float(strength), int(bool(wall)), float(dl))
    return fs, rep


def flow_rpy(x_flow, x_segs, f_segs, x_blobs, f_blobs, *, a_seg=1.0, a_blob=1.0, mu=1.0):
    """RPY-mode flow: point-tracer velocity from RPY sources."""
    lib = _load()
    xf = _arr(x_flow)
    nf = xf.size // 3
    xs, fs = _arr(x_segs), _arr(f_segs)
    xb, fb = _arr(x_blobs), _arr(f_blobs)
    v = np.zeros(3 * nf)
    lib.orc_flow_rpy(_p(v), _p(xf), _p(fs if fs.size else np.zeros(6)), _p(fb if fb.size else np.zeros(3)),
                     _p(xs if xs.size else np.zeros(3)), _p(xb if xb.size else np.zeros(3)), nf,
                     xs.size // 3, xb.size // 3, a_seg, a_blob, mu)
    return v
