"""The C-ABI called from C++ (tests/cpp/abi_consumer.cpp) in the reference's
evaluate_full_mobility sequence, on pinned host arrays of the build's Real
(double, and float as the reference ships, config.hpp:36,52), matches the FP64
oracle: at <= 1e-12 with double arrays; with float arrays the oracle runs on
the float-rounded inputs and the bar is the float rounding of the outputs
(prescribed and elastic-filament builds)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import geometry as G

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "abi_consumer")


@pytest.mark.parametrize("real", ["f64", "f32"])
@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (3, True)])
def test_cpp_consumer_matches_oracle(tmp_path, seg_dof, bs_acc, real):
    assert os.path.exists(EXE), "build first: python -m timfilsimmod_amd.build"
    b = G.small_body()
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        f.write(struct.pack("4i", b.num_segs, b.num_blobs, seg_dof, int(bs_acc)))
        for a in (b.x_segs, b.x_blobs, b.f_segs, b.f_blobs):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    out = tmp_path / "out.bin"
    res = subprocess.run([EXE, str(case), str(out)] + (["f32"] if real == "f32" else []), capture_output=True,
                         text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    v = np.fromfile(out, dtype=np.float32 if real == "f32" else np.float64).astype(np.float64)
    # the inputs as the build holds them
    cast = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if real == "f32" else (lambda a: a)
    xs, xb, fs, fb = (cast(a) for a in (b.x_segs, b.x_blobs, b.f_segs, b.f_blobs))
    rs = O.block("ss", xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rs = O.block("sb", xs, xb, fs, fb, rs, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rb = O.block("bb", xs, xb, fs, fb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rb = O.block("bs", xs, xb, fs, fb, rb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    ref = np.concatenate([rs, rb])
    err = np.linalg.norm(v - ref) / np.linalg.norm(ref)
    print(f"{real}: relative L2 error against the oracle {err:.2e}")
    if real == "f32":
        # every output is the FP64 result rounded once to float: within half an
        # ulp of float (2^-24 relative) plus the FP64 evaluation's own error
        assert np.all(np.abs(v - ref) <= 2.0 ** -24 * np.abs(ref) + 1e-12 * np.abs(ref).max())
        assert err <= 1e-7
    else:
        assert err <= 1e-12
