"""The C-ABI called from C++ (tests/cpp/abi_consumer.cpp) in the reference's
evaluate_full_mobility sequence, on pinned host arrays, matches the FP64
oracle at <= 1e-12 (prescribed and elastic-filament builds)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from timfilsimmod_amd import geometry as G

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "abi_consumer")


@pytest.mark.parametrize("seg_dof,bs_acc", [(6, False), (3, True)])
def test_cpp_consumer_matches_oracle(tmp_path, seg_dof, bs_acc):
    assert os.path.exists(EXE), "build first: python -m timfilsimmod_amd.build"
    b = G.small_body()
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        f.write(struct.pack("4i", b.num_segs, b.num_blobs, seg_dof, int(bs_acc)))
        for a in (b.x_segs, b.x_blobs, b.f_segs, b.f_blobs):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    out = tmp_path / "out.bin"
    res = subprocess.run([EXE, str(case), str(out)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    v = np.fromfile(out, dtype=np.float64)
    vs, vb = v[:6 * b.num_segs], v[6 * b.num_segs:]
    rs = O.block("ss", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rs = O.block("sb", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, rs, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rb = O.block("bb", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, seg_dof=seg_dof, bs_accumulate=bs_acc)
    rb = O.block("bs", b.x_segs, b.x_blobs, b.f_segs, b.f_blobs, rb, seg_dof=seg_dof, bs_accumulate=bs_acc)
    ref = np.concatenate([rs, rb])
    assert np.linalg.norm(v - ref) <= 1e-12 * np.linalg.norm(ref)
