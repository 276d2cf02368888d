"""The data-file layer (include/tfs_io.h) on the CPU: byte-exact writers,
strict readers and the state -> source-position reconstruction of the
flow-field program (flow_field_evaluator.cu:401-498).

Pins: the writers' bytes against a std::ofstream program that configures its
streams as the reference's writers do (tests/cpp/iostream_records.cpp); the
rotation matrix against scipy's Rotation; the segment chain against closed
forms (a straight filament, a planar arc) and a numpy restatement."""
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from timfilsimmod_amd import _lib as L
from timfilsimmod_amd import datafiles as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

EDGE = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 / 3.0, 123456.789, -9.87654321e-7, 1e-300,
                 5e-324, 1.7976931348623157e308, 1e22, 1e21, 1e16, 999999.5, 0.5e-4, 2.5e-5,
                 np.pi, -np.e, 1e100, 12345678901.0, 0.00012345678905, 9.99999999995e-5,
                 np.inf, -np.inf, np.nan])


@pytest.fixture(scope="module")
def iostream_tool(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("tool") / "iostream_records")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "iostream_records.cpp"),
                    "-o", out], check=True)
    return out


def _values():
    rng = np.random.default_rng(7)
    rand = rng.standard_normal(200) * 10.0 ** rng.integers(-12, 12, 200)
    return np.concatenate([EDGE, rand])


def test_record_bytes_match_iostream(tmp_path, iostream_tool):
    v = _values()
    v.tofile(tmp_path / "v.bin")
    subprocess.run([iostream_tool, str(tmp_path / "v.bin"), "37", str(tmp_path / "want.dat"),
                    str(tmp_path / "want_refs.dat"), str(tmp_path / "want.par")], check=True)
    D.write_record(tmp_path / "got.dat", v, step=37)
    D.write_record(tmp_path / "got_refs.dat", v, append=False, default_float=True)
    assert (tmp_path / "got.dat").read_bytes() == (tmp_path / "want.dat").read_bytes()
    assert (tmp_path / "got_refs.dat").read_bytes() == (tmp_path / "want_refs.dat").read_bytes()
    params = [("NFIL", 640), ("TOTAL_TIME_STEPS", 37)] + [(f"P{i}", float(x)) for i, x in enumerate(v)]
    D.write_par(tmp_path / "got.par", params)
    assert (tmp_path / "got.par").read_bytes() == (tmp_path / "want.par").read_bytes()


def test_record_format_matches_printf():
    # "%.10e" / "%g" are the printf conversions the stream settings select.
    import tempfile
    v = _values()[:-3]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "r.dat")
        D.write_record(p, v, step=5)
        D.write_record(p, v, step=6)
        text = open(p).read()
    line = "5 " + "".join("%.10e " % x for x in v) + "\n"
    assert text == line + line.replace("5 ", "6 ", 1)


def test_append_overwrite_and_read_back(tmp_path):
    p = tmp_path / "a.dat"
    rng = np.random.default_rng(1)
    recs = [rng.standard_normal(12) for _ in range(4)]
    for k, r in enumerate(recs):
        D.write_record(p, r, step=10 * k)
    steps, vals = D.read_records(p, 12)
    assert list(steps) == [0, 10, 20, 30]
    # 11 significant digits survive the text format
    want = np.array([[float("%.10e" % x) for x in r] for r in recs])
    assert np.array_equal(vals, want)
    np.testing.assert_allclose(vals, np.stack(recs), rtol=6e-11, atol=0)
    D.write_record(p, recs[0], step=99, append=False)
    assert D.count_records(p) == 1


def test_default_float_round_trip(tmp_path):
    v = np.array([1.5, -2.25, 1e-7, 123456.0, 1234567.0, 0.1])
    D.write_record(tmp_path / "r.dat", v, append=False, default_float=True)
    assert (tmp_path / "r.dat").read_text() == "1.5 -2.25 1e-07 123456 1.23457e+06 0.1 \n"
    _, back = D.read_records(tmp_path / "r.dat", 6, with_step=False)
    assert np.array_equal(back[0], [float("%g" % x) for x in v])


def test_reader_peek_rewind_blank_lines(tmp_path):
    p = tmp_path / "r.dat"
    p.write_text("\n3 1.0 2.0 \n\n   \n4 3.0 4.0 \n")
    with D.RecordReader(p) as r:
        assert r.peek_count() == 3
        assert r.peek_count() == 3
        s, v = r.read(2)
        assert s == 3 and list(v) == [1.0, 2.0]
        s, v = r.read(2)
        assert s == 4 and list(v) == [3.0, 4.0]
        assert r.read(2) is None and r.peek_count() is None
        r.rewind()
        assert r.read(2)[0] == 3
    assert D.count_records(p) == 2


def test_reader_parses_what_fscanf_reads(tmp_path):
    p = tmp_path / "r.dat"
    p.write_text("+7 1 -2 3.5e2 +4E-1 .5 nan -inf 1e400 -1e-400 4.9e-324 \n")
    with D.RecordReader(p) as r:
        s, v = r.read(10)
    assert s == 7
    assert list(v[:5]) == [1.0, -2.0, 350.0, 0.4, 0.5] and np.isnan(v[5]) and v[6] == -np.inf
    assert v[7] == np.inf and v[8] == 0.0 and v[9] == 5e-324


@pytest.mark.parametrize("text,count,why", [
    ("1 1.0 2.0 \n", 3, "holds 2 values"),
    ("1 1.0 2.0 3.0 \n", 2, "more than"),
    ("1 1.0 x2 \n", 2, "not a number"),
    ("1.5 1.0 2.0 \n", 2, "step"),
])
def test_reader_rejects_malformed_records(tmp_path, text, count, why):
    p = tmp_path / "r.dat"
    p.write_text(text)
    with D.RecordReader(p) as r:
        with pytest.raises(L.TfsError) as e:
            r.read(count)
    assert e.value.code == L.TFS_ERR_FORMAT and why in str(e.value)


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(L.TfsError) as e:
        D.RecordReader(tmp_path / "absent.dat")
    assert e.value.code == L.TFS_ERR_IO
    with pytest.raises(L.TfsError) as e:
        D.read_par(tmp_path / "absent.par")
    assert e.value.code == L.TFS_ERR_IO


def test_par_round_trip_and_layout(tmp_path):
    params = [("NFIL", 640), ("NSEG", 20), ("NSWIM", 1), ("NBLOB", 9000), ("MU", 1.0), ("RSEG", 1.0),
              ("DL", 2.2), ("FIL_LENGTH", 41.8), ("RBLOB", 1.0), ("KB", 1800.0), ("KT", 1800.0),
              ("TOL", 1e-4), ("TOTAL_TIME_STEPS", 3000000), ("DT", 1.0 / 300.0),
              ("PLOT_FREQUENCY_IN_STEPS", 10), ("STEPS_PER_PERIOD", 300), ("PRESCRIBED_CILIA", 1)]
    p = tmp_path / "run.par"
    D.write_par(p, params)
    lines = p.read_text().splitlines()
    assert lines[0] == "640 %% NFIL" and lines[4] == "1 %% MU" and lines[11] == "0.0001 %% TOL"
    assert lines[12] == "3000000 %% TOTAL_TIME_STEPS" and lines[13] == "0.00333333 %% DT"
    back = D.read_par(p)
    assert list(back) == [n for n, _ in params]
    assert back["TOTAL_TIME_STEPS"] == 3000000 and isinstance(back["NFIL"], int)
    assert back["DT"] == float("%g" % (1.0 / 300.0))
    (tmp_path / "bad.par").write_text("640 NFIL\n")
    with pytest.raises(L.TfsError) as e:
        D.read_par(tmp_path / "bad.par")
    assert e.value.code == L.TFS_ERR_FORMAT
    with pytest.raises(ValueError):
        D.write_par(tmp_path / "x.par", [("A", 1), ("", 1.0)])


def _unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _R(q):
    # scipy takes (x, y, z, w); the files hold (w, x, y, z).
    return Rotation.from_quat(np.roll(q, -1, axis=-1)).as_matrix()


def test_blob_and_base_positions_match_scipy_rotation():
    rng = np.random.default_rng(3)
    nbod, nfil, nseg, nblob = 3, 4, 5, 7
    qb = _unit_quats(rng, nbod)
    X = rng.standard_normal((nbod, 3)) * 20
    body = np.concatenate([X, qb], axis=1)
    fil_refs = rng.standard_normal((nfil, 3)) * 10
    blob_refs = rng.standard_normal((nblob, 3)) * 10
    qs = _unit_quats(rng, nbod * nfil * nseg)
    xs, xb, xbod = D.sources_from_states(body, qs, fil_refs, blob_refs, num_fil=nfil, num_seg=nseg, dl=2.2)
    R = _R(qb)
    want_b = X[:, None, :] + np.einsum("bij,kj->bki", R, blob_refs)
    np.testing.assert_allclose(xb.reshape(nbod, nblob, 3), want_b, rtol=0, atol=1e-12)
    np.testing.assert_allclose(xbod.reshape(nbod, 3), X, rtol=0, atol=0)
    base = X[:, None, :] + np.einsum("bij,kj->bki", R, fil_refs)
    np.testing.assert_allclose(xs.reshape(nbod, nfil, nseg, 3)[:, :, 0], base, rtol=0, atol=1e-12)
    # the chain: tangent = first column of R(q_seg), trapezoidal steps of dl
    t = _R(qs)[:, :, 0].reshape(nbod, nfil, nseg, 3)
    steps = 0.5 * 2.2 * (t[:, :, :-1] + t[:, :, 1:])
    chain = base[:, :, None, :] + np.concatenate([np.zeros((nbod, nfil, 1, 3)), np.cumsum(steps, axis=2)], axis=2)
    np.testing.assert_allclose(xs.reshape(nbod, nfil, nseg, 3), chain, rtol=0, atol=1e-11)


def test_straight_and_arc_filaments_closed_form():
    nseg, dl = 9, 2.2
    body = np.array([[1.0, 2.0, 3.0, 1.0, 0.0, 0.0, 0.0]])
    # straight: identity quaternions -> segments along e_x at spacing dl
    xs, _, _ = D.sources_from_states(body, np.tile([1.0, 0, 0, 0], nseg), [0.0, 0.0, 5.0], None,
                                     num_fil=1, num_seg=nseg, dl=dl)
    want = np.array([1.0, 2.0, 8.0]) + np.outer(np.arange(nseg) * dl, [1.0, 0, 0])
    np.testing.assert_allclose(xs.reshape(nseg, 3), want, atol=1e-13)
    # planar arc: segment k turned by k*a about e_z, tangent (cos ka, sin ka, 0);
    # the trapezoid sum has the closed form dl/2 (S_k-1 + S_k) with S_k the
    # partial sums of cos/sin over 0..k.
    a = 0.3
    ang = np.arange(nseg) * a
    q = np.stack([np.cos(ang / 2), 0 * ang, 0 * ang, np.sin(ang / 2)], axis=1)
    xs, _, _ = D.sources_from_states(body, q, [0.0, 0.0, 0.0], None, num_fil=1, num_seg=nseg, dl=dl)
    k = np.arange(nseg)
    c = np.array([np.sum(np.cos(ang[:j])) + np.sum(np.cos(ang[1:j + 1])) for j in k]) * dl / 2
    s = np.array([np.sum(np.sin(ang[:j])) + np.sum(np.sin(ang[1:j + 1])) for j in k]) * dl / 2
    np.testing.assert_allclose(xs.reshape(nseg, 3), np.stack([1 + c, 2 + s, 3 + 0 * c], axis=1), atol=1e-12)


def test_position_based_states_are_copied():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(2 * 3 * 4 * 3)
    body = np.array([[0, 0, 0, 1.0, 0, 0, 0], [5, 5, 5, 0.0, 1.0, 0, 0]])
    xs, _, _ = D.sources_from_states(body, x, np.zeros(9), None, num_fil=3, num_seg=4, dl=2.2,
                                     quaternion_based=False)
    assert np.array_equal(xs, x)


def test_flow_parameter_file_layout(tmp_path):
    p = tmp_path / "flow_field_simulation_parameters.input"
    D.write_flow_parameters(p, "run", 1000, periodic=True, cycles=3, remove_net_force=True)
    assert p.read_text().split() == ["SIMULATION_NAME", "run", "NFLOW", "1000", "QUATERNION_BASED_DATA", "1",
                                     "TREAT_DATA_AS_PERIODIC", "1", "NUMBER_OF_CYCLES_FOR_PERIODIC_DATA", "3",
                                     "REMOVE_NET_FORCE_STOKESLET", "1", "ADVANCE_TRACER_POSITIONS", "0"]


def test_flow_run_reports_missing_files_without_a_gpu(tmp_path):
    p = tmp_path / "in.input"
    D.write_flow_parameters(p, "absent", 10)
    with pytest.raises(L.TfsError) as e:
        D.flow_field_run(p)
    assert e.value.code == L.TFS_ERR_IO
