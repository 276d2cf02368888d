"""Pins the oracle's flow-field kernel (oracle/rpy_oracle.c, restating
src/flow_field/flow_field_evaluator.cu:13-228) against closed forms it does
not share code with: a single Stokeslet and rotlet; the no-slip condition of
Blake's image system on the wall z = 0; the 1/r^2 decay left after removing the
net-force Stokeslet; and the far-field approach of RPY tracer velocities to the
Stokeslet, with the Faxen correction a^2/6 Lap G."""
import numpy as np
import pytest

from oracle import oracle as O

MU = 0.7


def one_source(f6, x_src, pts, nb_sources=None, **kw):
    return O.flow_velocities(np.asarray(pts, float).reshape(-1), np.asarray(x_src, float),
                             np.asarray(f6, float), np.zeros(3) if nb_sources is None else nb_sources[0],
                             np.zeros(3) if nb_sources is None else nb_sources[1],
                             np.zeros(3), num_bodies=1, mu=MU, **kw).reshape(-1, 3)


def test_single_stokeslet_and_rotlet_closed_forms():
    r = 3.0
    # Stokeslet along x, seen on the x axis and on the y axis
    # one zero-force blob far away keeps Nb > 0 (no Blakelet images)
    v = O.flow_velocities(np.array([r, 0, 0, 0, r, 0.0]), np.zeros(3), np.array([1.0, 0, 0, 0, 0, 0]),
                          np.array([1e6, 0, 0]), np.zeros(3), np.zeros(3), num_bodies=1, mu=MU).reshape(-1, 3)
    c = 1.0 / (8 * np.pi * MU)
    np.testing.assert_allclose(v[0], [2 * c / r, 0, 0], atol=1e-15)
    np.testing.assert_allclose(v[1], [c / r, 0, 0], atol=1e-15)
    # rotlet: torque along z seen on the x axis: u = tau x r / (8 pi mu r^3)
    v = O.flow_velocities(np.array([r, 0, 0.0]), np.zeros(3), np.array([0, 0, 0, 0, 0, 1.0]),
                          np.array([1e6, 0, 0]), np.zeros(3), np.zeros(3), num_bodies=1, mu=MU)
    np.testing.assert_allclose(v, [0, c / r**2, 0], atol=1e-15)


def test_blakelet_satisfies_no_slip_on_the_wall():
    rng = np.random.default_rng(1)
    xs = np.column_stack([rng.uniform(-3, 3, 7), rng.uniform(-3, 3, 7), rng.uniform(0.5, 4, 7)])
    fs = np.zeros((7, 6))
    fs[:, :3] = rng.normal(size=(7, 3))
    wall = np.column_stack([rng.uniform(-6, 6, 20), rng.uniform(-6, 6, 20), np.zeros(20)])
    v = O.flow_velocities(wall.reshape(-1), xs.reshape(-1), fs.reshape(-1), np.zeros(0), np.zeros(0),
                          np.zeros(3), num_bodies=1, mu=MU)
    free = O.flow_velocities(wall.reshape(-1), xs.reshape(-1), fs.reshape(-1), np.array([1e9, 0, 0]),
                             np.zeros(3), np.zeros(3), num_bodies=1, mu=MU)
    assert np.abs(v).max() < 1e-12 * np.abs(free).max()
    # above the wall the image system stays finite
    above = O.flow_velocities(np.array([0.0, 0.0, 2.0]), xs.reshape(-1), fs.reshape(-1), np.zeros(0),
                              np.zeros(0), np.zeros(3), num_bodies=1, mu=MU)
    assert np.all(np.isfinite(above))


def test_net_force_removal_leaves_a_dipole_decay():
    rng = np.random.default_rng(2)
    xs = rng.uniform(-1, 1, 3 * 10)
    fs = rng.normal(size=60)
    xb = rng.uniform(-1, 1, 3 * 6)
    fb = rng.normal(size=18)
    far = lambda R: np.array([R, 0.3 * R, -0.2 * R])
    speeds = []
    for R in (100.0, 200.0):
        v0 = O.flow_velocities(far(R), xs, fs, xb, fb, np.zeros(3), num_bodies=1, mu=MU)
        v1 = O.flow_velocities(far(R), xs, fs, xb, fb, np.zeros(3), num_bodies=1, mu=MU, remove_net_force=True)
        speeds.append((np.linalg.norm(v0), np.linalg.norm(v1)))
    assert speeds[0][0] / speeds[1][0] == pytest.approx(2.0, rel=0.05)    # Stokeslet ~ 1/r
    assert speeds[0][1] / speeds[1][1] == pytest.approx(4.0, rel=0.05)    # dipole ~ 1/r^2


def test_rpy_tracer_tends_to_faxen_corrected_stokeslet():
    a = 1.3
    f = np.array([0.3, -1.0, 0.5, 0, 0, 0])
    for R in (5.0, 20.0):
        x = np.array([R, 0.5 * R, 0.2 * R])
        r = np.linalg.norm(x)
        G = (np.eye(3) / r + np.outer(x, x) / r**3) / (8 * np.pi * MU)
        lapG = (2 * np.eye(3) / r**3 - 6 * np.outer(x, x) / r**5) / (8 * np.pi * MU)
        v = O.flow_rpy(x, np.zeros(3), f, np.zeros(0), np.zeros(0), a_seg=a, mu=MU)
        np.testing.assert_allclose(v, (G + a * a / 6 * lapG) @ f[:3], rtol=1e-13)
    # inside the source sphere: rigid-body velocity F/(6 pi mu a)
    v = O.flow_rpy(np.array([0.2, 0.1, 0.0]), np.zeros(3), f, np.zeros(0), np.zeros(0), a_seg=a, mu=MU)
    np.testing.assert_allclose(v, f[:3] / (6 * np.pi * MU * a), rtol=1e-14)
