"""Prescribed-beat inputs: the Fulford & Blake beat table, filament frames,
and the reference generalised-force tables the dynamic-phase solve needs.

The beat is a Fourier fit in phase and a cubic in the shape parameter s
(src/cilia/filament.cpp:680-833, FULFORD_AND_BLAKE_BEAT):

    x(s, phi) = sum_{n=0..3} cos(n phi) Ax[n,p] s^(p+1) + sum_{n=1..3} sin(n phi) Bx[n-1,p] s^(p+1)

and likewise for y; x runs along the surface normal and y along the beat
direction.  The coefficients are published data (Ito, Omori & Ishikawa 2019,
Table 1, rescaled to unit mean length with the zero-mode halved, as the
reference transcribes them at filament.cpp:692-750) and are kept identical.

The reference reads its generalised-force tables from files written by a
separate WRITE_GENERALISED_FORCES run (mobility_solver.cpp:45-103); those files
are not part of the reference tree.  That run drives a filament at its natural
frequency with the shape angle held, and records the generalised forces of the
segment forces it needed, q_phi = K_phi . f and q_theta = K_theta . f
(cilia_sim_main.cu:370-380).  ``generalised_force_tables`` does the same for one
isolated filament in unbounded fluid: f = M^-1 K_phi * 2 pi with M the RPY
mobility of its segments, sampled over one period.  The two tables are then
consistent with each other: the filament alone beats at phi_dot = 2 pi with
theta_dot = 0 even where K_phi and K_theta are nearly parallel (the power
stroke, close to a rigid rotation about the base), where an unrelated angle
table would drive theta_dot without bound.
"""
from __future__ import annotations

import functools

import numpy as np

# [x|y][mode 0..3][power s, s^2, s^3]  (Ax, Ay of filament.cpp:692-736)
BEAT_A = np.array([
    [[9.7204e-01, -2.8315e-01, 4.9243e-02],
     [-1.8466e-02, -1.2926e-01, 2.6981e-01],
     [1.6209e-01, -3.4983e-01, 1.9082e-01],
     [1.0259e-02, 3.5907e-02, -6.8736e-02]],
    [[-3.3547e-01, 4.0369e-01, 1.0362e-01],
     [4.0318e-01, -1.5553e+00, 7.3455e-01],
     [-9.9513e-02, 3.2829e-02, -1.2106e-01],
     [8.1046e-02, -3.0982e-01, 1.4568e-01]],
])
# [x|y][mode 1..3][power]  (Bx, By of filament.cpp:724-750)
BEAT_B = np.array([
    [[1.9697e-01, -5.1193e-01, 3.4778e-01],
     [-5.1295e-02, 4.3396e-01, -3.3547e-01],
     [1.2311e-02, 1.4157e-01, -1.1695e-01]],
    [[2.9136e-01, 1.0721e+00, -1.0433e+00],
     [6.1554e-03, 3.2521e-01, -2.8315e-01],
     [-6.0528e-02, 2.3185e-01, -2.0108e-01]],
])


def shape(s, phase):
    """Local beat-plane shape (x, y) and its phase derivative at parameter s."""
    s = np.asarray(s, dtype=float)
    sv = np.stack([s, s * s, s * s * s], axis=-1)            # (..., 3)
    n = np.arange(4)
    cos_n, sin_n = np.cos(n * phase), np.sin(n * phase)
    pos = np.empty(s.shape + (2,))
    der = np.empty(s.shape + (2,))
    for ax in range(2):
        coef = cos_n @ BEAT_A[ax] + sin_n[1:] @ BEAT_B[ax]          # (3,)
        dcoef = (-n * sin_n) @ BEAT_A[ax] + (n[1:] * cos_n[1:]) @ BEAT_B[ax]
        pos[..., ax] = sv @ coef
        der[..., ax] = sv @ dcoef
    return pos, der


def frames(normals: np.ndarray) -> np.ndarray:
    """Per-filament frames (9 numbers each): e1 = normal, e2 = polar direction
    (the beat direction), e3 = e1 x e2."""
    nrm = normals / np.linalg.norm(normals, axis=1, keepdims=True)
    z = np.array([0.0, 0.0, 1.0])
    e2 = z[None, :] - (nrm @ z)[:, None] * nrm       # projection of e_z on the tangent plane
    e2 = -e2                                          # point towards the south pole (polar direction)
    bad = np.linalg.norm(e2, axis=1) < 1e-12
    e2[bad] = np.array([1.0, 0.0, 0.0])
    e2 /= np.linalg.norm(e2, axis=1, keepdims=True)
    e3 = np.cross(nrm, e2)
    return np.ascontiguousarray(np.concatenate([nrm, e2, e3], axis=1).reshape(-1))


def uniform_s(num_seg: int) -> np.ndarray:
    return np.linspace(0.0, 1.0, num_seg)


def rpy_translational(x: np.ndarray, a: float = 1.0, mu: float = 1.0) -> np.ndarray:
    """Dense translational RPY mobility of equal spheres at x (n x 3): far
    field for r > 2a, the Rotne-Prager overlap form below, 1/(6 pi mu a) on
    the diagonal."""
    n = len(x)
    M = np.zeros((3 * n, 3 * n))
    I3 = np.eye(3)
    for i in range(n):
        M[3 * i:3 * i + 3, 3 * i:3 * i + 3] = I3 / (6.0 * np.pi * mu * a)
        for j in range(i + 1, n):
            d = x[i] - x[j]
            r = np.linalg.norm(d)
            e = np.outer(d, d) / (r * r)
            if r > 2.0 * a:
                c = 1.0 / (8.0 * np.pi * mu * r)
                blk = c * ((1.0 + 2.0 * a * a / (3.0 * r * r)) * I3 + (1.0 - 2.0 * a * a / (r * r)) * e)
            else:
                c = 1.0 / (6.0 * np.pi * mu * a)
                blk = c * ((1.0 - 9.0 * r / (32.0 * a)) * I3 + (3.0 * r / (32.0 * a)) * e)
            M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
            M[3 * j:3 * j + 3, 3 * i:3 * i + 3] = blk
    return M


def generalised_force_tables(num_seg: int, fil_length: float, *, a: float = 1.0, mu: float = 1.0,
                             samples: int = 300):
    """(phase table, angle table) over one period, see the module docstring:
    an isolated filament (base at the origin, beat in the local x-y plane)
    driven at phi_dot = 2 pi, theta_dot = 0."""
    qp, qa = _tables(int(num_seg), float(fil_length), float(a), float(mu), int(samples))
    return qp.copy(), qa.copy()


@functools.lru_cache(maxsize=8)
def _tables(num_seg: int, fil_length: float, a: float, mu: float, samples: int):
    s = uniform_s(num_seg)
    q_phase = np.empty(samples)
    q_angle = np.empty(samples)
    for k in range(samples):
        phase = 2.0 * np.pi * k / samples
        pos, der = shape(s, phase)
        x = np.zeros((num_seg, 3))
        x[:, :2] = fil_length * pos
        kphi = np.zeros((num_seg, 3))
        kphi[1:, :2] = fil_length * der[1:]
        ktheta = np.zeros((num_seg, 3))
        ktheta[1:, 0] = -x[1:, 1]                 # e_z x ref
        ktheta[1:, 1] = x[1:, 0]
        f = np.linalg.solve(rpy_translational(x, a, mu), 2.0 * np.pi * kphi.reshape(-1))
        q_phase[k] = kphi.reshape(-1) @ f
        q_angle[k] = ktheta.reshape(-1) @ f
    return q_phase, q_angle
