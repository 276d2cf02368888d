"""Deterministic synthetic bodies: Fibonacci-sphere blobs plus straight filaments.

The reference seeds blobs and filaments with an equal-area MacQueen N-means
iteration on a random surface sample, seeded from ``std::random_device``
(src/cilia/seeding.cu:547-725, 160-162), so its geometry cannot be reproduced.
This module is our own deterministic generator for the same kind of body:

* blobs of radius ``a`` at the points of a Fibonacci sphere of radius ``R``;
* filament bases at a second Fibonacci set on the same sphere, lifted by
  ``a + 0.5*DL`` along the normal (the reference's RBLOB +
  BASE_HEIGHT_ABOVE_SURFACE, src/cilia/swimmer.cpp:384-388, config.hpp:341);
* segments straight along the local normal at spacing DL, perturbed by
  N(0, sigma^2) per component;
* forces and torques N(0,1) per component.

Random numbers come from splitmix64 with Box-Muller, written here, so the
arrays are identical on every machine and numpy version.  Layouts are the
reference's AoS host layouts (x_segs[3*Ns], f_segs[6*Ns], x_blobs[3*Nb], ...).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First ``n`` outputs of splitmix64 started at ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = state
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_open(seed: int, n: int) -> np.ndarray:
    """``n`` doubles uniform on the open interval (0, 1)."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    """``n`` standard normal deviates by Box-Muller on splitmix64 uniforms."""
    m = (n + 1) // 2
    u = uniform_open(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2.0 * np.pi * u2)
    out[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return out[:n]


def fibonacci_sphere(n: int, radius: float) -> np.ndarray:
    """``n`` points on a sphere: z_k = 1-(2k+1)/n, phi_k = k*pi*(3-sqrt(5))."""
    k = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / n
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return radius * np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


@dataclass
class Body:
    """One rigid body with its filaments, in the reference's AoS layouts."""

    num_fil: int
    num_seg: int
    num_blobs: int
    radius: float
    a: float
    dl: float
    centre: np.ndarray        # (3,)
    x_segs: np.ndarray        # (3*Ns,)  body -> filament -> segment
    x_blobs: np.ndarray       # (3*Nb,)
    f_segs: np.ndarray        # (6*Ns,)  fx fy fz tx ty tz
    f_blobs: np.ndarray       # (3*Nb,)
    fil_base: np.ndarray      # (Nfil, 3)
    fil_normal: np.ndarray    # (Nfil, 3)
    blob_refs: np.ndarray     # (Nb, 3) blob positions relative to the centre

    @property
    def num_segs(self) -> int:
        return self.num_fil * self.num_seg


def make_body(num_fil: int, num_seg: int, num_blobs: int, radius: float, *, a: float = 1.0,
              dl: float | None = None, sigma: float = 0.05, force_seed: int = 0,
              perturb_seed: int = 1, centre=(0.0, 0.0, 0.0)) -> Body:
    """Build one seeded body.

    ``dl`` defaults to 2.2*a.  ``force_seed`` seeds the N(0,1) forces/torques
    (segment forces first, then blob forces, from one stream);
    ``perturb_seed`` seeds the N(0, sigma*a) segment perturbations.
    """
    if dl is None:
        dl = 2.2 * a
    c = np.asarray(centre, dtype=np.float64)
    blobs = fibonacci_sphere(num_blobs, radius) + c if num_blobs > 0 else np.zeros((0, 3))
    normals = fibonacci_sphere(num_fil, 1.0) if num_fil > 0 else np.zeros((0, 3))
    bases = c + (radius + a + 0.5 * dl) * normals
    ns = num_fil * num_seg
    k = np.arange(num_seg, dtype=np.float64)
    segs = bases[:, None, :] + (k[None, :, None] * dl) * normals[:, None, :]
    segs = segs.reshape(ns, 3) + (sigma * a) * normal(perturb_seed, 3 * ns).reshape(ns, 3)
    fvals = normal(force_seed, 6 * ns + 3 * num_blobs)
    return Body(num_fil=num_fil, num_seg=num_seg, num_blobs=num_blobs, radius=radius, a=a, dl=dl,
                centre=c, x_segs=np.ascontiguousarray(segs.reshape(-1)),
                x_blobs=np.ascontiguousarray(blobs.reshape(-1)),
                f_segs=np.ascontiguousarray(fvals[: 6 * ns]),
                f_blobs=np.ascontiguousarray(fvals[6 * ns:]),
                fil_base=bases, fil_normal=normals, blob_refs=blobs - c)


# Named workloads, sizes as the benchmark configurations state them.
def small_body(**kw) -> Body:
    """2,000 blobs on R = 20a, 16 filaments x 20 segments."""
    return make_body(16, 20, 2000, 20.0, **kw)


def production_body(**kw) -> Body:
    """9,000 blobs on R = 50a, 640 filaments x 20 segments (12,800 segments)."""
    return make_body(640, 20, 9000, 50.0, **kw)


def flow_source_body(**kw) -> Body:
    """Flow-field sources: 9,000 blobs on R = 50a, 1,280 filaments x 20 segments."""
    return make_body(1280, 20, 9000, 50.0, **kw)


def ensemble_sizes(num_bodies: int = 64, seed: int = 7):
    """Per-body (num_fil, num_blobs, radius) for the uneven ensemble.

    Filament counts are uniform on [100, 1000] and blob counts uniform on
    [2000, 9000].  The radius keeps the area per blob of the production body
    (4*pi*50^2/9000), i.e. R = 50*sqrt(Nb/9000).
    """
    u = uniform_open(seed, 2 * num_bodies)
    nfil = 100 + np.floor(u[0::2] * 901).astype(int)
    nblob = 2000 + np.floor(u[1::2] * 7001).astype(int)
    radius = 50.0 * np.sqrt(nblob / 9000.0)
    return [(int(f), int(b), float(r)) for f, b, r in zip(nfil, nblob, radius)]


def plane_targets(n_side: int, half_width: float, centre=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n_side x n_side grid on the plane z = centre_z through the body centre,
    spanning [-half_width, half_width]^2 about the centre (AoS, 3*n_side^2)."""
    c = np.asarray(centre, dtype=np.float64)
    t = np.linspace(-half_width, half_width, n_side)
    gx, gy = np.meshgrid(t, t, indexing="xy")
    pts = np.stack([gx.ravel() + c[0], gy.ravel() + c[1], np.full(gx.size, c[2])], axis=1)
    return np.ascontiguousarray(pts.reshape(-1))


def volume_targets(n_side: int, half_width: float, centre=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n_side^3 grid over the cube [-half_width, half_width]^3 about the centre."""
    c = np.asarray(centre, dtype=np.float64)
    t = np.linspace(-half_width, half_width, n_side)
    gz, gy, gx = np.meshgrid(t, t, t, indexing="ij")
    pts = np.stack([gx.ravel() + c[0], gy.ravel() + c[1], gz.ravel() + c[2]], axis=1)
    return np.ascontiguousarray(pts.reshape(-1))
