"""An ensemble of independent bodies of uneven size, concatenated in the
reference's array layouts, for one block-diagonal (batched) launch.

Body b has uniform-random filament and blob counts (geometry.ensemble_sizes),
its own seeds (forces 1000+b, perturbations 2000+b) and its own centre, so no
two bodies overlap.  The reference couples all NSWIM bodies hydrodynamically
(cuda_functions.cu:399, 504, 580, 675); independent bodies are an addition.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import geometry as G


@dataclass
class Ensemble:
    bodies: list
    segs_per_body: np.ndarray
    blobs_per_body: np.ndarray
    x_segs: np.ndarray
    x_blobs: np.ndarray
    f_segs: np.ndarray
    f_blobs: np.ndarray

    @property
    def num_segs(self) -> int:
        return int(self.segs_per_body.sum())

    @property
    def num_blobs(self) -> int:
        return int(self.blobs_per_body.sum())

    def pairs_per_body(self) -> np.ndarray:
        n = self.segs_per_body + self.blobs_per_body
        return n.astype(np.float64) ** 2


def make_ensemble(num_bodies: int = 64, num_seg: int = 20, seed: int = 7, sizes=None) -> Ensemble:
    sizes = sizes if sizes is not None else G.ensemble_sizes(num_bodies, seed)
    bodies = []
    for b, (nfil, nblob, radius) in enumerate(sizes):
        centre = (400.0 * (b % 8), 400.0 * ((b // 8) % 8), 400.0 * (b // 64))
        bodies.append(G.make_body(nfil, num_seg, nblob, radius, force_seed=1000 + b, perturb_seed=2000 + b,
                                  centre=centre))
    cat = lambda k: np.ascontiguousarray(np.concatenate([getattr(x, k) for x in bodies]))
    return Ensemble(bodies=bodies,
                    segs_per_body=np.array([x.num_segs for x in bodies], dtype=np.int32),
                    blobs_per_body=np.array([x.num_blobs for x in bodies], dtype=np.int32),
                    x_segs=cat("x_segs"), x_blobs=cat("x_blobs"), f_segs=cat("f_segs"), f_blobs=cat("f_blobs"))
