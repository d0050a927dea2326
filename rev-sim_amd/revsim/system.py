"""Water systems, parameter vectors and the named workloads.

The reference builds its system from a PDB + OpenMM XML through Molly (water/water.jl:980-1007,
atom order O,H,H per residue, masses :43-44, parameters :1567-1572). Here the system is built
directly: oxygens on a cubic lattice, each molecule rigidly rotated by a seeded uniform-random
rotation, Maxwell-Boltzmann velocities (the role of Molly's random_velocity, water/water.jl:1593).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

KB = 0.00831446262  # kJ mol^-1 K^-1
MASS_O = 15.99943
MASS_H = 1.007947
R_OH = 0.09572
THETA0 = 1.82421813418  # 104.52 degrees

# θ = (σ_O, ε_O, q_H, k_bond, r0_bond, k_angle)
THETA_NAMES = ("sigma_O", "eps_O", "q_H", "k_bond", "r0_bond", "k_angle")
DEFAULT_THETA = np.array([0.315, 0.636, 0.417, 462750.4, 0.09572, 836.8], dtype=np.float64)

SEED_GEOMETRY = 11
SEED_VELOCITY = 12
SEED_NOISE = 13
SEED_TARGET = 14
SEED_MLP = 15


@dataclasses.dataclass
class WaterBox:
    coords: np.ndarray  # (3N,) float64, xyz per atom, O,H1,H2 per molecule
    velocities: np.ndarray  # (3N,) float64
    box: np.ndarray  # (3,) float64

    @property
    def n_molecules(self) -> int:
        return self.coords.size // 9

    @property
    def n_atoms(self) -> int:
        return self.coords.size // 3

    def masses(self) -> np.ndarray:
        m = np.empty(self.n_atoms)
        m[0::3] = MASS_O
        m[1::3] = MASS_H
        m[2::3] = MASS_H
        return m


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(n_molecules: int, box: float, seed_geometry: int = SEED_GEOMETRY,
              seed_velocity: int = SEED_VELOCITY, temperature: float = 300.0) -> WaterBox:
    """n_molecules waters with O on the sites of an n_side^3 cubic lattice (spacing box/n_side,
    sites at (i+1/2)*spacing); when n_side^3 > n_molecules a seeded random subset of sites is
    kept. Each molecule gets a uniform-random orientation; velocities are Maxwell-Boltzmann."""
    n_side = int(math.ceil(round(n_molecules ** (1.0 / 3.0), 9)))
    while n_side ** 3 < n_molecules:
        n_side += 1
    a = box / n_side
    rng = np.random.Generator(np.random.Philox(seed_geometry))
    idx = np.arange(n_side ** 3)
    if n_side ** 3 > n_molecules:
        idx = np.sort(rng.permutation(n_side ** 3)[:n_molecules])
    i, j, k = idx // (n_side * n_side), (idx // n_side) % n_side, idx % n_side
    O = (np.stack([i, j, k], axis=1) + 0.5) * a
    half = THETA0 / 2.0
    ref = np.array([[0.0, 0.0, 0.0],
                    [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                    [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    R = _random_rotations(rng, n_molecules)
    pos = O[:, None, :] + np.einsum("mij,aj->mai", R, ref)
    pos = np.mod(pos, box)
    coords = np.ascontiguousarray(pos.reshape(-1), dtype=np.float64)
    m = np.tile([MASS_O, MASS_H, MASS_H], n_molecules)
    vrng = np.random.Generator(np.random.Philox(seed_velocity))
    v = vrng.standard_normal((3 * n_molecules, 3)) * np.sqrt(KB * temperature / m)[:, None]
    return WaterBox(coords, np.ascontiguousarray(v.reshape(-1)), np.array([box] * 3, dtype=np.float64))


# Named workloads (water box, cutoff, steps) — the five water cases this library is measured on.
WORKLOADS = {
    "water216": dict(n_molecules=216, box=1.862, cutoff=0.8, n_steps=200),
    "water3000": dict(n_molecules=3000, box=4.48, cutoff=1.0, n_steps=1000),
    "water32768": dict(n_molecules=32768, box=9.95, cutoff=1.0, n_steps=1000,
                       msd_window_steps=500, msd_every=50, w_msd=1.0),
    "water3000_mlp": dict(n_molecules=3000, box=4.48, cutoff=1.0, n_steps=500, use_mlp=1),
    "water32768_mlp": dict(n_molecules=32768, box=9.95, cutoff=1.0, n_steps=5000,
                           msd_window_steps=500, msd_every=50, w_msd=1.0, use_mlp=1),
}


MLP_NF, MLP_NH = 16, 64
MLP_SIZES = (("W1", (MLP_NH, MLP_NF)), ("b1", (MLP_NH,)), ("W2", (MLP_NH, MLP_NH)), ("b2", (MLP_NH,)),
             ("w3", (MLP_NH,)), ("b3", (1,)))
MLP_NW = sum(int(np.prod(s)) for _, s in MLP_SIZES)


def mlp_init(seed: int = SEED_MLP) -> np.ndarray:
    """Xavier-uniform weights (zero biases) of the 16 -> 64 -> 64 -> 1 SiLU network, packed as
    W1 (row-major 64x16), b1, W2 (64x64), b2, w3, b3 — the layout rs_set_mlp takes."""
    rng = np.random.Generator(np.random.Philox(seed))
    parts = []
    for name, shape in MLP_SIZES:
        if name.startswith("b"):
            parts.append(np.zeros(shape))
        else:
            fan_out, fan_in = (shape if len(shape) == 2 else (1, shape[0]))
            a = np.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-a, a, size=shape))
    return np.ascontiguousarray(np.concatenate([p.reshape(-1) for p in parts]), dtype=np.float32)


def mlp_unpack(w: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in MLP_SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + n]).reshape(shape)
        o += n
    return out


def perturbed_target(g: np.ndarray, seed: int = SEED_TARGET, rel: float = 0.05) -> np.ndarray:
    """Synthetic target RDF: g * (1 + rel * N(0,1)) per bin with a seeded Philox stream."""
    rng = np.random.Generator(np.random.Philox(seed))
    return np.ascontiguousarray(g * (1.0 + rel * rng.standard_normal(g.shape)))
