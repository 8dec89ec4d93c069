"""Synthetic workloads of the benchmark configurations (seeded, no downloads).

* ``ar_fcc``: Ar fcc 4x4x4 conventional cells (256 atoms, a = 5.26 A).
* ``cu_batch``: periodic Cu fcc systems with random repeats n_k on {2..8}
  drawn with P(n) proportional to n (E[atoms] = 4 E[n]^3 ~ 780), positions
  jittered by N(0, sigma_jitter) -- 512 systems give ~4.0e5 atoms.
* ``lj_fluid``: fcc lattice at reduced density 0.8442 jittered by 0.05 sigma.
* ``soft_sphere_batch``: 50:50 binary soft spheres at packing fraction 0.64
  in random triclinic cells (strain <= 5 %).
* ``binary_lj_batch``: binary LJ liquids at reduced density 0.8.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .state import SimState

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])


def fcc_positions(a: float, reps: tuple[int, int, int]) -> np.ndarray:
    """Cartesian positions of an fcc crystal of nx x ny x nz cubic cells."""
    nx, ny, nz = reps
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 1, 3) + FCC_BASIS[None]
    return (g.reshape(-1, 3) * a).astype(np.float64)


def _state(pos_list, cells, masses, numbers, device, dtype=torch.float64) -> SimState:
    n = [len(p) for p in pos_list]
    pos = torch.tensor(np.concatenate(pos_list), dtype=dtype, device=device)
    cell = torch.tensor(np.stack(cells), dtype=dtype, device=device)  # column convention
    sys_idx = torch.repeat_interleave(torch.arange(len(n)), torch.tensor(n)).to(device)
    return SimState(
        positions=pos,
        masses=torch.tensor(np.concatenate(masses), dtype=dtype, device=device),
        cell=cell,
        pbc=True,
        atomic_numbers=torch.tensor(np.concatenate(numbers), dtype=torch.int64, device=device),
        system_idx=sys_idx,
    )


def ar_fcc(device="cuda", reps=(4, 4, 4), a=5.26, jitter=0.0, seed=0) -> SimState:
    """Argon fcc crystal (mass 39.948, Z = 18)."""
    rng = np.random.default_rng(seed)
    pos = fcc_positions(a, reps)
    if jitter:
        pos = pos + rng.normal(0.0, jitter, pos.shape)
    cell = np.diag([a * r for r in reps])
    n = len(pos)
    return _state([pos], [cell], [np.full(n, 39.948)], [np.full(n, 18)], device)


def cu_repeats(n_systems: int, seed: int = 0) -> np.ndarray:
    """Repeats (nx, ny, nz) i.i.d. on {2..8} with P(n) proportional to n."""
    rng = np.random.default_rng(seed)
    vals = np.arange(2, 9)
    prob = vals / vals.sum()
    return rng.choice(vals, size=(n_systems, 3), p=prob)


def cu_batch(
    n_systems: int = 512, device="cuda", a=3.61, jitter=0.05, seed=0, reps=None
) -> SimState:
    """Batch of jittered Cu fcc systems (mass 63.546, Z = 29)."""
    rng = np.random.default_rng(seed + 1)
    reps = cu_repeats(n_systems, seed) if reps is None else np.asarray(reps)
    pos, cells, ms, zs = [], [], [], []
    for r in reps:
        p = fcc_positions(a, tuple(int(v) for v in r))
        p = p + rng.normal(0.0, jitter, p.shape)
        pos.append(p)
        cells.append(np.diag(a * r.astype(np.float64)))
        ms.append(np.full(len(p), 63.546))
        zs.append(np.full(len(p), 29))
    return _state(pos, cells, ms, zs, device)


def lj_fluid(n_cells: int = 64, device="cuda", rho=0.8442, jitter=0.05, seed=0) -> SimState:
    """Reduced-unit LJ lattice fluid: 4 n^3 atoms at density rho (sigma = 1)."""
    rng = np.random.default_rng(seed)
    a = (4.0 / rho) ** (1.0 / 3.0)
    pos = fcc_positions(a, (n_cells,) * 3)
    pos = pos + rng.normal(0.0, jitter, pos.shape)
    n = len(pos)
    return _state([pos], [np.eye(3) * a * n_cells], [np.ones(n)], [np.ones(n, dtype=np.int64)], device)


def soft_sphere_batch(
    n_systems: int = 16, n_min: int = 64, n_max: int = 1024, phi: float = 0.64, device="cuda",
    seed: int = 0, strain: float = 0.05,
) -> SimState:
    """Binary soft-sphere systems (sigma_A = 1, sigma_B = 1.4, Z = 1 / 2)."""
    rng = np.random.default_rng(seed)
    pos, cells, ms, zs = [], [], [], []
    for _ in range(n_systems):
        n = int(rng.integers(n_min // 2, n_max // 2 + 1)) * 2
        z = np.array([1] * (n // 2) + [2] * (n // 2))
        d = np.where(z == 1, 1.0, 1.4)
        vol = (math.pi / 6.0 * d**3).sum() / phi
        side = vol ** (1.0 / 3.0)
        e = rng.uniform(-strain, strain, (3, 3))
        e = 0.5 * (e + e.T)
        cell = side * (np.eye(3) + e)
        frac = rng.uniform(0.0, 1.0, (n, 3))
        pos.append(frac @ cell.T)
        cells.append(cell)
        ms.append(np.ones(n))
        zs.append(z)
    return _state(pos, cells, ms, zs, device)


def binary_lj_batch(
    n_systems: int = 8, n_min: int = 128, n_max: int = 512, rho: float = 0.8, device="cuda",
    seed: int = 0,
) -> SimState:
    """Binary (80:20) LJ liquids on a jittered lattice at density rho."""
    rng = np.random.default_rng(seed)
    pos, cells, ms, zs = [], [], [], []
    for _ in range(n_systems):
        n = int(rng.integers(n_min, n_max + 1))
        side = (n / rho) ** (1.0 / 3.0)
        m = int(math.ceil(n ** (1.0 / 3.0)))
        g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
        p = (g + 0.5) * (side / m) + rng.normal(0.0, 0.05, (n, 3))
        z = np.where(rng.uniform(size=n) < 0.8, 1, 2)
        pos.append(p)
        cells.append(np.eye(3) * side)
        ms.append(np.ones(n))
        zs.append(z)
    return _state(pos, cells, ms, zs, device)
