"""Oracle MD integrators (CPU, float64).  Test infrastructure only.

Restates torch_sim/integrators/md.py:113-208, nve.py:19-130 and
nvt.py:27-240.  The Langevin noise is a caller-supplied tensor (the reference
draws ``torch.randn(..., generator=state.rng)`` once per step, nvt.py:75-80),
so trajectories can be compared with an identical noise stream.
"""

from __future__ import annotations

import math

import torch

# MetalUnits (torch_sim/units.py:94-114), CODATA 2014 constants (units.py:38-62)
E_CHARGE = 1.6021766208e-19
AMU = 1.660539040e-27
K_B = 1.38064852e-23
TIME_UNIT = math.sqrt(1.0 * E_CHARGE / (AMU * 1e-20)) * 1e-12  # internal units per ps
KB_EV = K_B / E_CHARGE  # eV / K
PRESSURE_UNIT = 1e5 * (1.0 * 1e-30 / E_CHARGE)  # bar in eV/A^3


def initialize_momenta(masses, system_idx, kT, noise):
    """p = xi sqrt(m kT), minus the per-system mean when n > 1 (md.py:113-166).

    ``noise`` [N,3] replaces ``torch.randn(positions.shape, generator=...)``.
    """
    if kT.ndim > 0:
        kT = kT[system_idx]
    p = noise * torch.sqrt(masses * kT).unsqueeze(-1)
    n_sys = int(system_idx[-1]) + 1
    cnt = torch.bincount(system_idx, minlength=n_sys)
    mean = torch.zeros((n_sys, 3), dtype=p.dtype).index_add(0, system_idx, p)
    mean = mean / cnt.view(-1, 1)
    multi = (cnt > 1)[system_idx].view(-1, 1)
    return torch.where(multi, p - mean[system_idx], p)


def momentum_step(p, f, dt):
    """p + F dt (md.py:169-187)."""
    return p + f * dt


def position_step(x, p, m, dt):
    """x + (p/m) dt (md.py:190-208)."""
    return x + (p / m.unsqueeze(-1)) * dt


def ou_step(p, m, system_idx, dt, kT, gamma, noise):
    """p = c1 p + c2 sqrt(m) xi, c1 = exp(-g dt), c2 = sqrt(kT (1 - c1^2)).

    Restates nvt.py:27-86 with the noise supplied.
    """
    c1 = torch.exp(torch.as_tensor(-gamma * dt, dtype=p.dtype))
    kT = torch.as_tensor(kT, dtype=p.dtype)
    if kT.ndim > 0:
        kT = kT[system_idx]
    if c1.ndim > 0:
        c1 = c1[system_idx]
    c2 = torch.sqrt(kT * (1 - torch.square(c1))).unsqueeze(-1)
    c1 = c1.unsqueeze(-1) if c1.ndim > 0 else c1
    return c1 * p + c2 * torch.sqrt(m).unsqueeze(-1) * noise


def nve_step(x, p, f, m, force_fn, dt):
    """Velocity Verlet B(dt/2) A(dt) F B(dt/2) (nve.py:74-130).

    ``force_fn(x) -> (energy, forces)``.  Returns (x, p, f, energy).
    """
    p = momentum_step(p, f, dt / 2)
    x = position_step(x, p, m, dt)
    e, f = force_fn(x)
    p = momentum_step(p, f, dt / 2)
    return x, p, f, e


def langevin_step(x, p, f, m, system_idx, force_fn, dt, kT, gamma, noise):
    """BAOAB: B(dt/2) A(dt/2) O(dt) A(dt/2) F B(dt/2) (nvt.py:144-240)."""
    p = momentum_step(p, f, dt / 2)
    x = position_step(x, p, m, dt / 2)
    p = ou_step(p, m, system_idx, dt, kT, gamma, noise)
    x = position_step(x, p, m, dt / 2)
    e, f = force_fn(x)
    p = momentum_step(p, f, dt / 2)
    return x, p, f, e


def kinetic_energy(p, m, system_idx, n_sys):
    """Per-system 0.5 sum p^2/m (quantities.py:104-139)."""
    ke = 0.5 * (p.square().sum(1) / m)
    return torch.zeros(n_sys, dtype=p.dtype).index_add(0, system_idx, ke)


def calc_kT(p, m, system_idx, n_sys):
    """Per-system sum p^2/m / (3 n) (quantities.py:16-66, dof = 3n)."""
    s = p.square().sum(1) / m
    tot = torch.zeros(n_sys, dtype=p.dtype).index_add(0, system_idx, s)
    cnt = torch.bincount(system_idx, minlength=n_sys).to(p.dtype)
    return tot / (3.0 * cnt)
