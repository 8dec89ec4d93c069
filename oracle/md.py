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


# ------------------------------------------------------------- NPT (GJF)
# Isotropic NPT Langevin with the logarithmic strain eps = ln(V/V0)/3 as the
# cell coordinate (torch_sim/integrators/npt.py:653-1107): the
# Gronbech-Jensen-Farago integrator on particles and on eps.

def gjf_particle_position(x, v, f, m, dt, alpha, beta, scale):
    """x' = s x + (2 s / (s + 1)) b dt (v + dt F / (2m) + beta / (2m)),
    b = 1 / (1 + alpha dt / (2m)) (npt.py:838-885).  ``alpha``, ``scale``
    per atom [N]; ``beta`` [N,3]."""
    m2 = 2 * m.unsqueeze(-1)
    b = 1 / (1 + (alpha * dt) / (2 * m))
    c1 = scale.unsqueeze(-1) * x
    c2 = (2 * scale / (scale + 1)) * b * dt
    c3 = v + dt * f / m2 + beta / m2
    return c1 + c2.unsqueeze(-1) * c3


def gjf_particle_velocity(v, f_old, f_new, m, dt, alpha, beta):
    """v' = a v + dt (a F_old + F_new) / (2m) + b beta / m with
    a = (1 - alpha dt / (2m)) / (1 + alpha dt / (2m)) (npt.py:316-367)."""
    m2 = 2 * m
    a = ((1 - (alpha * dt) / m2) / (1 + (alpha * dt) / m2)).unsqueeze(-1)
    b = (1 / (1 + (alpha * dt) / m2)).unsqueeze(-1)
    return a * v + dt * ((a * f_old) + f_new) / m2.unsqueeze(-1) + b * beta / m.unsqueeze(-1)


def gjf_strain_position(eps, eps_dot, force, q, dt, cell_alpha, beta):
    """eps' = eps + b dt eps_dot + b dt^2 F / (2Q) + b dt beta / (2Q),
    b = 1 / (1 + c_alpha dt / (2Q)) (npt.py:769-798)."""
    q2 = 2 * q
    b = 1 / (1 + (cell_alpha * dt) / q2)
    return eps + b * dt * eps_dot + b * dt * dt * force / q2 + b * dt * beta / q2


def gjf_strain_velocity(eps_dot, f_old, f_new, q, dt, cell_alpha, beta):
    """eps_dot' = a eps_dot + dt (a F_old + F_new) / (2Q) + b beta / Q
    (npt.py:801-835)."""
    a = (1 - (cell_alpha * dt) / (2 * q)) / (1 + (cell_alpha * dt) / (2 * q))
    b = 1 / (1 + (cell_alpha * dt) / (2 * q))
    return a * eps_dot + dt * ((a * f_old) + f_new) / (2 * q) + b * beta / q


def strain_force(stress, volume, n_atoms, kT, p_ext):
    """F_eps = 3 V (P - P_ext), P = -tr(sigma)/3 + n kT / V (npt.py:704-749)."""
    p_vir = -torch.einsum("nii->n", stress) / 3
    return 3 * volume * (p_vir + n_atoms * kT / volume - p_ext)


def npt_langevin_isotropic_step(x, p, f, m, system_idx, ref_cell, eps, eps_dot, stress,
                                force_fn, dt, kT, p_ext, alpha, cell_alpha, b_tau, cell_noise,
                                noise):
    """One isotropic NPT Langevin step (npt.py:988-1107) with the noise
    supplied: ``cell_noise`` [S] and ``noise`` [N,3] standard normals, in the
    reference's draw order.  Per-system tensors: kT, p_ext, alpha,
    cell_alpha, b_tau [S]; ref_cell, stress [S,3,3] (column cells).
    ``force_fn(x, cell) -> (energy, forces, stress)``.  Returns
    (x, p, f, energy, stress, cell, eps, eps_dot)."""
    n_sys = ref_cell.shape[0]
    cnt = torch.bincount(system_idx, minlength=n_sys).to(x.dtype)
    q = (cnt + 1) * kT * b_tau * b_tau
    v0 = torch.linalg.det(ref_cell)
    f_eps_n = strain_force(stress, v0 * torch.exp(3 * eps), cnt, kT, p_ext)
    cell_beta = torch.sqrt(2.0 * cell_alpha * kT * dt) * cell_noise
    beta = torch.sqrt(2 * alpha[system_idx] * kT[system_idx] * dt).unsqueeze(-1) * noise
    eps_new = gjf_strain_position(eps, eps_dot, f_eps_n, q, dt, cell_alpha, cell_beta)
    cell = torch.exp(eps_new)[:, None, None] * ref_cell
    scale = torch.exp(eps_new - eps)[system_idx]
    v = p / m.unsqueeze(-1)
    x = gjf_particle_position(x, v, f, m, dt, alpha[system_idx], beta, scale)
    e, f_new, stress_new = force_fn(x, cell)
    f_eps_new = strain_force(stress_new, v0 * torch.exp(3 * eps_new), cnt, kT, p_ext)
    eps_dot = gjf_strain_velocity(eps_dot, f_eps_n, f_eps_new, q, dt, cell_alpha, cell_beta)
    v = gjf_particle_velocity(v, f, f_new, m, dt, alpha[system_idx], beta)
    return x, v * m.unsqueeze(-1), f_new, e, stress_new, cell, eps_new, eps_dot
