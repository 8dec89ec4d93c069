"""Per-system thermodynamic quantities (torch_sim/quantities.py:16-139, 327-342).

These are per-system segment reductions used for reporting and convergence;
they are small torch reductions outside the per-step hot loop.  The argument
rules are the reference's: exactly one of ``momenta`` / ``velocities``; with
``system_idx`` the result is per system, without it a scalar over all atoms.
"""

from __future__ import annotations

import torch

from .units import MetalUnits


def _squared_term(masses, momenta, velocities) -> torch.Tensor:
    if (momenta is None) == (velocities is None):
        raise ValueError("Must pass either one of momenta or velocities")
    if momenta is not None:
        return momenta.square() / masses.unsqueeze(-1)
    return velocities.square() * masses.unsqueeze(-1)


def _per_system(values: torch.Tensor, system_idx: torch.Tensor, n_systems) -> torch.Tensor:
    s = int(n_systems) if n_systems is not None else int(system_idx.max().item()) + 1
    out = torch.zeros(s, dtype=values.dtype, device=values.device)
    return out.index_add(0, system_idx, values)


def calc_kT(  # noqa: N802
    *,
    masses: torch.Tensor,
    momenta: torch.Tensor | None = None,
    velocities: torch.Tensor | None = None,
    system_idx: torch.Tensor | None = None,
    dof_per_system: torch.Tensor | None = None,
    n_systems: int | None = None,
) -> torch.Tensor:
    """Temperature in energy units, sum m v^2 / dof (quantities.py:16-66);
    dof = 3 n per system unless given."""
    sq = _squared_term(masses, momenta, velocities)
    if system_idx is None:
        return sq.sum() / sq.numel()
    tot = _per_system(sq.sum(dim=-1), system_idx, n_systems)
    if dof_per_system is None:
        dof_per_system = sq.shape[-1] * torch.bincount(system_idx, minlength=tot.shape[0])
    return tot / dof_per_system.to(sq.dtype)


def calc_temperature(
    *,
    masses: torch.Tensor,
    momenta: torch.Tensor | None = None,
    velocities: torch.Tensor | None = None,
    system_idx: torch.Tensor | None = None,
    dof_per_system: torch.Tensor | None = None,
    units: float = MetalUnits.temperature,
) -> torch.Tensor:
    """Temperature in kelvin (quantities.py:69-101): calc_kT / units."""
    kT = calc_kT(masses=masses, momenta=momenta, velocities=velocities, system_idx=system_idx,
                 dof_per_system=dof_per_system)
    return kT / float(units)


def calc_kinetic_energy(
    *,
    masses: torch.Tensor,
    momenta: torch.Tensor | None = None,
    velocities: torch.Tensor | None = None,
    system_idx: torch.Tensor | None = None,
    n_systems: int | None = None,
) -> torch.Tensor:
    """0.5 sum m v^2, per system with ``system_idx`` (quantities.py:104-139)."""
    sq = _squared_term(masses, momenta, velocities)
    if system_idx is None:
        return 0.5 * sq.sum()
    return 0.5 * _per_system(sq.sum(dim=-1), system_idx, n_systems)


def system_wise_max_force(state) -> torch.Tensor:
    """Per-system max |F_i| (quantities.py:327-342)."""
    norms = state.forces.norm(dim=1)
    out = torch.zeros(state.n_systems, dtype=norms.dtype, device=norms.device)
    return out.scatter_reduce(0, state.system_idx, norms, reduce="amax", include_self=False)
