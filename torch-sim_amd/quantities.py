"""Per-system thermodynamic quantities (torch_sim/quantities.py:16-139, 327-342).

These are per-system segment reductions used for reporting and convergence;
they are small torch reductions outside the per-step hot loop.
"""

from __future__ import annotations

import torch


def _n_systems(system_idx: torch.Tensor, n_systems: int | None) -> int:
    return int(n_systems) if n_systems is not None else int(system_idx.max().item()) + 1


def calc_kinetic_energy(
    *, masses: torch.Tensor, momenta: torch.Tensor, system_idx: torch.Tensor,
    n_systems: int | None = None,
) -> torch.Tensor:
    """Per-system 0.5 sum |p|^2 / m (quantities.py:104-139)."""
    s = _n_systems(system_idx, n_systems)
    ke = 0.5 * momenta.square().sum(dim=1) / masses
    return torch.zeros(s, dtype=momenta.dtype, device=momenta.device).index_add(0, system_idx, ke)


def calc_kT(  # noqa: N802
    *, masses: torch.Tensor, momenta: torch.Tensor, system_idx: torch.Tensor,
    dof_per_system: torch.Tensor | None = None,
) -> torch.Tensor:
    """Per-system sum |p|^2/m / dof (quantities.py:16-66), dof = 3 n by default."""
    s = int(system_idx.max().item()) + 1
    v = momenta.square().sum(dim=1) / masses
    tot = torch.zeros(s, dtype=momenta.dtype, device=momenta.device).index_add(0, system_idx, v)
    if dof_per_system is None:
        dof_per_system = 3 * torch.bincount(system_idx, minlength=s)
    return tot / dof_per_system.to(momenta.dtype)


def system_wise_max_force(state) -> torch.Tensor:
    """Per-system max |F_i| (quantities.py:327-342)."""
    norms = state.forces.norm(dim=1)
    out = torch.zeros(state.n_systems, dtype=norms.dtype, device=norms.device)
    return out.scatter_reduce(0, state.system_idx, norms, reduce="amax", include_self=False)
