"""MD integrators with the reference (init_fn, step_fn) surface.

Mirrors torch_sim/integrators/md.py:31-208, nve.py:19-130, nvt.py:27-240 and
the registry of integrators/__init__.py:125-217.  The per-atom updates run in
the HIP kernels of tsx_md.hip (FP64); forces come from ``model(state)`` once
per step exactly as in the reference.  The Langevin noise of each step is
``torch.randn(momenta.shape, generator=state.rng)`` as in nvt.py:75-80, so a
seeded run draws the same stream as the reference on the same device.

For long runs on one of the pair models, :mod:`torch_sim_amd.engine`
provides the fused, graph-captured path with a Verlet skin.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any

import torch

from . import _lib
from ._lib import call, ptr, stream
from .quantities import calc_kT
from .state import SimState
from .units import MetalUnits


@dataclass(kw_only=True)
class MDState(SimState):
    """MD state: SimState plus momenta, energy and forces (md.py:31-110)."""

    momenta: torch.Tensor
    energy: torch.Tensor
    forces: torch.Tensor

    _atom_attributes = SimState._atom_attributes | {"momenta", "forces"}
    _system_attributes = SimState._system_attributes | {"energy"}

    @property
    def velocities(self) -> torch.Tensor:
        return self.momenta / self.masses.unsqueeze(-1)

    def calc_kT(self) -> torch.Tensor:  # noqa: N802
        return calc_kT(
            masses=self.masses,
            momenta=self.momenta,
            system_idx=self.system_idx,
            dof_per_system=self.get_number_of_degrees_of_freedom(),
        )

    def calc_temperature(self) -> torch.Tensor:
        return self.calc_kT() / MetalUnits.temperature


def _check64(state: SimState) -> None:
    _lib.require_gpu(state.positions, "state.positions")
    if state.dtype != torch.float64:
        raise TypeError("torch_sim_amd integrates positions and momenta in float64")


def _per_system(v, state: SimState) -> torch.Tensor:
    t = torch.as_tensor(v, dtype=torch.float64, device=state.device)
    if t.ndim == 0:
        t = t.expand(state.n_systems)
    return t.contiguous()


def _scalar(v) -> float:
    return float(v.item()) if isinstance(v, torch.Tensor) else float(v)


def initialize_momenta(
    positions: torch.Tensor,
    masses: torch.Tensor,
    system_idx: torch.Tensor,
    kT: torch.Tensor,
    generator: torch.Generator | None = None,
) -> torch.Tensor:
    """Maxwell-Boltzmann momenta with per-system COM removal (md.py:113-166)."""
    kT = torch.as_tensor(kT, dtype=positions.dtype, device=positions.device)
    if kT.ndim > 0:
        kT = kT[system_idx]
    p = torch.randn(
        positions.shape, device=positions.device, dtype=positions.dtype, generator=generator
    ) * torch.sqrt(masses * kT).unsqueeze(-1)
    n_sys = int(system_idx[-1]) + 1
    cnt = torch.bincount(system_idx, minlength=n_sys)
    # segmented sum over the contiguous systems (an index_add would serialise
    # a single large system on one address)
    mean = torch.segment_reduce(p, "sum", lengths=cnt, axis=0) / cnt.view(-1, 1)
    return torch.where(
        torch.repeat_interleave(cnt > 1, cnt).view(-1, 1), p - mean[system_idx], p
    )


def momentum_step(state: MDState, dt) -> MDState:
    """p <- p + F dt (md.py:169-187), in place on the GPU."""
    _check64(state)
    call("tsx_md_kick", ptr(state.momenta), ptr(state.forces), _scalar(dt), state.n_atoms, stream())
    return state


def position_step(state: MDState, dt) -> MDState:
    """x <- x + (p/m) dt (md.py:190-208), in place on the GPU."""
    _check64(state)
    call(
        "tsx_md_drift", ptr(state.positions), ptr(state.momenta), ptr(state.masses),
        _scalar(dt), state.n_atoms, stream(),
    )
    return state


def _prepare(state: SimState) -> None:
    for name in ("positions", "momenta", "forces", "masses"):
        t = getattr(state, name, None)
        if t is not None and not t.is_contiguous():
            setattr(state, name, t.contiguous())


def _md_init(state: SimState, model, kT) -> MDState:
    out = model(state)
    kT_t = torch.as_tensor(kT, device=state.device, dtype=state.dtype)
    momenta = getattr(state, "momenta", None)
    if momenta is None:
        momenta = initialize_momenta(
            state.positions, state.masses, state.system_idx, kT_t, state.rng
        )
    md = MDState.from_state(
        state,
        momenta=momenta.contiguous(),
        energy=out["energy"],
        forces=out["forces"].contiguous(),
    )
    md.positions = md.positions.clone().contiguous()
    return md


def nve_init(state: SimState, model, *, kT, **_kwargs: Any) -> MDState:
    """NVE initial state: forces + Maxwell-Boltzmann momenta (nve.py:19-71)."""
    return _md_init(state, model, kT)


def nve_step(state: MDState, model, *, dt, **_kwargs: Any) -> MDState:
    """Velocity Verlet B(dt/2) A(dt) F B(dt/2) (nve.py:74-130)."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    call(
        "tsx_md_vv_half", ptr(state.positions), ptr(state.momenta), ptr(state.forces),
        ptr(state.masses), d, state.n_atoms, stream(),
    )
    out = model(state)
    state.energy = out["energy"]
    state.forces = out["forces"].contiguous()
    call("tsx_md_kick", ptr(state.momenta), ptr(state.forces), 0.5 * d, state.n_atoms, stream())
    return state


def nvt_langevin_init(state: SimState, model, *, kT, **_kwargs: Any) -> MDState:
    """NVT Langevin initial state (nvt.py:89-141)."""
    return _md_init(state, model, kT)


def nvt_langevin_step(
    state: MDState,
    model,
    *,
    dt,
    kT,
    gamma=None,
    noise: torch.Tensor | None = None,
) -> MDState:
    """BAOAB Langevin step (nvt.py:144-240).

    ``noise`` [N,3] replaces the step's standard-normal draw (for trajectory
    comparisons); by default it is drawn from ``state.rng`` like the reference.
    """
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    g = (1.0 / (100.0 * d)) if gamma is None else gamma
    kT_s = _per_system(kT, state)
    g_s = _per_system(g, state)
    if noise is None:
        noise = torch.randn(
            state.momenta.shape, device=state.device, dtype=state.dtype, generator=state.rng
        )
    noise = noise.to(torch.float64).contiguous()
    atom_sys = state.system_idx.to(torch.int32)
    call(
        "tsx_md_baoa", ptr(state.positions), ptr(state.momenta), ptr(state.forces),
        ptr(state.masses), ptr(atom_sys), ptr(kT_s), ptr(g_s), d, ptr(noise), 0, 0,
        state.n_atoms, stream(),
    )
    out = model(state)
    state.energy = out["energy"]
    state.forces = out["forces"].contiguous()
    call("tsx_md_kick", ptr(state.momenta), ptr(state.forces), 0.5 * d, state.n_atoms, stream())
    return state


@dataclass(kw_only=True)
class NPTState(MDState):
    """MD state with the stress of each system (npt.py:40-59)."""

    stress: torch.Tensor

    _system_attributes = MDState._system_attributes | {"stress"}


@dataclass(kw_only=True)
class NPTLangevinIsotropicState(NPTState):
    """Isotropic NPT Langevin state (npt.py:653-701): the cell coordinate is
    the logarithmic strain eps = ln(V / V0) / 3 of each system."""

    alpha: torch.Tensor
    cell_alpha: torch.Tensor
    b_tau: torch.Tensor
    reference_cell: torch.Tensor
    cell_positions: torch.Tensor  # eps [S]
    cell_velocities: torch.Tensor  # d eps / dt [S]
    cell_masses: torch.Tensor

    _system_attributes = NPTState._system_attributes | {
        "cell_positions",
        "cell_velocities",
        "cell_masses",
        "reference_cell",
        "alpha",
        "cell_alpha",
        "b_tau",
    }

    @property
    def current_cell(self) -> torch.Tensor:
        """exp(eps) h0 (npt.py:690-694)."""
        return torch.exp(self.cell_positions)[:, None, None] * self.reference_cell

    @property
    def volume(self) -> torch.Tensor:
        """V0 exp(3 eps) (npt.py:696-701)."""
        return torch.linalg.det(self.reference_cell) * torch.exp(3 * self.cell_positions)


def npt_langevin_isotropic_init(
    state: SimState,
    model,
    *,
    kT,
    dt,
    alpha=None,
    cell_alpha=None,
    b_tau=None,
    **_kwargs: Any,
) -> NPTLangevinIsotropicState:
    """Isotropic NPT Langevin initial state, eps = 0 (npt.py:888-985).
    Defaults: alpha = 1 / (5 dt), cell_alpha = 1 / (30 dt), b_tau = 300 dt."""
    _check64(state)
    d = _scalar(dt)
    out = model(state)
    if "stress" not in out:
        raise ValueError("NPT needs a model built with compute_stress=True")
    kT_s = _per_system(kT, state)
    alpha_s = _per_system(1.0 / (5 * d) if alpha is None else alpha, state)
    calpha_s = _per_system(1.0 / (30 * d) if cell_alpha is None else cell_alpha, state)
    btau_s = _per_system(300 * d if b_tau is None else b_tau, state)
    momenta = getattr(state, "momenta", None)
    if momenta is None:
        momenta = initialize_momenta(state.positions, state.masses, state.system_idx, kT_s,
                                     state.rng)
    n_per = torch.bincount(state.system_idx, minlength=state.n_systems).to(torch.float64)
    zeros = torch.zeros(state.n_systems, dtype=torch.float64, device=state.device)
    st = NPTLangevinIsotropicState.from_state(
        state,
        momenta=momenta.contiguous(),
        energy=out["energy"],
        forces=out["forces"].contiguous(),
        stress=out["stress"].contiguous(),
        alpha=alpha_s,
        cell_alpha=calpha_s,
        b_tau=btau_s,
        reference_cell=state.cell.clone().contiguous(),
        cell_positions=zeros.clone(),
        cell_velocities=zeros.clone(),
        cell_masses=(n_per + 1) * kT_s * btau_s * btau_s,
    )
    st.positions = st.positions.clone().contiguous()
    st.cell = st.cell.clone().contiguous()
    return st


def npt_langevin_isotropic_step(
    state: NPTLangevinIsotropicState,
    model,
    *,
    dt,
    kT,
    external_pressure,
    noise=None,
) -> NPTLangevinIsotropicState:
    """One isotropic NPT Langevin (GJF) step (npt.py:987-1107), in place.

    The noise is drawn like the reference's: standard normals for the
    strains [S], then for the particles [N,3], from ``state.rng``; ``noise``
    = (cell_noise [S], particle_noise [N,3]) supplies them instead."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    n, s = state.n_atoms, state.n_systems
    kT_s = _per_system(kT, state)
    pext = torch.as_tensor(external_pressure, dtype=torch.float64, device=state.device)
    if pext.ndim >= 2:
        raise ValueError(
            f"External pressure tensor provided with shape {tuple(pext.shape)}. "
            "Only scalar or per-system external pressure is supported."
        )
    pext_s = (pext.expand(s) if pext.ndim == 0 else pext).contiguous()
    if noise is None:
        cell_noise = torch.randn(s, device=state.device, dtype=state.dtype, generator=state.rng)
        part_noise = torch.randn(state.momenta.shape, device=state.device, dtype=state.dtype,
                                 generator=state.rng)
    else:
        cell_noise, part_noise = noise
    cell_noise = cell_noise.to(torch.float64).contiguous()
    part_noise = part_noise.to(torch.float64).contiguous()
    for name in ("alpha", "cell_alpha", "b_tau", "cell_positions", "cell_velocities",
                 "cell_masses", "stress", "reference_cell"):
        setattr(state, name, getattr(state, name).to(torch.float64).contiguous())
    atom_sys = state.system_idx.to(torch.int32)
    sys_start = torch.searchsorted(
        state.system_idx, torch.arange(s + 1, device=state.device, dtype=state.system_idx.dtype)
    ).to(torch.int32)
    ws = torch.empty(int(_lib.lib().tsx_npt_ws_bytes(n, s)) // 8, dtype=torch.float64,
                     device=state.device)
    cell = torch.empty_like(state.reference_cell)
    call("tsx_npt_langevin_iso_pre", s, ptr(sys_start), ptr(atom_sys), n, ptr(state.positions),
         ptr(state.momenta), ptr(state.forces), ptr(state.masses), ptr(kT_s), ptr(pext_s),
         ptr(state.alpha), ptr(state.cell_alpha), ptr(state.b_tau), d, ptr(state.stress),
         ptr(state.reference_cell), ptr(state.cell_positions), ptr(state.cell_velocities),
         ptr(state.cell_masses), ptr(cell), ptr(cell_noise), ptr(part_noise), ptr(ws), stream())
    state.cell = cell
    f_old = state.forces
    out = model(state)
    if "stress" not in out:
        raise ValueError("NPT needs a model built with compute_stress=True")
    state.energy = out["energy"]
    state.forces = out["forces"].contiguous()
    state.stress = out["stress"].contiguous()
    call("tsx_npt_langevin_iso_post", s, ptr(sys_start), ptr(atom_sys), n, ptr(state.momenta),
         ptr(f_old), ptr(state.forces), ptr(state.masses), ptr(kT_s), ptr(pext_s),
         ptr(state.alpha), ptr(state.cell_alpha), d, ptr(state.stress), ptr(state.reference_cell),
         ptr(state.cell_positions), ptr(state.cell_velocities), ptr(state.cell_masses), ptr(ws),
         stream())
    return state


class Integrator(str, Enum):
    """Integrator names (integrators/__init__.py:125-163): NVE, Langevin NVT
    and the isotropic Langevin barostat are the ones on this path."""

    nve = "nve"
    nvt_langevin = "nvt_langevin"
    npt_langevin_isotropic = "npt_langevin_isotropic"


INTEGRATOR_REGISTRY = {
    Integrator.nve: (nve_init, nve_step),
    Integrator.nvt_langevin: (nvt_langevin_init, nvt_langevin_step),
    Integrator.npt_langevin_isotropic: (npt_langevin_isotropic_init, npt_langevin_isotropic_step),
}

# unit-carrying kwargs per call channel (integrators/__init__.py:231-255 reads
# them off the reference's TimeArg / InverseTimeArg / PressureArg annotations):
# times in ps, rates in 1/ps, pressures in bar
INTEGRATOR_KWARG_UNITS = {
    Integrator.nvt_langevin: {"gamma": (("step", 1.0 / MetalUnits.time),)},
    Integrator.npt_langevin_isotropic: {
        "alpha": (("init", 1.0 / MetalUnits.time),),
        "cell_alpha": (("init", 1.0 / MetalUnits.time),),
        "b_tau": (("init", MetalUnits.time),),
        "external_pressure": (("step", MetalUnits.pressure),),
    },
}
