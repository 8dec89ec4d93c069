"""FIRE relaxation with the unit and Frechet cell filters.

Mirrors torch_sim.optimizers: ``fire_init`` / ``fire_step``
(optimizers/fire.py:21-175), the ``FireState`` / ``CellFireState`` classes
(optimizers/state.py:38-51, cell_filters.py:426-470), the ``CellFilter`` names
(cell_filters.py:287-292) and ``Optimizer`` / ``OPTIM_REGISTRY``
(optimizers/__init__.py:37-52).  The arithmetic of a step runs in the HIP
kernels behind ``tsx_fire_pre`` (before the model call) and
``tsx_fire_post`` (after it); the model is any ``ModelInterface``.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any

import torch

from . import _lib
from .neighbors import SystemIndex, state_system_index
from .state import SimState

FIRE_FLAVORS = ("vv_fire", "ase_fire")


class Optimizer(str, Enum):
    """Optimizer names (optimizers/__init__.py:37-43).  Only FIRE is on this path."""

    fire = "fire"


class CellFilter(str, Enum):
    """Cell filter names (cell_filters.py:287-292)."""

    unit = "unit"
    frechet = "frechet"


_FILTER_CODE = {None: 0, CellFilter.unit: 1, CellFilter.frechet: 2}


@dataclass(kw_only=True)
class OptimState(SimState):
    """SimState plus forces, energy, stress (optimizers/state.py:11-36)."""

    forces: torch.Tensor
    energy: torch.Tensor
    stress: torch.Tensor | None = None

    _atom_attributes = SimState._atom_attributes | {"forces"}
    _system_attributes = SimState._system_attributes | {"energy", "stress"}


@dataclass(kw_only=True)
class FireState(OptimState):
    """FIRE state (optimizers/state.py:38-51)."""

    velocities: torch.Tensor
    dt: torch.Tensor
    alpha: torch.Tensor
    n_pos: torch.Tensor
    fmax: torch.Tensor | None = None  # [S, 2]: max |F_i|, max cell-force row norm

    _atom_attributes = OptimState._atom_attributes | {"velocities"}
    _system_attributes = OptimState._system_attributes | {"dt", "alpha", "n_pos", "fmax"}


@dataclass(kw_only=True)
class CellFireState(FireState):
    """FIRE state with a cell filter (cell_filters.py:426-470)."""

    reference_cell: torch.Tensor
    cell_filter: CellFilter = CellFilter.frechet
    cell_factor: torch.Tensor | None = None
    pressure: torch.Tensor | None = None
    hydrostatic_strain: bool = False
    constant_volume: bool = False
    cell_positions: torch.Tensor | None = None
    cell_forces: torch.Tensor | None = None
    cell_masses: torch.Tensor | None = None
    cell_velocities: torch.Tensor | None = None

    _system_attributes = FireState._system_attributes | {
        "reference_cell",
        "cell_factor",
        "pressure",
        "cell_positions",
        "cell_forces",
        "cell_masses",
        "cell_velocities",
    }
    _global_attributes = FireState._global_attributes | {
        "cell_filter",
        "hydrostatic_strain",
        "constant_volume",
    }


def _as_filter(cell_filter) -> CellFilter | None:
    if cell_filter is None:
        return None
    if isinstance(cell_filter, CellFilter):
        return cell_filter
    if isinstance(cell_filter, str):
        return CellFilter(cell_filter)
    raise ValueError(
        f"Unknown cell_filter={cell_filter!r}; torch_sim_amd implements "
        f"{[c.value for c in CellFilter]} (custom filter callables stay on the reference path)"
    )


def _params(flavor: str, cell_filter, *, dt_max=1.0, n_min=5, f_inc=1.1, f_dec=0.5,
            alpha_start=0.1, f_alpha=0.99, max_step=0.2, eps=1e-16, hydrostatic=False,
            constant_volume=False, first_step=False) -> _lib.FireParams:
    if flavor not in FIRE_FLAVORS:
        raise ValueError(f"Unknown fire_flavor={flavor!r}, must be one of {FIRE_FLAVORS}")
    return _lib.FireParams(
        dt_max=float(dt_max), f_inc=float(f_inc), f_dec=float(f_dec),
        alpha_start=float(alpha_start), f_alpha=float(f_alpha), max_step=float(max_step),
        eps=float(eps), n_min=int(n_min), flavor=0 if flavor == "ase_fire" else 1,
        cell_filter=_FILTER_CODE[cell_filter], hydrostatic_strain=int(bool(hydrostatic)),
        constant_volume=int(bool(constant_volume)), first_step=int(bool(first_step)),
    )


def _check(state: SimState) -> None:
    _lib.require_gpu(state.positions, "state.positions")
    if state.dtype != torch.float64:
        raise TypeError("torch_sim_amd runs FIRE in float64")


def _sysidx(state: SimState) -> SystemIndex:
    return state_system_index(state)


def _post(state: FireState, prm: _lib.FireParams, *, init: bool, active=None,
          flags=None, cell_build=None) -> None:
    si = _sysidx(state)
    cellf = isinstance(state, CellFireState)
    stress = state.stress.contiguous() if (cellf and state.stress is not None) else None
    if cellf and stress is None:
        raise ValueError("a cell filter needs a model that returns stress")
    if state.fmax is None or state.fmax.shape != (state.n_systems, 2):
        state.fmax = torch.zeros((state.n_systems, 2), dtype=torch.float64, device=state.device)
    _lib.call(
        "tsx_fire_post", ctypes_ref(prm), state.n_systems, _lib.ptr(si.sys_start),
        _lib.ptr(state.velocities), _lib.ptr(state.forces), _lib.ptr(state.masses),
        _lib.ptr(state.dt), _lib.ptr(state.alpha), _lib.ptr(state.n_pos), _lib.ptr(state.cell),
        _lib.ptr(state.reference_cell) if cellf else None,
        _lib.ptr(state.cell_factor) if cellf else None,
        _lib.ptr(state.pressure) if cellf else None, _lib.ptr(stress),
        _lib.ptr(state.cell_velocities) if cellf else None,
        _lib.ptr(state.cell_forces) if cellf else None,
        _lib.ptr(state._cell_mass_scalar) if cellf else None,
        _lib.ptr(active), _lib.ptr(state.fmax), _lib.ptr(cell_build), _lib.ptr(flags),
        int(init), _lib.stream(),
    )


def ctypes_ref(prm):
    import ctypes

    return ctypes.byref(prm)


def fire_init(
    state: SimState,
    model,
    *,
    dt_start: float = 0.1,
    alpha_start: float = 0.1,
    fire_flavor: str = "ase_fire",
    cell_filter: CellFilter | str | None = None,
    cell_factor: float | torch.Tensor | None = None,
    hydrostatic_strain: bool = False,
    constant_volume: bool = False,
    scalar_pressure: float = 0.0,
    **_filter_kwargs: Any,
) -> FireState | CellFireState:
    """Initialise a FIRE state (fire.py:21-114; filter inits cell_filters.py:180-284)."""
    if fire_flavor not in FIRE_FLAVORS:
        raise ValueError(f"Unknown fire_flavor={fire_flavor!r}, must be one of {FIRE_FLAVORS}")
    _check(state)
    cf = _as_filter(cell_filter)
    dev, dt_ = state.device, torch.float64
    n_sys = state.n_systems
    out = model(state)
    dt = torch.as_tensor(dt_start, device=dev, dtype=dt_)
    dt = dt.expand(n_sys).clone() if dt.ndim == 0 else dt.clone()
    alpha = torch.as_tensor(alpha_start, device=dev, dtype=dt_)
    alpha = alpha.expand(n_sys).clone() if alpha.ndim == 0 else alpha.clone()
    attrs = dict(
        forces=out["forces"].contiguous(),
        energy=out["energy"],
        stress=out.get("stress"),
        velocities=torch.full_like(state.positions, float("nan")),
        dt=dt,
        alpha=alpha,
        n_pos=torch.zeros(n_sys, device=dev, dtype=torch.int32),
    )
    base = {n: getattr(state, n) for n in ("positions", "masses", "cell", "pbc", "atomic_numbers",
                                           "system_idx")}
    base["positions"] = state.positions.clone().contiguous()
    base["cell"] = state.cell.clone().contiguous()
    if cf is None:
        fs = FireState(**base, **attrs)
        _post(fs, _params(fire_flavor, None), init=True)
        return fs
    counts = torch.bincount(state.system_idx, minlength=n_sys).to(dt_)
    if cell_factor is None:
        cfac = counts
    else:
        cfac = torch.as_tensor(cell_factor, device=dev, dtype=dt_)
        cfac = cfac.expand(n_sys).clone() if cfac.ndim == 0 else cfac
        if cfac.numel() != n_sys:
            raise ValueError(f"cell_factor tensor must have {n_sys} elements, got {cfac.numel()}")
    masses_sum = torch.zeros(n_sys, device=dev, dtype=dt_).index_add_(0, state.system_idx, state.masses)
    eye = torch.eye(3, device=dev, dtype=dt_)
    fs = CellFireState(
        **base,
        **attrs,
        reference_cell=state.cell.clone().contiguous(),
        cell_filter=cf,
        cell_factor=cfac.contiguous(),
        pressure=(scalar_pressure * eye).expand(n_sys, 3, 3).contiguous(),
        hydrostatic_strain=hydrostatic_strain,
        constant_volume=constant_volume,
        cell_positions=(torch.zeros((n_sys, 3, 3), device=dev, dtype=dt_) if cf == CellFilter.frechet
                        else (eye * cfac.view(n_sys, 1, 1)).contiguous()),
        cell_forces=torch.zeros((n_sys, 3, 3), device=dev, dtype=dt_),
        cell_masses=masses_sum.unsqueeze(-1).expand(-1, 3).contiguous(),
        cell_velocities=torch.full((n_sys, 3, 3), float("nan"), device=dev, dtype=dt_),
    )
    prm = _params(fire_flavor, cf, hydrostatic=hydrostatic_strain, constant_volume=constant_volume)
    _post(fs, prm, init=True)
    return fs


def fire_step(
    state: FireState | CellFireState,
    model,
    *,
    dt_max: float = 1.0,
    n_min: int = 5,
    f_inc: float = 1.1,
    f_dec: float = 0.5,
    alpha_start: float = 0.1,
    f_alpha: float = 0.99,
    max_step: float = 0.2,
    fire_flavor: str = "ase_fire",
) -> FireState | CellFireState:
    """One FIRE step (fire.py:117-175, 179-499); mutates and returns ``state``."""
    _check(state)
    cellf = isinstance(state, CellFireState)
    cf = state.cell_filter if cellf else None
    first = False
    if fire_flavor == "ase_fire":
        # fire.py:313-315: the raw-force first step when every velocity is NaN.
        # Velocities that this function's previous call left, untouched since
        # (same tensor, same in-place version), are not all NaN: skip the
        # reduction and its host synchronisation.
        vel, seen = state.velocities, getattr(state, "_fire_velocities_seen", None)
        if seen is None or seen[0] is not vel or seen[1] != vel._version:
            first = bool(vel.isnan().any(dim=1).all())
    prm = _params(fire_flavor, cf, dt_max=dt_max, n_min=n_min, f_inc=f_inc, f_dec=f_dec,
                  alpha_start=alpha_start, f_alpha=f_alpha, max_step=max_step,
                  hydrostatic=state.hydrostatic_strain if cellf else False,
                  constant_volume=state.constant_volume if cellf else False, first_step=first)
    si = _sysidx(state)
    for name in ("positions", "velocities", "forces", "cell"):
        t = getattr(state, name)
        if not t.is_contiguous():
            setattr(state, name, t.contiguous())
    ws = torch.empty(max(1, int(_lib.lib().tsx_fire_ws_bytes(state.n_systems)) // 8),
                     dtype=torch.float64, device=state.device)
    _lib.call(
        "tsx_fire_pre", ctypes_ref(prm), state.n_systems, _lib.ptr(si.sys_start),
        _lib.ptr(si.atom_sys), state.n_atoms, _lib.ptr(state.positions), 3,
        _lib.ptr(state.velocities), _lib.ptr(state.forces), _lib.ptr(state.masses),
        _lib.ptr(state.dt), _lib.ptr(state.alpha), _lib.ptr(state.n_pos), _lib.ptr(state.cell),
        _lib.ptr(state.reference_cell) if cellf else None,
        _lib.ptr(state.cell_factor) if cellf else None,
        _lib.ptr(state.cell_positions) if cellf else None,
        _lib.ptr(state.cell_velocities) if cellf else None,
        _lib.ptr(state.cell_forces) if cellf else None,
        None, None, None, 0.0, 0.0, None, _lib.ptr(ws), _lib.stream(),
    )
    out = model(state)
    state.forces = out["forces"].contiguous()
    state.energy = out["energy"]
    if "stress" in out:
        state.stress = out["stress"]
    _post(state, prm, init=False)
    state._fire_velocities_seen = (state.velocities, state.velocities._version)
    return state


def _cell_mass_scalar(self) -> torch.Tensor:
    return self.cell_masses[:, 0].contiguous()


CellFireState._cell_mass_scalar = property(_cell_mass_scalar)

OPTIM_REGISTRY = {Optimizer.fire: (fire_init, fire_step)}

__all__ = [
    "CellFilter",
    "CellFireState",
    "FireState",
    "OPTIM_REGISTRY",
    "OptimState",
    "Optimizer",
    "fire_init",
    "fire_step",
]
