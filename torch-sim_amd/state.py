"""Batched simulation state, mirroring torch_sim.state.SimState.

Field layout and conventions follow torch_sim/state.py:120-326:
``positions [N,3]``, ``masses [N]``, ``cell [S,3,3]`` in the *column-vector*
convention (lattice vector k is ``cell[s, :, k]``), ``pbc [3]`` bool shared by
all systems, ``atomic_numbers [N]`` and ``system_idx [N]`` (int64, sorted,
consecutive from 0).  ``n_systems`` is ``cell.shape[0]`` (the reference
computes it with ``torch.unique``, state.py:437-440; the invariant
``cell.shape[0] == n_systems`` is validated at construction, so both agree
without a device synchronisation).
"""

from __future__ import annotations

import copy
from dataclasses import dataclass, field, fields
from typing import Any, ClassVar

import torch


@dataclass(kw_only=True)
class SimState:
    """State of a batch of periodic atomistic systems (state.py:120-326)."""

    positions: torch.Tensor
    masses: torch.Tensor
    cell: torch.Tensor
    pbc: torch.Tensor | list | bool
    atomic_numbers: torch.Tensor
    system_idx: torch.Tensor | None = None
    _rng: Any = field(default=None, repr=False)
    # model outputs kept on the state (store_model_extras, state.py:376-404):
    # per-system and per-atom tensors, read back as attributes
    _system_extras: dict = field(default_factory=dict, repr=False)
    _atom_extras: dict = field(default_factory=dict, repr=False)

    _atom_attributes: ClassVar[set[str]] = {
        "positions",
        "masses",
        "atomic_numbers",
        "system_idx",
    }
    _system_attributes: ClassVar[set[str]] = {"cell"}
    _global_attributes: ClassVar[set[str]] = {"pbc", "_rng"}
    _extras_names: ClassVar[tuple[str, str]] = ("_system_extras", "_atom_extras")

    def __getattr__(self, name: str):
        """Extras read as attributes (state.py:337-361); only reached when
        the ordinary lookup fails."""
        if not name.startswith("_"):
            for ex in SimState._extras_names:
                try:
                    d = object.__getattribute__(self, ex)
                except AttributeError:
                    continue
                if name in d:
                    return d[name]
        raise AttributeError(f"'{type(self).__name__}' has no attribute or extra '{name}'")

    def __setattr__(self, name: str, value) -> None:
        """A write to an existing extra updates it (None removes it)."""
        if not name.startswith("_"):
            for ex in SimState._extras_names:
                d = self.__dict__.get(ex)
                if d is not None and name in d:
                    if value is None:
                        del d[name]
                    else:
                        d[name] = value
                    return
        object.__setattr__(self, name, value)

    def __post_init__(self) -> None:
        self._system_extras = dict(self._system_extras or {})
        self._atom_extras = dict(self._atom_extras or {})
        if not isinstance(self.pbc, torch.Tensor):
            v = [self.pbc] * 3 if isinstance(self.pbc, bool) else list(self.pbc)
            self.pbc = torch.tensor(v, dtype=torch.bool, device=self.positions.device)
        if self.system_idx is None:
            self.system_idx = torch.zeros(
                self.positions.shape[0], dtype=torch.int64, device=self.positions.device
            )
        if self.cell.ndim == 2:
            self.cell = self.cell.unsqueeze(0)
        n = {self.positions.shape[0], self.masses.shape[0], self.atomic_numbers.shape[0]}
        if len(n) > 1:
            raise ValueError(
                f"Incompatible shapes: positions {self.positions.shape[0]}, "
                f"masses {self.masses.shape[0]}, atomic_numbers {self.atomic_numbers.shape[0]}"
            )
        if self.cell.shape[-2:] != (3, 3):
            raise ValueError("Cell must have shape (n_systems, 3, 3)")

    # ------------------------------------------------------------- basics
    @property
    def device(self) -> torch.device:
        return self.positions.device

    @property
    def dtype(self) -> torch.dtype:
        return self.positions.dtype

    @property
    def n_atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_systems(self) -> int:
        return int(self.cell.shape[0])

    @property
    def n_atoms_per_system(self) -> torch.Tensor:
        return torch.bincount(self.system_idx, minlength=self.n_systems)

    @property
    def row_vector_cell(self) -> torch.Tensor:
        return self.cell.mT

    @row_vector_cell.setter
    def row_vector_cell(self, value: torch.Tensor) -> None:
        self.cell = value.mT

    @property
    def column_vector_cell(self) -> torch.Tensor:
        return self.cell

    @property
    def volume(self) -> torch.Tensor:
        return torch.det(self.cell)

    @property
    def rng(self) -> torch.Generator:
        """Per-state generator on the state's device (state.py:186, 213-221)."""
        if not isinstance(self._rng, torch.Generator):
            g = torch.Generator(device=self.device)
            if self._rng is None:
                g.seed()
            else:
                g.manual_seed(int(self._rng))
            self._rng = g
        return self._rng

    @rng.setter
    def rng(self, value: Any) -> None:
        self._rng = value

    @property
    def system_extras(self) -> dict:
        return self._system_extras

    @property
    def atom_extras(self) -> dict:
        return self._atom_extras

    def has_extras(self, key: str) -> bool:
        return key in self._system_extras or key in self._atom_extras

    def get_number_of_degrees_of_freedom(self) -> torch.Tensor:
        """3 n per system (no constraints; constraints.py:691-699)."""
        return 3 * self.n_atoms_per_system

    # ------------------------------------------------------ manipulation
    def _field_names(self) -> list[str]:
        return [f.name for f in fields(self)]

    def clone(self) -> SimState:
        kw = {}
        for name in self._field_names():
            v = getattr(self, name)
            if isinstance(v, torch.Tensor):
                kw[name] = v.clone()
            elif name in self._extras_names:
                kw[name] = {k: t.clone() for k, t in v.items()}
            elif name == "_rng":
                kw[name] = v if not isinstance(v, torch.Generator) else None
            else:
                kw[name] = copy.copy(v)
        out = type(self)(**kw)
        if isinstance(self._rng, torch.Generator):
            g = torch.Generator(device=self.device)
            g.set_state(self._rng.get_state())
            out._rng = g
        return out

    def to(self, device: torch.device | str | None = None, dtype: torch.dtype | None = None):
        kw = {}
        for name in self._field_names():
            v = getattr(self, name)
            if isinstance(v, torch.Tensor):
                if v.is_floating_point() and dtype is not None:
                    v = v.to(device=device, dtype=dtype)
                else:
                    v = v.to(device=device)
            elif name in self._extras_names:
                v = {k: t.to(device=device) for k, t in v.items()}
            elif name == "_rng":
                v = None if isinstance(v, torch.Generator) else v
            kw[name] = v
        return type(self)(**kw)

    def set_cell(self, cell: torch.Tensor, scale_atoms: bool = False) -> None:
        """Set the cell, optionally keeping fractional coordinates (state.py:537-560)."""
        if cell.shape != self.cell.shape:
            raise ValueError(f"New cell must have shape {self.cell.shape}, got {cell.shape}")
        if scale_atoms:
            m = torch.linalg.solve(self.cell.mT, cell.mT)
            self.positions = torch.bmm(self.positions.unsqueeze(1), m[self.system_idx]).squeeze(1)
        self.cell = cell

    def store_model_extras(self, model_output: dict[str, torch.Tensor]) -> None:
        """Keep the model outputs other than energy / forces / stress
        (state.py:376-404): leading dimension n_atoms -> per-atom extras
        (also when n_atoms == n_systems), n_systems -> per-system extras;
        scalars and other shapes are skipped."""
        n, s = self.n_atoms, self.n_systems
        for key, val in model_output.items():
            if key in ("energy", "forces", "stress"):
                continue
            if not isinstance(val, torch.Tensor) or val.ndim == 0:
                continue
            if val.shape[0] == n:
                self._atom_extras[key] = val
            elif val.shape[0] == s:
                self._system_extras[key] = val

    @classmethod
    def from_state(cls, state: SimState, **kwargs: Any):
        """New state of class ``cls`` from the fields of ``state`` plus ``kwargs``."""
        kw = {n: getattr(state, n) for n in state._field_names() if n in {f.name for f in fields(cls)}}
        kw.update(kwargs)
        return cls(**kw)

    def split(self) -> list[SimState]:
        """One state per system (state.py:1140-1315)."""
        counts = self.n_atoms_per_system.tolist()
        out, a = [], 0
        for s, c in enumerate(counts):
            kw = {}
            for name in self._field_names():
                v = getattr(self, name)
                if name in self._atom_attributes and isinstance(v, torch.Tensor):
                    v = torch.zeros_like(v[a : a + c]) if name == "system_idx" else v[a : a + c]
                elif name in self._system_attributes and isinstance(v, torch.Tensor):
                    v = v[s : s + 1]
                elif name == "_atom_extras":
                    v = {k: t[a : a + c] for k, t in v.items()}
                elif name == "_system_extras":
                    v = {k: t[s : s + 1] for k, t in v.items()}
                elif name == "_rng":
                    v = None
                kw[name] = v
            out.append(type(self)(**kw))
            a += c
        return out


def concatenate_states(states: list[SimState]) -> SimState:
    """Concatenate states along atoms and systems (state.py:1317-1477)."""
    if not states:
        raise ValueError("Cannot concatenate an empty list of states")
    first = states[0]
    cls = type(first)
    kw: dict[str, Any] = {}
    offset = 0
    sys_idx = []
    for st in states:
        sys_idx.append(st.system_idx + offset)
        offset += st.n_systems
    for f in fields(cls):
        name = f.name
        v0 = getattr(first, name)
        if name == "system_idx":
            kw[name] = torch.cat(sys_idx)
        elif name == "pbc":
            kw[name] = v0
        elif name == "_rng":
            kw[name] = None
        elif name in SimState._extras_names:
            # extras every state carries are concatenated; the others dropped
            keys = [k for k in v0 if all(k in getattr(s, name) for s in states)]
            kw[name] = {k: torch.cat([getattr(s, name)[k] for s in states]) for k in keys}
        elif isinstance(v0, torch.Tensor) and v0.ndim >= 1:
            kw[name] = torch.cat([getattr(s, name) for s in states])
        else:
            kw[name] = v0
    return cls(**kw)


def select_systems(state: SimState, sys_ids) -> SimState:
    """New state holding the systems ``sys_ids`` of ``state`` in that order.

    Device gathers over the atom and system attributes; the counterpart of the
    reference's ``state[sys_ids]`` / ``pop`` (state.py:1140-1315) without a
    per-system Python loop.
    """
    dev = state.device
    ids = torch.as_tensor(sys_ids, device=dev, dtype=torch.int64).reshape(-1)
    counts = torch.bincount(state.system_idx, minlength=state.n_systems)
    starts = torch.cumsum(counts, 0) - counts
    sel = counts[ids]
    total = int(sel.sum().item())
    new_sys = torch.repeat_interleave(torch.arange(ids.numel(), device=dev), sel, output_size=total)
    offs = torch.cumsum(sel, 0) - sel
    atom_idx = starts[ids][new_sys] + (torch.arange(total, device=dev) - offs[new_sys])
    kw: dict[str, Any] = {}
    for f in fields(state):
        name = f.name
        v = getattr(state, name)
        if name == "system_idx":
            kw[name] = new_sys
        elif name == "_rng":
            kw[name] = None
        elif isinstance(v, torch.Tensor) and name in state._atom_attributes:
            kw[name] = v[atom_idx]
        elif isinstance(v, torch.Tensor) and name in state._system_attributes:
            kw[name] = v[ids]
        elif name == "_atom_extras":
            kw[name] = {k: t[atom_idx] for k, t in v.items()}
        elif name == "_system_extras":
            kw[name] = {k: t[ids] for k, t in v.items()}
        else:
            kw[name] = v
    return type(state)(**kw)


def initialize_state(system: SimState | list[SimState], device=None, dtype=None) -> SimState:
    """Coerce a SimState or a list of SimStates to one state (state.py:1480-1547)."""
    if isinstance(system, SimState):
        return system.to(device=device, dtype=dtype) if (device or dtype) else system
    if isinstance(system, (list, tuple)) and all(isinstance(s, SimState) for s in system):
        st = concatenate_states(list(system))
        return st.to(device=device, dtype=dtype) if (device or dtype) else st
    raise TypeError(
        "torch_sim_amd builds states from SimState objects; ASE/pymatgen conversion "
        "is outside this package"
    )
