"""MD integrators with the reference (init_fn, step_fn) surface.

Mirrors torch_sim/integrators/md.py:31-565, nve.py:19-130, nvt.py:27-800,
npt.py:35-2986 and the registry of integrators/__init__.py:125-255: NVE,
Langevin, Nose-Hoover chain and CSVR thermostats; Langevin (isotropic and
anisotropic), MTK Nose-Hoover and stochastic cell-rescaling barostats.  The
per-atom updates run in the HIP kernels of tsx_md.hip, tsx_npt.hip and
tsx_thermo.hip (FP64), with the kinetic sums a step needs fused into the
momentum pass that produces them; per-system chains, volumes and cell maps
run one thread per system.  Forces come from ``model(state)`` once per step
exactly as in the reference.  The Langevin noise of each step is
``torch.randn(momenta.shape, generator=state.rng)`` as in nvt.py:75-80, so a
seeded run draws the same stream as the reference on the same device.

For long runs on one of the pair models, :mod:`torch_sim_amd.engine`
provides the fused, graph-captured path with a Verlet skin.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any, Callable

import torch

from . import _lib
from ._lib import call, ptr, stream
from .neighbors import state_system_index
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


# ---------------------------------------------------------------- helpers
def _si(state: SimState):
    si = state_system_index(state)
    if getattr(si, "_mom_ws", None) is None:
        nbytes = int(_lib.lib().tsx_mom_reduce_ws_bytes(si.max_atoms, si.n_systems))
        si._mom_ws = torch.zeros((nbytes + 7) // 8, dtype=torch.float64, device=state.device)
    return si


def _mom_update(state: SimState, *, scale=None, forces=None, h: float = 0.0, reduce: bool = True,
                tensor: bool = False, factor: float = 1.0) -> torch.Tensor | None:
    """p <- scale_s p + h F on the GPU; returns factor * sum p^2 / m [S] (or
    sum p p^T / m as [S,6]) of the updated momenta when ``reduce``."""
    si = _si(state)
    out = None
    if reduce:
        out = torch.empty((state.n_systems, 6) if tensor else (state.n_systems,),
                          dtype=torch.float64, device=state.device)
    call("tsx_md_scale_kick_reduce", state.n_systems, ptr(si.sys_start), si.max_atoms,
         ptr(state.momenta), ptr(state.masses), ptr(scale), ptr(forces), float(h),
         1 if tensor else 0, float(factor), ptr(out), ptr(si._mom_ws), stream())
    return out


def _store(state: MDState, out: dict, stress: bool = False) -> None:
    state.energy = out["energy"]
    state.forces = out["forces"].contiguous()
    if stress:
        if "stress" not in out:
            raise ValueError("NPT needs a model built with compute_stress=True")
        state.stress = out["stress"].contiguous()
    state.store_model_extras(out)


def _f64(t) -> torch.Tensor:
    return t.to(torch.float64).contiguous()


# ------------------------------------------------------ Nose-Hoover chain
@dataclass
class NoseHooverChain:
    """Chain state (md.py:249-273): positions, momenta and masses
    [S, chain_length], tau, kinetic_energy and degrees_of_freedom [S]."""

    positions: torch.Tensor
    momenta: torch.Tensor
    masses: torch.Tensor
    tau: torch.Tensor
    kinetic_energy: torch.Tensor
    degrees_of_freedom: torch.Tensor
    system_idx: torch.Tensor | None = None


@dataclass
class NoseHooverChainFns:
    """Chain functions (md.py:276-287), with the chain's fixed parameters."""

    initialize: Callable
    half_step: Callable
    update_mass: Callable
    dt: float = 0.0
    chain_length: int = 3
    chain_steps: int = 3
    sy_steps: int = 3


def _nhc_call(chain: NoseHooverChain, fns: NoseHooverChainFns, kT_s: torch.Tensor, *,
              update_mass: bool, scale_out=None, sys_mom=None) -> None:
    """tsx_nhc_half_step on the chain's tensors, in place."""
    for name in ("positions", "momenta", "masses"):
        t = getattr(chain, name)
        if t.dtype != torch.float64 or not t.is_contiguous():
            setattr(chain, name, _f64(t))
    for name in ("tau", "kinetic_energy", "degrees_of_freedom"):
        t = getattr(chain, name)
        if t.dtype != torch.float64 or not t.is_contiguous() or t.ndim == 0:
            setattr(chain, name, _f64(t.expand(kT_s.shape[0]) if t.ndim == 0 else t))
    call("tsx_nhc_half_step", kT_s.shape[0], fns.chain_length, fns.chain_steps, fns.sy_steps,
         float(fns.dt), ptr(kT_s), ptr(chain.tau), ptr(chain.degrees_of_freedom),
         ptr(chain.positions), ptr(chain.momenta), ptr(chain.masses), ptr(chain.kinetic_energy),
         1 if update_mass else 0, ptr(scale_out), ptr(sys_mom), stream())


def construct_nose_hoover_chain(dt, chain_length: int, chain_steps: int, sy_steps: int,
                                tau) -> NoseHooverChainFns:
    """Nose-Hoover chain functions (md.py:338-565): Martyna et al.'s
    direct translation with chain_steps x sy_steps Suzuki-Yoshida substeps,
    run per system in one HIP thread (tsx_nhc_half_step)."""
    if sy_steps not in (1, 3, 5, 7):
        raise KeyError(sy_steps)
    d = _scalar(dt)

    def init_fn(degrees_of_freedom, KE, kT) -> NoseHooverChain:  # noqa: N803
        n = KE.shape[0] if KE.dim() > 0 else 1
        dev = KE.device
        kt = torch.as_tensor(kT, device=dev, dtype=KE.dtype)
        kt = kt.expand(n) if kt.ndim == 0 else kt
        t = torch.as_tensor(tau, device=dev, dtype=KE.dtype)
        t = t.expand(n) if t.ndim == 0 else t
        q = kt.unsqueeze(-1) * torch.square(t).unsqueeze(-1) * torch.ones(
            (n, chain_length), dtype=KE.dtype, device=dev)
        q[:, 0] *= degrees_of_freedom
        z = torch.zeros((n, chain_length), dtype=KE.dtype, device=dev)
        return NoseHooverChain(z.clone(), z.clone(), q.contiguous(), _f64(t), _f64(KE),
                               _f64(torch.as_tensor(degrees_of_freedom, device=dev)))

    def half_step_fn(P, chain: NoseHooverChain, kT, system_idx):  # noqa: N803
        kT_s = _f64(torch.as_tensor(kT, device=P.device, dtype=torch.float64).expand(
            chain.kinetic_energy.shape[0]))
        lam = torch.empty_like(kT_s)
        _nhc_call(chain, fns, kT_s, update_mass=False, scale_out=lam)
        return P * lam[system_idx].unsqueeze(-1), chain

    def update_mass_fn(chain: NoseHooverChain, kT) -> NoseHooverChain:
        n = chain.kinetic_energy.shape[0]
        kt = torch.as_tensor(kT, device=chain.masses.device, dtype=chain.masses.dtype)
        kt = kt.expand(n) if kt.ndim == 0 else kt
        q = kt.unsqueeze(-1) * torch.square(chain.tau).unsqueeze(-1) * torch.ones_like(chain.masses)
        q[:, 0] *= chain.degrees_of_freedom
        return NoseHooverChain(chain.positions, chain.momenta, q.contiguous(), chain.tau,
                               chain.kinetic_energy, chain.degrees_of_freedom)

    fns = NoseHooverChainFns(init_fn, half_step_fn, update_mass_fn, d, int(chain_length),
                             int(chain_steps), int(sy_steps))
    return fns


# --------------------------------------------------------- NVT Nose-Hoover
@dataclass(kw_only=True)
class NVTNoseHooverState(MDState):
    """NVT state with a Nose-Hoover chain (nvt.py:243-290)."""

    chain: NoseHooverChain
    _chain_fns: NoseHooverChainFns

    _global_attributes = MDState._global_attributes | {"chain", "_chain_fns"}

    def get_number_of_degrees_of_freedom(self) -> torch.Tensor:
        return super().get_number_of_degrees_of_freedom() - 3


def nvt_nose_hoover_init(state: SimState, model, *, kT, dt, tau=None, chain_length: int = 3,
                         chain_steps: int = 3, sy_steps: int = 3,
                         **kwargs: Any) -> NVTNoseHooverState:
    """NVT Nose-Hoover initial state (nvt.py:293-378); tau defaults to 10 dt."""
    _check64(state)
    d = _scalar(dt)
    kT_t = torch.as_tensor(kT, device=state.device, dtype=state.dtype)
    fns = construct_nose_hoover_chain(d, chain_length, chain_steps, sy_steps,
                                      10.0 * d if tau is None else tau)
    out = model(state)
    momenta = kwargs.get("momenta")
    if momenta is None:
        momenta = getattr(state, "momenta", None)
    if momenta is None:
        momenta = initialize_momenta(state.positions, state.masses, state.system_idx, kT_t,
                                     state.rng)
    si = _si(state)
    ke = torch.empty(state.n_systems, dtype=torch.float64, device=state.device)
    st = NVTNoseHooverState.from_state(
        state, momenta=momenta.contiguous().clone(), energy=out["energy"],
        forces=out["forces"].contiguous(),
        atomic_numbers=kwargs.get("atomic_numbers", state.atomic_numbers),
        chain=None, _chain_fns=fns)
    call("tsx_md_scale_kick_reduce", st.n_systems, ptr(si.sys_start), si.max_atoms,
         ptr(st.momenta), ptr(st.masses), None, None, 0.0, 0, 0.5, ptr(ke), ptr(si._mom_ws),
         stream())
    dof = state.get_number_of_degrees_of_freedom() - 3
    st.chain = fns.initialize(dof, ke, kT_t)
    st.positions = st.positions.clone().contiguous()
    st.store_model_extras(out)
    return st


def nvt_nose_hoover_step(state: NVTNoseHooverState, model, *, dt, kT) -> NVTNoseHooverState:
    """One NHC step (nvt.py:381-488), in place: chain masses and half step,
    velocity Verlet (the chain's momentum scale fused into its first kick
    and drift), the chain's kinetic energy fused into the closing kick,
    chain half step."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    si = _si(state)
    kT_s = _per_system(kT, state)
    chain, fns = state.chain, state._chain_fns
    lam = torch.empty(state.n_systems, dtype=torch.float64, device=state.device)
    _nhc_call(chain, fns, kT_s, update_mass=True, scale_out=lam)
    call("tsx_md_scale_kick_drift", state.n_atoms, ptr(si.atom_sys), ptr(state.positions),
         ptr(state.momenta), ptr(lam), ptr(state.forces), 0.5 * d, ptr(state.masses), d, stream())
    _store(state, model(state))
    chain.kinetic_energy = _mom_update(state, forces=state.forces, h=0.5 * d, factor=0.5)
    _nhc_call(chain, fns, kT_s, update_mass=False, scale_out=lam)
    _mom_update(state, scale=lam, reduce=False)
    return state


def nvt_nose_hoover_invariant(state: NVTNoseHooverState, kT) -> torch.Tensor:
    """Extended energy of NVT with a Nose-Hoover chain (nvt.py:491-562)."""
    kT_s = _per_system(kT, state)
    ke = _mom_update(state, factor=0.5)
    dof = state.get_number_of_degrees_of_freedom().to(torch.float64)
    c = state.chain
    e = state.energy + ke
    e = e + torch.square(c.momenta[:, 0]) / (2 * c.masses[:, 0]) + dof * kT_s * c.positions[:, 0]
    for i in range(1, c.positions.shape[1]):
        e = e + c.momenta[:, i] ** 2 / (2 * c.masses[:, i]) + kT_s * c.positions[:, i]
    return e


# ----------------------------------------------------------- NVT V-Rescale
@dataclass(kw_only=True)
class NVTVRescaleState(MDState):
    """NVT state of the CSVR thermostat (nvt.py:565-591)."""

    def get_number_of_degrees_of_freedom(self) -> torch.Tensor:
        return super().get_number_of_degrees_of_freedom() - 3


def _csvr_draws(state: SimState) -> tuple[torch.Tensor, torch.Tensor]:
    """r1 ~ N(0, 1) and r2 = 2 Gamma((dof - 1) / 2) per system, drawn from
    the state's generator as the reference draws them (nvt.py:635-638)."""
    dof = state.get_number_of_degrees_of_freedom()
    r1 = torch.randn(state.n_systems, device=state.device, dtype=state.dtype,
                     generator=state.rng)
    r2 = torch._standard_gamma((dof - 1) / 2, generator=state.rng) * 2  # noqa: SLF001
    return r1, _f64(r2)


def _vrescale(state: SimState, kT_s, tau_s, dt: float, r1, r2) -> torch.Tensor:
    """The per-system CSVR momentum scale of the current momenta."""
    ke2 = _mom_update(state)
    dof = _f64(state.get_number_of_degrees_of_freedom())
    lam = torch.empty_like(ke2)
    call("tsx_vrescale_factor", state.n_systems, ptr(ke2), ptr(dof), ptr(kT_s), ptr(tau_s),
         float(dt), ptr(_f64(r1)), ptr(_f64(r2)), ptr(lam), stream())
    return lam


def nvt_vrescale_init(state: SimState, model, *, kT, **_kwargs: Any) -> NVTVRescaleState:
    """NVT CSVR initial state (nvt.py:657-712)."""
    md = _md_init(state, model, kT)
    return NVTVRescaleState.from_state(md)


def nvt_vrescale_step(model, state: NVTVRescaleState, *, dt, kT, tau=None,
                      noise=None) -> NVTVRescaleState:
    """One CSVR step (nvt.py:715-800), in place: the rescale fused into the
    first kick and drift of velocity Verlet.  tau defaults to 10 dt;
    ``noise`` = (r1, r2) supplies the draws."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    si = _si(state)
    kT_s = _per_system(kT, state)
    tau_s = _per_system(10.0 * d if tau is None else tau, state)
    r1, r2 = _csvr_draws(state) if noise is None else noise
    lam = _vrescale(state, kT_s, tau_s, d, r1, r2)
    call("tsx_md_scale_kick_drift", state.n_atoms, ptr(si.atom_sys), ptr(state.positions),
         ptr(state.momenta), ptr(lam), ptr(state.forces), 0.5 * d, ptr(state.masses), d, stream())
    _store(state, model(state))
    call("tsx_md_kick", ptr(state.momenta), ptr(state.forces), 0.5 * d, state.n_atoms, stream())
    return state


# ------------------------------------------------ NPT Langevin anisotropic
@dataclass(kw_only=True)
class NPTLangevinAnisotropicState(NPTState):
    """Anisotropic NPT Langevin state (npt.py:62-131): one logarithmic strain
    eps_k = ln(L_k / L_k0) per Cartesian axis, [S,3]."""

    alpha: torch.Tensor
    cell_alpha: torch.Tensor
    b_tau: torch.Tensor
    reference_cell: torch.Tensor
    cell_positions: torch.Tensor  # [S,3]
    cell_velocities: torch.Tensor  # [S,3]
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
        """Row k of the column cell scaled by exp(eps_k) (npt.py:120-124)."""
        return torch.exp(self.cell_positions).unsqueeze(-1) * self.reference_cell

    @property
    def volume(self) -> torch.Tensor:
        return torch.linalg.det(self.cell)


def npt_langevin_anisotropic_init(state: SimState, model, *, kT, dt, alpha=None, cell_alpha=None,
                                  b_tau=None, **_kwargs: Any) -> NPTLangevinAnisotropicState:
    """Anisotropic NPT Langevin initial state, eps = 0 (npt.py:409-510).
    Defaults: alpha = 1 / (5 dt), cell_alpha = 1 / (30 dt), b_tau = 300 dt."""
    iso = npt_langevin_isotropic_init(state, model, kT=kT, dt=dt, alpha=alpha,
                                      cell_alpha=cell_alpha, b_tau=b_tau)
    zeros = torch.zeros((state.n_systems, 3), dtype=torch.float64, device=state.device)
    kw = {n: getattr(iso, n) for n in ("momenta", "energy", "forces", "stress", "alpha",
                                       "cell_alpha", "b_tau", "reference_cell", "cell_masses")}
    st = NPTLangevinAnisotropicState.from_state(iso, cell_positions=zeros.clone(),
                                                cell_velocities=zeros.clone(), **kw)
    return st


def _pext3(external_pressure, state: SimState) -> torch.Tensor:
    p = torch.as_tensor(external_pressure, dtype=torch.float64, device=state.device)
    if p.ndim == 0:
        p = p.expand(state.n_systems, 3)
    elif p.ndim == 1 and p.shape[0] == 3:
        p = p.unsqueeze(0).expand(state.n_systems, 3)
    if p.shape != (state.n_systems, 3):
        raise ValueError(
            f"external_pressure must be a scalar, [3] or [n_systems, 3]; got {tuple(p.shape)}")
    return p.contiguous()


def npt_langevin_anisotropic_step(state: NPTLangevinAnisotropicState, model, *, dt, kT,
                                  external_pressure, noise=None) -> NPTLangevinAnisotropicState:
    """One anisotropic NPT Langevin (GJF) step (npt.py:513-650), in place.
    The draws are the reference's: standard normals for the strains [S,3],
    then for the particles [N,3]; ``noise`` = (cell_noise, particle_noise)
    supplies them."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    si = _si(state)
    n, s = state.n_atoms, state.n_systems
    kT_s = _per_system(kT, state)
    pext = _pext3(external_pressure, state)
    if noise is None:
        cell_noise = torch.randn((s, 3), device=state.device, dtype=state.dtype,
                                 generator=state.rng)
        part_noise = torch.randn(state.momenta.shape, device=state.device, dtype=state.dtype,
                                 generator=state.rng)
    else:
        cell_noise, part_noise = noise
    cell_noise, part_noise = _f64(cell_noise), _f64(part_noise)
    for name in ("alpha", "cell_alpha", "b_tau", "cell_positions", "cell_velocities",
                 "cell_masses", "stress", "reference_cell"):
        setattr(state, name, _f64(getattr(state, name)))
    ws = torch.empty(int(_lib.lib().tsx_npt_ws_bytes(n, s)) // 8, dtype=torch.float64,
                     device=state.device)
    cell = _f64(state.cell).clone()
    call("tsx_npt_langevin_aniso_pre", s, ptr(si.sys_start), ptr(si.atom_sys), n,
         ptr(state.positions), ptr(state.momenta), ptr(state.forces), ptr(state.masses),
         ptr(kT_s), ptr(pext), ptr(state.alpha), ptr(state.cell_alpha), ptr(state.b_tau), d,
         ptr(state.stress), ptr(state.reference_cell), ptr(cell), ptr(state.cell_positions),
         ptr(state.cell_velocities), ptr(state.cell_masses), ptr(cell_noise), ptr(part_noise),
         ptr(ws), stream())
    state.cell = cell
    f_old = state.forces
    _store(state, model(state), stress=True)
    call("tsx_npt_langevin_aniso_post", s, ptr(si.sys_start), ptr(si.atom_sys), n,
         ptr(state.momenta), ptr(f_old), ptr(state.forces), ptr(state.masses), ptr(kT_s),
         ptr(pext), ptr(state.alpha), ptr(state.cell_alpha), d, ptr(state.stress), ptr(cell),
         ptr(state.cell_velocities), ptr(state.cell_masses), ptr(ws), stream())
    return state


# ----------------------------------------------- NPT Nose-Hoover (MTK)
@dataclass(kw_only=True)
class NPTNoseHooverIsotropicState(NPTState):
    """MTK NPT state (npt.py:1110-1218): cell_position = ln(V / V0) / 3
    [S], cell_momentum [S,1], cell_mass [S], thermostat and barostat
    chains."""

    reference_cell: torch.Tensor
    cell_position: torch.Tensor
    cell_momentum: torch.Tensor
    cell_mass: torch.Tensor
    thermostat: NoseHooverChain
    thermostat_fns: NoseHooverChainFns
    barostat: NoseHooverChain
    barostat_fns: NoseHooverChainFns

    _system_attributes = NPTState._system_attributes | {
        "reference_cell", "cell_position", "cell_momentum", "cell_mass"}
    _global_attributes = NPTState._global_attributes | {
        "thermostat", "barostat", "thermostat_fns", "barostat_fns"}

    @property
    def current_cell(self) -> torch.Tensor:
        """(V / V0)^(1/3) h0 with V = V0 exp(3 cell_position) (npt.py:1194-1212)."""
        v0 = torch.det(self.reference_cell)
        v = v0 * torch.exp(3 * self.cell_position)
        return torch.pow(v / v0, 1.0 / 3)[:, None, None] * self.reference_cell

    def get_number_of_degrees_of_freedom(self) -> torch.Tensor:
        return super().get_number_of_degrees_of_freedom() - 3


def npt_nose_hoover_isotropic_init(state: SimState, model, *, kT, dt, chain_length: int = 3,
                                   chain_steps: int = 2, sy_steps: int = 3, t_tau=None,
                                   b_tau=None, **kwargs: Any) -> NPTNoseHooverIsotropicState:
    """MTK NPT initial state (npt.py:1642-1809): t_tau defaults to 10 dt,
    b_tau to 100 dt."""
    _check64(state)
    d = _scalar(dt)
    dev, s = state.device, state.n_systems
    kT_t = torch.as_tensor(kT, device=dev, dtype=state.dtype)
    kT_s = _per_system(kT, state)
    t_tau = 10.0 * d if t_tau is None else t_tau
    b_tau = 100.0 * d if b_tau is None else b_tau
    baro_fns = construct_nose_hoover_chain(d, chain_length, chain_steps, sy_steps, b_tau)
    therm_fns = construct_nose_hoover_chain(d, chain_length, chain_steps, sy_steps, t_tau)
    n_per = torch.bincount(state.system_idx, minlength=s)
    cell_mass = _f64(3 * (n_per + 1) * kT_s * torch.square(_per_system(b_tau, state)))
    cell_momentum = torch.zeros((s, 1), dtype=torch.float64, device=dev)
    momenta = kwargs.get("momenta")
    if momenta is None:
        momenta = getattr(state, "momenta", None)
    if momenta is None:
        momenta = initialize_momenta(state.positions, state.masses, state.system_idx, kT_t,
                                     state.rng)
    out = model(state)
    if "stress" not in out:
        raise ValueError("NPT needs a model built with compute_stress=True")
    st = NPTNoseHooverIsotropicState.from_state(
        state, momenta=momenta.contiguous().clone(), energy=out["energy"],
        forces=out["forces"].contiguous(), stress=out["stress"].contiguous(),
        atomic_numbers=kwargs.get("atomic_numbers", state.atomic_numbers),
        reference_cell=_f64(state.cell).clone(),
        cell_position=torch.zeros(s, dtype=torch.float64, device=dev),
        cell_momentum=cell_momentum, cell_mass=cell_mass,
        thermostat=None, thermostat_fns=therm_fns, barostat=None, barostat_fns=baro_fns)
    ke = _mom_update(st, factor=0.5)
    ke_cell = cell_momentum.squeeze(-1) ** 2 / (2 * cell_mass)
    st.barostat = baro_fns.initialize(torch.ones(s, dtype=torch.float64, device=dev), ke_cell,
                                      kT_t)
    st.thermostat = therm_fns.initialize(state.get_number_of_degrees_of_freedom() - 3, ke, kT_t)
    st.positions = st.positions.clone().contiguous()
    st.cell = st.cell.clone().contiguous()
    st.store_model_extras(out)
    return st


def npt_nose_hoover_isotropic_step(state: NPTNoseHooverIsotropicState, model, *, dt, kT,
                                   external_pressure) -> NPTNoseHooverIsotropicState:
    """One MTK NPT step with Nose-Hoover chains on particles and cell
    (npt.py:1812-1963), in place."""
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    si = _si(state)
    n, s = state.n_atoms, state.n_systems
    kT_s = _per_system(kT, state)
    pext = _per_system(external_pressure, state)
    for name in ("cell_position", "cell_mass", "stress", "reference_cell"):
        setattr(state, name, _f64(getattr(state, name)))
    state.cell_momentum = _f64(state.cell_momentum).reshape(s, 1)
    cm = state.cell_momentum.view(s)
    therm, baro = state.thermostat, state.barostat
    lam = torch.empty(s, dtype=torch.float64, device=state.device)
    _nhc_call(baro, state.barostat_fns, kT_s, update_mass=True, sys_mom=cm)
    _nhc_call(therm, state.thermostat_fns, kT_s, update_mass=True, scale_out=lam)
    ke2 = _mom_update(state, scale=lam)
    dof = _f64(state.get_number_of_degrees_of_freedom())
    ws = torch.empty(int(_lib.lib().tsx_npt_nh_ws_bytes(s)) // 8, dtype=torch.float64,
                     device=state.device)
    cell = torch.empty_like(state.reference_cell)
    call("tsx_npt_nh_pre", s, ptr(si.sys_start), ptr(si.atom_sys), n, ptr(state.positions),
         ptr(state.momenta), ptr(state.forces), ptr(state.masses), ptr(kT_s), ptr(baro.tau),
         ptr(dof), ptr(ke2), ptr(state.stress), ptr(pext), ptr(state.reference_cell),
         ptr(state.cell_position), ptr(cm), ptr(state.cell_mass), ptr(cell), d, ptr(ws), stream())
    state.cell = cell
    _store(state, model(state), stress=True)
    call("tsx_npt_nh_post", s, ptr(si.sys_start), si.max_atoms, ptr(state.momenta),
         ptr(state.forces), ptr(state.masses), ptr(state.stress), ptr(pext),
         ptr(state.reference_cell), ptr(state.cell_position), ptr(cm), ptr(state.cell_mass), d,
         ptr(ke2), ptr(therm.kinetic_energy), ptr(baro.kinetic_energy), ptr(ws),
         ptr(si._mom_ws), stream())
    _nhc_call(therm, state.thermostat_fns, kT_s, update_mass=False, scale_out=lam)
    _mom_update(state, scale=lam, reduce=False)
    _nhc_call(baro, state.barostat_fns, kT_s, update_mass=False, sys_mom=cm)
    return state


def npt_nose_hoover_isotropic_invariant(state: NPTNoseHooverIsotropicState, kT,
                                        external_pressure) -> torch.Tensor:
    """Extended Hamiltonian of the MTK NPT dynamics (npt.py:1997-2059)."""
    kT_s = _per_system(kT, state)
    pext = _per_system(external_pressure, state)
    vol = torch.det(state.current_cell)
    ke = _mom_update(state, factor=0.5)
    dof = state.get_number_of_degrees_of_freedom().to(torch.float64)

    def chain_e(c, nf):
        e = torch.square(c.momenta[:, 0]) / (2 * c.masses[:, 0]) + nf * kT_s * c.positions[:, 0]
        for i in range(1, c.positions.shape[1]):
            e = e + torch.square(c.momenta[:, i]) / (2 * c.masses[:, i]) + kT_s * c.positions[:, i]
        return e

    e = state.energy + ke
    e = e + chain_e(state.thermostat, dof)
    e = e + chain_e(state.barostat, torch.ones_like(dof))
    e = e + pext * vol
    return e + torch.square(state.cell_momentum.reshape(-1)) / (2 * state.cell_mass)


# ------------------------------------------------ stochastic cell rescaling
@dataclass(kw_only=True)
class NPTCRescaleState(NPTState):
    """State of the cell-rescaling barostats (npt.py:2062-2090)."""

    isothermal_compressibility: torch.Tensor
    tau_p: torch.Tensor
    initial_cell: torch.Tensor
    initial_cell_inv: torch.Tensor
    initial_volume: torch.Tensor

    _system_attributes = NPTState._system_attributes | {
        "isothermal_compressibility", "tau_p", "initial_cell", "initial_cell_inv",
        "initial_volume"}

    def get_number_of_degrees_of_freedom(self) -> torch.Tensor:
        return super().get_number_of_degrees_of_freedom() - 3


def npt_crescale_init(state: SimState, model, *, kT, dt, tau_p=None,
                      isothermal_compressibility=None, **_kwargs: Any) -> NPTCRescaleState:
    """C-rescale initial state (npt.py:2904-2986): tau_p defaults to 3 dt,
    the compressibility to 1e-6 / bar."""
    _check64(state)
    d = _scalar(dt)
    kT_t = torch.as_tensor(kT, device=state.device, dtype=state.dtype)
    tp = _per_system(3 * d if tau_p is None else tau_p, state)
    bt = _per_system(1e-6 / MetalUnits.pressure if isothermal_compressibility is None
                     else isothermal_compressibility, state)
    out = model(state)
    if "stress" not in out:
        raise ValueError("NPT needs a model built with compute_stress=True")
    momenta = getattr(state, "momenta", None)
    if momenta is None:
        momenta = initialize_momenta(state.positions, state.masses, state.system_idx, kT_t,
                                     state.rng)
    c0 = _f64(state.cell).clone()
    st = NPTCRescaleState.from_state(
        state, momenta=momenta.contiguous().clone(), energy=out["energy"],
        forces=out["forces"].contiguous(), stress=out["stress"].contiguous(), tau_p=tp,
        isothermal_compressibility=bt, initial_cell=c0,
        initial_cell_inv=torch.inverse(c0).contiguous(), initial_volume=torch.det(c0))
    st.positions = st.positions.clone().contiguous()
    st.cell = st.cell.clone().contiguous()
    st.store_model_extras(out)
    return st


_CR_MODES = {"isotropic": 0, "anisotropic": 1, "triclinic": 2, "average": 3}


def _crescale_draws(state: SimState, mode: str) -> dict:
    """The step's draws in the reference's order (npt.py:2517-2901):
    CSVR (r1, r2), the barostat's normals, CSVR again."""
    g, s, dev = state.rng, state.n_systems, state.device
    out = {}
    out["r1a"], out["r2a"] = _csvr_draws(state)
    out["ra"] = torch.randn(s, device=dev, dtype=torch.float64, generator=g)
    if mode != "isotropic":
        shape = (s, 3) if mode == "anisotropic" else (s, 3, 3)
        out["R"] = torch.randn(shape, device=dev, dtype=torch.float64, generator=g)
        out["rb"] = torch.randn(s, device=dev, dtype=torch.float64, generator=g)
    out["r1b"], out["r2b"] = _csvr_draws(state)
    return out


def _crescale_step(mode: str, state: NPTCRescaleState, model, dt, kT, external_pressure, tau,
                   noise) -> NPTCRescaleState:
    _check64(state)
    _prepare(state)
    d = _scalar(dt)
    si = _si(state)
    n, s = state.n_atoms, state.n_systems
    kT_s = _per_system(kT, state)
    tau_s = _per_system(tau, state)
    pe = torch.as_tensor(external_pressure, dtype=torch.float64, device=state.device)
    pext_t = None
    if pe.ndim >= 2:
        if mode != "triclinic":
            raise ValueError("a target stress tensor needs the triclinic barostat")
        pext_t = (pe.unsqueeze(0).expand(s, 3, 3) if pe.ndim == 2 else pe).contiguous()
        pext = torch.zeros(s, dtype=torch.float64, device=state.device)
    else:
        pext = _per_system(pe, state)
    z = _crescale_draws(state, mode) if noise is None else noise
    packed = torch.zeros((s, 11), dtype=torch.float64, device=state.device)
    packed[:, 0] = z["ra"]
    if mode == "anisotropic":
        packed[:, 1:4] = z["R"]
    elif mode != "isotropic":
        packed[:, 1:10] = z["R"].reshape(s, 9)
    if mode != "isotropic":
        packed[:, 10] = z["rb"]
    for name in ("isothermal_compressibility", "tau_p", "stress", "initial_cell_inv",
                 "initial_volume"):
        setattr(state, name, _f64(getattr(state, name)))
    state.cell = _f64(state.cell).clone()
    lam = _vrescale(state, kT_s, tau_s, 0.5 * d, z["r1a"], z["r2a"])
    kin = _mom_update(state, scale=lam, forces=state.forces, h=0.5 * d, tensor=True,
                      reduce=mode != "average")
    dof = _f64(state.get_number_of_degrees_of_freedom())
    ws = torch.empty(int(_lib.lib().tsx_crescale_ws_bytes(s)) // 8, dtype=torch.float64,
                     device=state.device)
    call("tsx_crescale_step", s, _CR_MODES[mode], ptr(si.atom_sys), n, ptr(state.positions),
         ptr(state.momenta), ptr(state.masses), ptr(kin), ptr(state.stress), ptr(state.cell),
         ptr(kT_s), ptr(state.isothermal_compressibility), ptr(state.tau_p), d, ptr(pext),
         ptr(pext_t), ptr(state.initial_cell_inv), ptr(state.initial_volume), ptr(dof),
         ptr(packed), ptr(ws), stream())
    _store(state, model(state), stress=True)
    _mom_update(state, forces=state.forces, h=0.5 * d, reduce=False)
    lam = _vrescale(state, kT_s, tau_s, 0.5 * d, z["r1b"], z["r2b"])
    _mom_update(state, scale=lam, reduce=False)
    return state


def npt_crescale_isotropic_step(state: NPTCRescaleState, model, *, dt, kT, external_pressure,
                                tau=None, noise=None) -> NPTCRescaleState:
    """Isotropic C-rescale step (npt.py:2797-2901): CSVR(dt/2), B(dt/2),
    barostat, forces, B(dt/2), CSVR(dt/2); tau defaults to dt."""
    d = _scalar(dt)
    return _crescale_step("isotropic", state, model, d, kT, external_pressure,
                          tau or 1 * d, noise)


def npt_crescale_anisotropic_step(state: NPTCRescaleState, model, *, dt, kT, external_pressure,
                                  tau=None, noise=None) -> NPTCRescaleState:
    """Anisotropic (diagonal) C-rescale step (npt.py:2645-2718); tau
    defaults to 100 dt."""
    d = _scalar(dt)
    return _crescale_step("anisotropic", state, model, d, kT, external_pressure,
                          tau or 100 * d, noise)


def npt_crescale_triclinic_step(state: NPTCRescaleState, model, *, dt, kT, external_pressure,
                                tau=None, noise=None) -> NPTCRescaleState:
    """Triclinic C-rescale step (npt.py:2517-2642), with a deviatoric target
    when external_pressure is a [3,3] / [S,3,3] tensor; tau defaults to dt."""
    d = _scalar(dt)
    return _crescale_step("triclinic", state, model, d, kT, external_pressure,
                          1 * d if tau is None else tau, noise)


def npt_crescale_triclinic_average_step(state: NPTCRescaleState, model, *, dt, kT,
                                        external_pressure, tau=None,
                                        noise=None) -> NPTCRescaleState:
    """Triclinic C-rescale with the equipartition kinetic pressure, positions
    scaled only (npt.py:2721-2794); tau defaults to dt."""
    d = _scalar(dt)
    return _crescale_step("average", state, model, d, kT, external_pressure, tau or 1 * d, noise)


class Integrator(str, Enum):
    """Integrator names (integrators/__init__.py:125-163)."""

    nve = "nve"
    nvt_vrescale = "nvt_vrescale"
    nvt_langevin = "nvt_langevin"
    nvt_nose_hoover = "nvt_nose_hoover"
    npt_langevin_anisotropic = "npt_langevin_anisotropic"
    npt_langevin_isotropic = "npt_langevin_isotropic"
    npt_nose_hoover_isotropic = "npt_nose_hoover_isotropic"
    npt_crescale_isotropic = "npt_crescale_isotropic"
    npt_crescale_triclinic = "npt_crescale_triclinic"


INTEGRATOR_REGISTRY = {
    Integrator.nve: (nve_init, nve_step),
    Integrator.nvt_vrescale: (nvt_vrescale_init, nvt_vrescale_step),
    Integrator.nvt_langevin: (nvt_langevin_init, nvt_langevin_step),
    Integrator.nvt_nose_hoover: (nvt_nose_hoover_init, nvt_nose_hoover_step),
    Integrator.npt_langevin_anisotropic: (npt_langevin_anisotropic_init,
                                          npt_langevin_anisotropic_step),
    Integrator.npt_langevin_isotropic: (npt_langevin_isotropic_init, npt_langevin_isotropic_step),
    Integrator.npt_nose_hoover_isotropic: (npt_nose_hoover_isotropic_init,
                                           npt_nose_hoover_isotropic_step),
    Integrator.npt_crescale_isotropic: (npt_crescale_init, npt_crescale_isotropic_step),
    Integrator.npt_crescale_triclinic: (npt_crescale_init, npt_crescale_triclinic_step),
}

# unit-carrying kwargs per call channel (integrators/__init__.py:221-255 reads
# them off the reference's TimeArg / InverseTimeArg / PressureArg /
# InversePressureArg annotations): times in ps, rates in 1/ps, pressures in
# bar, compressibilities in 1/bar
_T, _IT = MetalUnits.time, 1.0 / MetalUnits.time
_P, _IP = MetalUnits.pressure, 1.0 / MetalUnits.pressure
_LANGEVIN_NPT_UNITS = {
    "alpha": (("init", _IT),),
    "cell_alpha": (("init", _IT),),
    "b_tau": (("init", _T),),
    "external_pressure": (("step", _P),),
}
_CRESCALE_UNITS = {
    "tau_p": (("init", _T),),
    "isothermal_compressibility": (("init", _IP),),
    "tau": (("step", _T),),
    "external_pressure": (("step", _P),),
}
INTEGRATOR_KWARG_UNITS = {
    Integrator.nvt_vrescale: {"tau": (("step", _T),)},
    Integrator.nvt_langevin: {"gamma": (("step", _IT),)},
    Integrator.nvt_nose_hoover: {"tau": (("init", _T),)},
    Integrator.npt_langevin_anisotropic: dict(_LANGEVIN_NPT_UNITS),
    Integrator.npt_langevin_isotropic: dict(_LANGEVIN_NPT_UNITS),
    Integrator.npt_nose_hoover_isotropic: {
        "t_tau": (("init", _T),),
        "b_tau": (("init", _T),),
        "external_pressure": (("step", _P),),
    },
    Integrator.npt_crescale_isotropic: dict(_CRESCALE_UNITS),
    Integrator.npt_crescale_triclinic: dict(_CRESCALE_UNITS),
}
