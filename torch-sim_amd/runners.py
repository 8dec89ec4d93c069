"""High-level entry points with the reference signatures (torch_sim/runners.py).

``integrate`` (runners.py:252-413): unit conversion (temperature in K,
timestep in ps, Langevin gamma in 1/ps), ``init_fn`` then ``n_steps`` calls
of ``step_fn``.  When the model is a closed-form pair model, the integrator
is NVE or Langevin NVT, the temperature is constant over the run and no
reporter is attached, ``engine=None`` (default) routes the run through the
fused, graph-captured :class:`~torch_sim_amd.engine.BatchedMD`; ``engine=False``
keeps the reference's per-step loop.  A trajectory reporter
(runners.py:34-70, 108-250, 338-404) is called after every step of the
per-step loop; on the engine the run is cut at the steps where the reporter
writes something and the closed state of that step is reported.

``optimize`` (runners.py:567-753): ``init_fn`` on the systems in flight,
``steps_between_swaps`` steps, a per-system convergence check, removal of
converged systems and refill from the queue, results returned in the input
order.  ``generate_force_convergence_fn`` / ``generate_energy_convergence_fn``
follow runners.py:501-564; the force check reads the per-system maxima that
``tsx_fire_post`` writes with the step.
"""

from __future__ import annotations

import warnings
from typing import Any, Callable

import torch

from dataclasses import fields as dc_fields

import numpy as np

from .engine import BatchedMD, estimate_skin
from .integrators import INTEGRATOR_KWARG_UNITS, INTEGRATOR_REGISTRY, Integrator, MDState
from .models import PairPotentialModel
from .neighbors import SystemIndex, TensorKey, VerletList, assemble_verlet, segment_index
from .optimizers import OPTIM_REGISTRY, Optimizer, fire_init, fire_step
from .quantities import calc_kinetic_energy, system_wise_max_force
from .state import SimState, concatenate_states, initialize_state, select_systems
from .trajectory import TrajectoryReporter, _per_system
from .units import MetalUnits


def _configure_reporter(reporter, *, properties=None, prop_frequency: int = 10,
                        state_frequency: int = 100) -> TrajectoryReporter:
    """A reporter, or one built from a dict of its keywords with the named
    default properties (runners.py:34-70)."""
    if isinstance(reporter, TrajectoryReporter):
        return reporter
    possible = {
        "potential_energy": _per_system(lambda state: state.energy),
        "forces": _per_system(lambda state: state.forces),
        "stress": _per_system(lambda state: state.stress),
        "kinetic_energy": _per_system(lambda state: calc_kinetic_energy(
            masses=state.masses, momenta=state.momenta, system_idx=state.system_idx)),
        "temperature": _per_system(lambda state: state.calc_temperature()),
        "max_force": _per_system(system_wise_max_force),
    }
    calcs = {k: fn for k, fn in possible.items() if k in (properties or ())}
    kw = dict(reporter)
    return TrajectoryReporter(
        prop_calculators=kw.pop("prop_calculators", {prop_frequency: calcs}),
        state_frequency=kw.pop("state_frequency", state_frequency),
        **kw,
    )


def _initial_step(reporter) -> int:
    """1, or one past the last step of appended files (runners.py:108-141)."""
    if reporter is None or reporter.mode != "a":
        return 1
    last = [s if s is not None else 0 for s in reporter.last_steps]
    if len(set(last)) != 1:
        raise ValueError(
            f"Trajectory files have different last steps: {set(last)} Cannot resume "
            "integration from inconsistent states. You can truncate the trajectories to the "
            "same step using:\n\n    reporter.truncate_to_step(min(reporter.last_step))\n\n"
            "before calling integrate again.")
    return 1 + min(last)


def _write_initial_state(reporter, state, model, *, with_props: bool = True) -> None:
    """Step 0 when every file is still empty (runners.py:228-250)."""
    if reporter and all(t.last_step is None for t in reporter.trajectories):
        reporter.report(state, 0, model=model, with_props=with_props)


def _report_steps(reporter) -> list[int]:
    return sorted({f for f in (reporter.state_frequency, *reporter.prop_calculators) if f})


def _normalize_temperature_tensor(temperature, n_steps: int, state: SimState) -> torch.Tensor:
    """(n_steps,) or (n_steps, n_systems) temperatures (runners.py:173-218)."""
    temps = torch.atleast_1d(torch.as_tensor(temperature, dtype=state.dtype, device=state.device))
    if temps.ndim > 2:
        raise ValueError(f"Temperature tensor must be 1D or 2D, got shape {temps.shape}.")
    if temps.numel() == 1:
        return temps.expand(n_steps)
    if state.n_systems == n_steps:
        warnings.warn(
            "n_systems is equal to n_steps. Interpreting temperature array of length "
            "n_systems as temperatures for each system, broadcasted over steps.",
            stacklevel=2,
        )
        return temps.expand(n_steps)
    if temps.shape[0] == state.n_systems:
        if temps.ndim == 2:
            raise ValueError("If temperature tensor is 2D, first dimension must be n_steps.")
        return temps.unsqueeze(0).expand(n_steps, -1)
    if temps.shape[0] == n_steps:
        return temps
    raise ValueError(
        f"Temperature length ({temps.shape[0]}) must be either n_steps ({n_steps}), "
        f"n_systems ({state.n_systems}) or 1 (scalar)."
    )


def integrate(
    system,
    model,
    *,
    integrator: Integrator | tuple[Callable, Callable],
    n_steps: int,
    temperature,
    timestep: float,
    trajectory_reporter=None,
    autobatcher: bool = False,
    pbar: bool = False,
    init_kwargs: dict[str, Any] | None = None,
    engine: bool | None = None,
    skin: float | None = None,
    seed: int | None = None,
    **integrator_kwargs: Any,
):
    """Run MD (runners.py:252-413); see the module docstring for ``engine``."""
    if autobatcher:
        raise NotImplementedError("autobatching is outside this package")
    state = initialize_state(system, model.device, model.dtype)
    kTs = _normalize_temperature_tensor(temperature, n_steps, state) * MetalUnits.temperature
    if isinstance(integrator, Integrator) or integrator in INTEGRATOR_REGISTRY:
        integrator = Integrator(integrator)
        init_fn, step_fn = INTEGRATOR_REGISTRY[integrator]
    elif isinstance(integrator, tuple) and len(integrator) == 2 and all(map(callable, integrator)):
        init_fn, step_fn = integrator
    else:
        raise ValueError(
            "integrator must be key from Integrator or a tuple of (init_func, step_func), "
            f"got {type(integrator)}"
        )
    init_kwargs = {} if init_kwargs is None else dict(init_kwargs)
    dt = float(timestep * MetalUnits.time)
    channels = {"init": init_kwargs, "step": integrator_kwargs}
    for key, factors in INTEGRATOR_KWARG_UNITS.get(integrator, {}).items():
        for channel, factor in factors:
            kw = channels[channel]
            if kw.get(key) is not None:
                kw[key] = kw[key] * factor

    engine_ok = (
        integrator in (Integrator.nve, Integrator.nvt_langevin)
        and isinstance(model, PairPotentialModel)
        and model.fused_kernels
        and state.dtype == torch.float64
        and n_steps > 0
        and bool((kTs == kTs[0]).all())
        and set(integrator_kwargs) <= {"gamma"}
    )
    use_engine = engine_ok if engine is None else bool(engine)
    if use_engine and not engine_ok:
        raise ValueError("this model / integrator / schedule is not supported by the engine")

    reporter = None
    if trajectory_reporter is not None:
        reporter = _configure_reporter(
            trajectory_reporter, properties=["kinetic_energy", "potential_energy", "temperature"])
    first = _initial_step(reporter)
    state = init_fn(state=state, model=model, kT=kTs[0], dt=dt, **init_kwargs)
    _write_initial_state(reporter, state, model)
    if use_engine:
        md = BatchedMD(
            model,
            state,
            dt=dt,
            kT=kTs[0],
            gamma=integrator_kwargs.get("gamma"),
            langevin=integrator == Integrator.nvt_langevin,
            skin=estimate_skin(model.cutoff) if skin is None else skin,
            seed=int(torch.randint(0, 2**62, (1,)).item()) if seed is None else seed,
            tpa=getattr(model, "tpa", 4),
        )
        if reporter is None:
            md.run(n_steps)
            return md.state()
        freqs = _report_steps(reporter)
        step, last = first, first + n_steps - 1
        while step <= last:
            k = min([((step + f - 1) // f) * f for f in freqs] + [last])
            md.run(k - step + 1)
            if any(k % f == 0 for f in freqs):
                reporter.report(md.state(), k, model=model)
            step = k + 1
        out = md.state()
        reporter.finish()
        return out
    for step in range(first, first + n_steps):
        state = step_fn(state=state, model=model, dt=dt, kT=kTs[step - first],
                        **integrator_kwargs)
        if reporter:
            reporter.report(state, step, model=model)
    if reporter:
        reporter.finish()
    return state


def generate_force_convergence_fn(force_tol: float = 1e-1, *, include_cell_forces: bool = False):
    """Per-system ``max |F_i| < force_tol`` (and cell forces) (runners.py:501-538)."""

    def convergence_fn(state, last_energy=None):  # noqa: ARG001
        fmax = getattr(state, "fmax", None)
        f_atoms = fmax[:, 0] if fmax is not None else system_wise_max_force(state)
        conv = f_atoms < force_tol
        if include_cell_forces:
            cf = getattr(state, "cell_forces", None)
            if cf is None:
                raise ValueError("cell_forces not found in state")
            f_cell = fmax[:, 1] if fmax is not None else cf.norm(dim=2).max(dim=1).values
            conv = conv & (f_cell < force_tol)
        return conv

    return convergence_fn


def generate_energy_convergence_fn(energy_tol: float = 1e-3):
    """Per-system ``|E - E_last| < energy_tol`` (runners.py:541-564)."""

    def convergence_fn(state, last_energy):
        return torch.abs(state.energy - last_energy) < energy_tol

    return convergence_fn


class InFlightAutoBatcher:
    """Keeps at most ``max_atoms`` atoms in flight and refills converged slots
    from the queue in input order (the role of autobatching.py:1040-1152;
    memory estimation is not part of this package: the budget is given)."""

    def __init__(self, model=None, *, max_atoms: int, max_iterations: int | None = None):
        self.max_atoms = int(max_atoms)
        self.max_iterations = max_iterations


def _gather_systems(parts, cls, template, counts_new: np.ndarray):
    """A state of class ``cls`` holding, in order, the systems of several
    states: ``parts`` is a sequence of (state, sys_ids (device int64),
    atom_rows (device int64), n_atoms).  Per-atom and per-system tensors are
    device gathers; ``system_idx`` and the system offsets come from the
    host-side ``counts_new`` (no synchronisation)."""
    dev = template.device
    kw: dict[str, Any] = {}
    n_new = int(counts_new.sum())
    cnt = torch.as_tensor(counts_new, device=dev, dtype=torch.int64)
    for f in dc_fields(cls):
        name = f.name
        v0 = getattr(template, name)
        if name == "system_idx":
            kw[name] = torch.repeat_interleave(torch.arange(len(counts_new), device=dev), cnt,
                                               output_size=n_new)
        elif name == "_rng":
            kw[name] = None
        elif name in ("_atom_extras", "_system_extras"):
            kw[name] = {}
        elif isinstance(v0, torch.Tensor) and name in cls._atom_attributes:
            kw[name] = torch.cat([p[0].__dict__[name].index_select(0, p[2]) for p in parts
                                  if p[3]]).contiguous()
        elif isinstance(v0, torch.Tensor) and name in cls._system_attributes:
            kw[name] = torch.cat([p[0].__dict__[name].index_select(0, p[1]) for p in parts
                                  if p[1].numel()]).contiguous()
        else:
            kw[name] = v0
    return cls(**kw)


def _scatter_systems(dst, src, dst_ids, dst_rows, src_ids, src_rows) -> None:
    """dst's systems ``dst_ids`` (atoms ``dst_rows``) <- src's systems
    ``src_ids`` (atoms ``src_rows``), every per-atom and per-system tensor."""
    cls = type(dst)
    for f in dc_fields(cls):
        name = f.name
        v = dst.__dict__.get(name)
        if not isinstance(v, torch.Tensor) or name == "system_idx":
            continue
        if name in cls._atom_attributes:
            v.index_copy_(0, dst_rows, src.__dict__[name].index_select(0, src_rows))
        elif name in cls._system_attributes:
            v.index_copy_(0, dst_ids, src.__dict__[name].index_select(0, src_ids))


def _resident_ok(model, init_fn, step_fn) -> bool:
    return (isinstance(model, PairPotentialModel) and model.fused_kernels
            and init_fn is fire_init and step_fn is fire_step)


def _optimize_resident(state0, model, *, convergence_fn, max_steps, steps_between_swaps,
                       budget, init_kwargs, optimizer_kwargs, stats):
    """In-flight FIRE with a device-resident refill (autobatching.py:1040-1152,
    runners.py:567-753).

    Every structure is FIRE-initialised once, as the reference does before it
    starts swapping (runners.py:650-667): the resulting pool state, with its
    Verlet list, holds each system at its input position.  The batch in
    flight is gathered from the pool; every ``steps_between_swaps`` steps one
    host read of the convergence mask decides the swap: converged systems
    are written back into the pool, and the next batch is gathered on the
    device -- the kept systems from the current batch, then as many queued
    systems as the atom budget allows, in queue order.  Their neighbor-list
    rows move with them (``assemble_verlet``), so a swap rebuilds no list,
    and nothing is concatenated on the host.  The pool ends up as the result,
    in input order."""
    dev = state0.device
    pool = fire_init(state0, model, **init_kwargs)
    n_total = pool.n_systems
    psi = model._system_index(pool)
    counts = psi.counts.astype(np.int64)
    pstart = np.concatenate([[0], np.cumsum(counts)])
    pool_vl = model._vl
    pool_species = model.species_of(pool.atomic_numbers)
    if pool_vl is not None:  # the pool keeps its list; the batch gets its own
        model._vl = VerletList(model.cutoff, model.skin, tpa=model.tpa)
    steps = np.ones(n_total, dtype=np.int64)  # the reference counts from 1 (runners.py:160-162)
    queue = list(range(n_total))

    def take(n_free_atoms: int, *, at_least_one: bool) -> list[int]:
        out, used = [], 0
        while queue:
            c = int(counts[queue[0]])
            if used + c > n_free_atoms and (out or not at_least_one):
                break
            used += c
            out.append(queue.pop(0))
        return out

    def rows_of(start: np.ndarray, cnt: np.ndarray) -> torch.Tensor:
        return segment_index(torch.as_tensor(start, device=dev, dtype=torch.int64),
                             torch.as_tensor(cnt, device=dev, dtype=torch.int64),
                             int(cnt.sum()))

    def assemble(flight, fsi, keep_local: list[int], new: list[int], current: list[int]):
        keep_pool = [current[k] for k in keep_local]
        counts_new = counts[keep_pool + new]
        k_ids = torch.as_tensor(keep_local, device=dev, dtype=torch.int64)
        n_ids = torch.as_tensor(new, device=dev, dtype=torch.int64)
        parts = []
        if keep_local:
            fstart = np.concatenate([[0], np.cumsum(counts[current])])
            kl = np.asarray(keep_local)
            parts.append((flight, k_ids, rows_of(fstart[kl], counts[current][kl]),
                          int(counts[keep_pool].sum())))
        if new:
            na = np.asarray(new)
            parts.append((pool, n_ids, rows_of(pstart[na], counts[na]), int(counts[na].sum())))
        nxt = _gather_systems(parts, type(pool), pool, counts_new)
        si = SystemIndex(nxt.system_idx, len(counts_new))
        si._counts = counts_new.copy()
        object.__setattr__(nxt, "_tsx_sysidx", si)
        model._sysidx = si
        if pool_species is not None:
            sp = []
            if keep_local:
                sp.append(model._species.index_select(0, parts[0][2]))
            if new:
                sp.append(pool_species.index_select(0, parts[-1][2]))
            model._species_key, model._species = TensorKey(nxt.atomic_numbers), torch.cat(sp)
        if pool_vl is not None:
            vparts = []
            if keep_local:
                vparts.append((model._vl, fsi, k_ids, parts[0][3]))
            if new:
                vparts.append((pool_vl, psi, n_ids, parts[-1][3]))
            dst = VerletList(model.cutoff, model.skin, tpa=model.tpa)
            assemble_verlet(dst, vparts, si)
            model._vl = dst
        if keep_local:  # kept systems carry velocities: not a first step
            nxt._fire_velocities_seen = (nxt.velocities, nxt.velocities._version)
        return nxt, si

    current = take(budget, at_least_one=True)
    state, fsi = assemble(None, None, [], current, [])
    n_swaps = n_refills = 0
    while True:
        exceeded = np.zeros(len(current), dtype=bool)
        last_energy = state.energy
        for _ in range(steps_between_swaps):
            last_energy = state.energy
            state = fire_step(state, model, **optimizer_kwargs)
            stats["atom_steps"] += state.n_atoms
            stats["batch_steps"] += 1
            steps[current] += 1
            exceeded = steps[current] > max_steps
            if bool((steps > max_steps).all()):
                warnings.warn(f"All systems have reached the maximum number of steps: {max_steps}.",
                              stacklevel=3)
                break
        conv = convergence_fn(state, last_energy).to("cpu").numpy() | exceeded
        n_swaps += 1
        if not conv.any():
            continue
        conv_local = np.nonzero(conv)[0]
        keep_local = np.nonzero(~conv)[0].tolist()
        # converged systems back into the pool (the result, in input order)
        cur = np.asarray(current)
        fstart = np.concatenate([[0], np.cumsum(counts[cur])])
        done = cur[conv_local]
        _scatter_systems(pool, state, torch.as_tensor(done, device=dev),
                         rows_of(pstart[done], counts[done]),
                         torch.as_tensor(conv_local, device=dev),
                         rows_of(fstart[conv_local], counts[done]))
        kept_atoms = int(counts[cur[keep_local]].sum()) if keep_local else 0
        new = take(budget - kept_atoms, at_least_one=not keep_local) if queue else []
        if not keep_local and not new:
            break
        n_refills += 1
        state, fsi = assemble(state, fsi, keep_local, new, current)
        current = [current[k] for k in keep_local] + new
    stats["swaps"] += n_swaps
    stats["refills"] += n_refills
    # the pool keeps the model's list; later calls on other states rebuild
    model._sysidx = None
    model._species_key = None
    return pool


def optimize(
    system,
    model,
    *,
    optimizer: Optimizer | tuple[Callable, Callable] = Optimizer.fire,
    convergence_fn: Callable | None = None,
    max_steps: int = 10_000,
    steps_between_swaps: int = 5,
    trajectory_reporter=None,
    autobatcher: InFlightAutoBatcher | bool = False,
    pbar: bool = False,  # noqa: ARG001
    init_kwargs: dict[str, Any] | None = None,
    stats: dict | None = None,
    **optimizer_kwargs: Any,
):
    """Relax every system to convergence (runners.py:567-753).

    The returned state holds all systems in input order; systems that hit
    ``max_steps`` are returned as they are (check ``convergence_fn`` on the
    result).  FIRE with one of this package's pair models runs the
    device-resident refill of ``_optimize_resident``; other optimizer pairs
    or models run the same schedule with host-side selection and
    concatenation.  ``stats`` (optional dict) receives the atom-steps, batch
    steps, swap checks and refills of the run."""
    reporter = None
    if trajectory_reporter is not None:
        reporter = _configure_reporter(trajectory_reporter, properties=["potential_energy"])
    if convergence_fn is None:
        convergence_fn = generate_energy_convergence_fn(energy_tol=1e-3)
    state0 = initialize_state(system, model.device, model.dtype)
    if isinstance(optimizer, Optimizer) or optimizer in OPTIM_REGISTRY:
        init_fn, step_fn = OPTIM_REGISTRY[Optimizer(optimizer)]
    elif isinstance(optimizer, tuple) and len(optimizer) == 2 and all(map(callable, optimizer)):
        init_fn, step_fn = optimizer
    else:
        raise TypeError(
            f"Invalid optimizer={type(optimizer).__name__}, must be key from Optimizer or a "
            "tuple of (init_func, step_func)"
        )
    init_kwargs = dict(init_kwargs or {})
    n_total = state0.n_systems
    counts = torch.bincount(state0.system_idx, minlength=n_total).tolist()
    if isinstance(autobatcher, InFlightAutoBatcher):
        budget = autobatcher.max_atoms
    else:
        budget = sum(counts)  # everything in flight (one GPU holds these batches)
    if stats is None:
        stats = {}
    for k in ("atom_steps", "batch_steps", "swaps", "refills"):
        stats.setdefault(k, 0)
    if reporter is None and _resident_ok(model, init_fn, step_fn):
        return _optimize_resident(state0, model, convergence_fn=convergence_fn,
                                  max_steps=max_steps, steps_between_swaps=steps_between_swaps,
                                  budget=budget, init_kwargs=init_kwargs,
                                  optimizer_kwargs=optimizer_kwargs, stats=stats)
    queue = list(range(n_total))
    done: dict[int, SimState] = {}
    steps = torch.ones(n_total, dtype=torch.int64)  # counted from 1 (runners.py:160-162)
    og_files = None
    if reporter is not None:
        og_files = reporter.filenames
        if reporter.mode == "a":  # resume after the files' last steps
            steps += torch.tensor([s or 0 for s in reporter.last_steps], dtype=torch.int64)
        # the input state at step 0 (it has no energies yet)
        _write_initial_state(reporter, state0, model, with_props=False)

    def take(n_free_atoms: int, *, at_least_one: bool) -> list[int]:
        out, used = [], 0
        while queue:
            c = counts[queue[0]]
            if used + c > n_free_atoms and (out or not at_least_one):
                break
            used += c
            out.append(queue.pop(0))
        return out

    current = take(budget, at_least_one=True)
    state = init_fn(state=select_systems(state0, current), model=model, **init_kwargs)
    files_of = None
    while True:
        exceeded = None
        last_energy = state.energy
        if og_files is not None and files_of != current:
            reporter.reopen_trajectories([og_files[i] for i in current])
            files_of = list(current)
        for _ in range(steps_between_swaps):
            last_energy = state.energy
            state = step_fn(state=state, model=model, **optimizer_kwargs)
            if reporter is not None:
                reporter.report(state, steps[current].tolist(), model=model)
            stats["atom_steps"] += state.n_atoms
            stats["batch_steps"] += 1
            steps[current] += 1
            exceeded = steps > max_steps
            if bool(exceeded.all()):
                warnings.warn(f"All systems have reached the maximum number of steps: {max_steps}.",
                              stacklevel=2)
                break
        conv = convergence_fn(state, last_energy).to("cpu")
        conv = conv | exceeded[current]
        stats["swaps"] += 1
        if not bool(conv.any()):
            continue
        conv_local = torch.nonzero(conv).flatten().tolist()
        keep_local = torch.nonzero(~conv).flatten().tolist()
        finished = select_systems(state, conv_local)
        for k, loc in enumerate(conv_local):
            done[current[loc]] = (finished, k)
        kept_atoms = sum(counts[current[loc]] for loc in keep_local)
        new = take(budget - kept_atoms, at_least_one=not keep_local) if queue else []
        if not keep_local and not new:
            if reporter is not None:
                reporter.finish()
            break
        parts = []
        if keep_local:
            parts.append(select_systems(state, keep_local))
        if new:
            parts.append(init_fn(state=select_systems(state0, new), model=model, **init_kwargs))
        state = parts[0] if len(parts) == 1 else concatenate_states(parts)
        current = [current[loc] for loc in keep_local] + new
        stats["refills"] += 1
    # input order
    groups: dict[int, tuple] = {}
    for sys_id, (st, k) in done.items():
        groups.setdefault(id(st), (st, []))[1].append((sys_id, k))
    chunks, order = [], []
    for st, members in groups.values():
        members.sort()
        chunks.append(select_systems(st, [k for _, k in members]))
        order.extend(sys_id for sys_id, _ in members)
    merged = chunks[0] if len(chunks) == 1 else concatenate_states(chunks)
    perm = sorted(range(len(order)), key=lambda j: order[j])
    return select_systems(merged, perm)


__all__ = [
    "InFlightAutoBatcher",
    "MDState",
    "generate_energy_convergence_fn",
    "generate_force_convergence_fn",
    "integrate",
    "optimize",
]
