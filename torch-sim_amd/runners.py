"""High-level entry points with the reference signatures (torch_sim/runners.py).

``integrate`` (runners.py:252-413): unit conversion (temperature in K,
timestep in ps, Langevin gamma in 1/ps), ``init_fn`` then ``n_steps`` calls
of ``step_fn``.  When the model is a closed-form pair model, the integrator
is NVE or Langevin NVT, the temperature is constant over the run and no
reporter is attached, ``engine=None`` (default) routes the run through the
fused, graph-captured :class:`~torch_sim_amd.engine.BatchedMD`; ``engine=False``
keeps the reference's per-step loop.

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

from .engine import BatchedMD, estimate_skin
from .integrators import INTEGRATOR_KWARG_UNITS, INTEGRATOR_REGISTRY, Integrator, MDState
from .models import PairPotentialModel
from .optimizers import OPTIM_REGISTRY, Optimizer
from .quantities import system_wise_max_force
from .state import SimState, concatenate_states, initialize_state, select_systems
from .units import MetalUnits


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
    if trajectory_reporter is not None:
        raise NotImplementedError("trajectory reporting is outside this package")
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
        isinstance(integrator, Integrator)
        and isinstance(model, PairPotentialModel)
        and state.dtype == torch.float64
        and n_steps > 0
        and bool((kTs == kTs[0]).all())
        and set(integrator_kwargs) <= {"gamma"}
    )
    use_engine = engine_ok if engine is None else bool(engine)
    if use_engine and not engine_ok:
        raise ValueError("this model / integrator / schedule is not supported by the engine")

    state = init_fn(state=state, model=model, kT=kTs[0], dt=dt, **init_kwargs)
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
        md.run(n_steps)
        return md.state()
    for step in range(n_steps):
        state = step_fn(state=state, model=model, dt=dt, kT=kTs[step], **integrator_kwargs)
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
    **optimizer_kwargs: Any,
):
    """Relax every system to convergence (runners.py:567-753).

    The returned state holds all systems in input order; systems that hit
    ``max_steps`` are returned as they are (check ``convergence_fn`` on the
    result)."""
    if trajectory_reporter is not None:
        raise NotImplementedError("trajectory reporting is outside this package")
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
    queue = list(range(n_total))
    done: dict[int, SimState] = {}
    steps = torch.zeros(n_total, dtype=torch.int64)

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
    while True:
        exceeded = None
        last_energy = state.energy
        for _ in range(steps_between_swaps):
            last_energy = state.energy
            state = step_fn(state=state, model=model, **optimizer_kwargs)
            steps[current] += 1
            exceeded = steps > max_steps
            if bool(exceeded.all()):
                warnings.warn(f"All systems have reached the maximum number of steps: {max_steps}.",
                              stacklevel=2)
                break
        conv = convergence_fn(state, last_energy).to("cpu")
        conv = conv | exceeded[current]
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
            break
        parts = []
        if keep_local:
            parts.append(select_systems(state, keep_local))
        if new:
            parts.append(init_fn(state=select_systems(state0, new), model=model, **init_kwargs))
        state = parts[0] if len(parts) == 1 else concatenate_states(parts)
        current = [current[loc] for loc in keep_local] + new
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
