"""High-level entry points with the reference signatures (torch_sim/runners.py).

``integrate`` (runners.py:252-413): unit conversion (temperature in K,
timestep in ps, Langevin gamma in 1/ps), ``init_fn`` then ``n_steps`` calls
of ``step_fn``.  When the model is a closed-form pair model, the integrator
is NVE or Langevin NVT, the temperature is constant over the run and no
reporter is attached, ``engine=None`` (default) routes the run through the
fused, graph-captured :class:`~torch_sim_amd.engine.BatchedMD`; ``engine=False``
keeps the reference's per-step loop.

``optimize`` (runners.py:567-753) is provided by :mod:`torch_sim_amd.optimizers`.
"""

from __future__ import annotations

import warnings
from typing import Any, Callable

import torch

from .engine import BatchedMD, estimate_skin
from .integrators import INTEGRATOR_KWARG_UNITS, INTEGRATOR_REGISTRY, Integrator, MDState
from .models import PairPotentialModel
from .state import SimState, initialize_state
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


__all__ = ["integrate", "MDState"]
