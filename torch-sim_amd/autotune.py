"""Autotuner for the fused MD engine: lanes per atom and Verlet skin.

For a given (model, batch) it times short graph-replayed Langevin runs for
each candidate (tpa, skin) and keeps the fastest.  The skin trades list
length (4 k bytes per atom-step in the force kernel, k ~ (rc + skin)^3)
against rebuild frequency, so trials run long enough (``trial_steps``) for
rebuilds to show.  Results are cached in memory and in
``$TSX_AUTOTUNE_CACHE`` (JSON, default ~/.cache/torch_sim_amd/autotune.json)
keyed by model kind, cutoff, atom / system counts and number density.
"""

from __future__ import annotations

import json
import math
import os
import time

import torch

_MEM: dict[str, dict] = {}


def _cache_path() -> str:
    return os.environ.get(
        "TSX_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "torch_sim_amd", "autotune.json")
    )


def _key(state, model) -> str:
    vol = float(torch.det(state.cell).abs().sum())
    rho = state.n_atoms / vol if vol > 0 else 0.0
    return f"k{model.kind}-rc{model.cutoff:.3f}-n{state.n_atoms}-s{state.n_systems}-rho{rho:.4f}"


def _load() -> dict:
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key: str, cfg: dict) -> None:
    data = _load()
    data[key] = cfg
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as fh:
            json.dump(data, fh, indent=1)
    except OSError:
        pass


def tune_md(
    state,
    model,
    *,
    kT_kelvin: float = 1000.0,
    tpas=(2, 4, 8),
    skins=None,
    warmup_steps: int = 40,
    trial_steps: int = 160,
    log=None,
) -> dict:
    """Pick (tpa, skin) for a Langevin run of ``state`` with ``model``."""
    from . import integrators
    from .engine import BatchedMD
    from .units import MetalUnits

    key = _key(state, model)
    if key in _MEM:
        return _MEM[key]
    disk = _load()
    if key in disk:
        _MEM[key] = disk[key]
        return disk[key]
    rc = model.cutoff
    if skins is None:
        skins = tuple(round(f * rc, 3) for f in (0.05, 0.08, 0.10, 0.11, 0.12, 0.16))
    kT = kT_kelvin * MetalUnits.temperature
    dt = 0.002 * MetalUnits.time
    st = state.clone()
    st.rng = 0
    md = integrators.nvt_langevin_init(state=st, model=model, kT=kT)
    def rate_of(skin, tpa, steps):
        eng = BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=skin, tpa=tpa, seed=1,
                        chunk_steps=steps)
        eng.run(warmup_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(steps)
        torch.cuda.synchronize()
        return state.n_atoms * steps / (time.perf_counter() - t0)

    trials = []
    for skin in skins:
        for tpa in tpas:
            rate = rate_of(skin, tpa, trial_steps)
            if log:
                log(f"  autotune skin={skin} tpa={tpa}: {rate:.3e} atom-steps/s")
            trials.append((rate, skin, tpa))
    # the two fastest again, over twice the steps: neighbouring skins are
    # often within the noise of one short trial
    trials.sort(reverse=True)
    best = None
    for _, skin, tpa in trials[:2]:
        rate = rate_of(skin, tpa, 2 * trial_steps)
        if log:
            log(f"  autotune re-timed skin={skin} tpa={tpa}: {rate:.3e} atom-steps/s")
        if best is None or rate > best[0]:
            best = (rate, skin, tpa)
    cfg = {"skin": best[1], "tpa": best[2], "rate": best[0]}
    _MEM[key] = cfg
    _store(key, cfg)
    return cfg


def default_tpa(mean_neighbors: float) -> int:
    """Heuristic lanes per atom when no tuning data exists."""
    return 8 if mean_neighbors > 120 else (4 if mean_neighbors > 40 else 2)


__all__ = ["tune_md", "default_tpa", "math"]
