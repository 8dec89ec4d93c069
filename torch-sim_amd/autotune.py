"""Autotuner for the fused MD engine: lanes per atom, Verlet skin and cell size.

For a given (model, batch) it times short graph-replayed Langevin runs and
keeps the fastest configuration:

* lanes per atom ``tpa`` (the width of an atom's tile in the pair kernel)
  and the skin, over their grid at two bins per list radius;
* then the cell-list bin size (``bins_per_rlist`` 1, 2 or 3 bins per list
  radius) for the two fastest (tpa, skin) pairs, each re-timed over twice
  the steps.

The skin trades list length (4 k bytes per atom-step in the force kernel,
k ~ (rc + skin)^3) against rebuild frequency, so trials run long enough
(``trial_steps``) for rebuilds to show; the bin size trades the fill's
stencil volume against its per-bin work.  Results are cached in memory and
in ``$TSX_AUTOTUNE_CACHE`` (JSON, default
~/.cache/torch_sim_amd/autotune.json) keyed by everything that changes the
outcome: model kind and parameters, cutoff, atom / system counts, number
density, temperature, the search grid, the library build and the GPU.
"""

from __future__ import annotations

import hashlib
import json
import math
import os
import time

import torch

from . import _lib

_MEM: dict[str, dict] = {}


def _cache_path() -> str:
    return os.environ.get(
        "TSX_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "torch_sim_amd", "autotune.json")
    )


def _library_id() -> str:
    """Version, size and content hash of the loaded libtsx.so (a pick made
    with other kernels is not reused)."""
    path = _lib._LIB_PATH
    h = hashlib.sha1()
    with open(path, "rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 20), b""):
            h.update(chunk)
    return f"v{_lib.lib().tsx_version()}-{h.hexdigest()[:12]}"


def _key(state, model, kT_kelvin, tpas, skins, bins) -> str:
    vol = float(torch.det(state.cell).abs().sum())
    rho = state.n_atoms / vol if vol > 0 else 0.0
    params = ",".join(f"{v:.6g}" for v in model._params.flatten().tolist())
    grid = (f"tpa{'/'.join(map(str, sorted(tpas)))}-skin{'/'.join(f'{x:g}' for x in sorted(skins))}"
            f"-bins{'/'.join(map(str, sorted(bins)))}")
    dev = torch.cuda.get_device_name(state.device) if state.device.type == "cuda" else "cpu"
    return (f"k{model.kind}-p{params}-rc{model.cutoff:.3f}-n{state.n_atoms}-s{state.n_systems}"
            f"-rho{rho:.4f}-T{kT_kelvin:g}-{grid}-{_library_id()}-{dev}")


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
    bins=(1, 2, 3),
    warmup_steps: int = 40,
    trial_steps: int = 160,
    log=None,
) -> dict:
    """Pick (tpa, skin, bins_per_rlist) for a Langevin run of ``state`` with
    ``model``; returns {"skin", "tpa", "bins_per_rlist", "rate"}."""
    from . import integrators
    from .engine import BatchedMD
    from .units import MetalUnits

    rc = model.cutoff
    if skins is None:
        skins = tuple(round(f * rc, 3) for f in (0.05, 0.08, 0.10, 0.11, 0.12, 0.16))
    key = _key(state, model, kT_kelvin, tpas, skins, bins)
    if key in _MEM:
        return _MEM[key]
    disk = _load()
    if key in disk:
        _MEM[key] = disk[key]
        return disk[key]
    kT = kT_kelvin * MetalUnits.temperature
    dt = 0.002 * MetalUnits.time
    st = state.clone()
    st.rng = 0
    md = integrators.nvt_langevin_init(state=st, model=model, kT=kT)
    def rate_of(skin, tpa, steps, nb=2):
        eng = BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=skin, tpa=tpa, seed=1,
                        chunk_steps=steps, bins_per_rlist=nb)
        eng.run(warmup_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(steps)
        torch.cuda.synchronize()
        return state.n_atoms * steps / (time.perf_counter() - t0)

    nb0 = 2 if 2 in bins else bins[0]
    trials = []
    for skin in skins:
        for tpa in tpas:
            rate = rate_of(skin, tpa, trial_steps, nb0)
            if log:
                log(f"  autotune skin={skin} tpa={tpa} bins={nb0}: {rate:.3e} atom-steps/s")
            trials.append((rate, skin, tpa))
    # the two fastest again over twice the steps (neighbouring skins are often
    # within the noise of one short trial), each at every bin size
    trials.sort(reverse=True)
    best = None
    for _, skin, tpa in trials[:2]:
        for nb in bins:
            rate = rate_of(skin, tpa, 2 * trial_steps, nb)
            if log:
                log(f"  autotune re-timed skin={skin} tpa={tpa} bins={nb}: {rate:.3e} atom-steps/s")
            if best is None or rate > best[0]:
                best = (rate, skin, tpa, nb)
    cfg = {"skin": best[1], "tpa": best[2], "bins_per_rlist": best[3], "rate": best[0]}
    _MEM[key] = cfg
    _store(key, cfg)
    return cfg


def default_tpa(mean_neighbors: float) -> int:
    """Heuristic lanes per atom when no tuning data exists."""
    return 8 if mean_neighbors > 120 else (4 if mean_neighbors > 40 else 2)


__all__ = ["tune_md", "default_tpa", "math"]
