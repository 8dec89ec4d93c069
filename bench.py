"""Benchmark: MD atom-steps/s on one MI355X, batched (graph-captured) vs plain.

Workload (default): 512 periodic Cu fcc systems with random repeats (32-2048
atoms each, ~4.0e5 atoms), positions jittered N(0, 0.05 A), Morse
D = 0.3429 eV, alpha = 1.3588 1/A, r0 = 2.866 A, cutoff 6 A, Langevin NVT
T = 1000 K, gamma = 10 / ps, dt = 2 fs.  FP32 pair arithmetic, FP64
accumulation / integration, synthetic seeded data.

Legs reported on one JSON line:
  value        batched fused engine, hipGraph replay, K timed steps
  plain_gpu    the per-step path (nvt_langevin_step + model(state) with a
               neighbor list rebuilt every step) on the same GPU
  roofline     fused-step kernel: algorithmic bytes / measured duration
  cpu_baseline the FP64 CPU oracle (reference algorithm) on a bounded sample
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import torch_sim_amd as tsa  # noqa: E402

HBM_PEAK_GBS = 8000.0
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)

# HBM bytes per dispatch of the fused-step kernel, from two rocprofv3 --pmc runs of this
# script (FETCH_SIZE, WRITE_SIZE in KiB per dispatch, averaged over the dispatches), each
# scaled by what the same counter reads for a 256 MiB device copy (FETCH_SIZE x 2.000,
# WRITE_SIZE x 1.000).  Keyed by (workload, skin, tpa); the summaries with their commands
# are profiles/k_md_fused__cu_morse_nvt_512sys.md and profiles/k_md_fused__lj_fluid_1m_nve.md.
MEASURED_TRAFFIC = {
    ("cu_morse_nvt", 0.6, 2): {"fetch_kib": 111967.1, "fetch_factor": 2.000,
                               "write_kib": 41957.1, "write_factor": 1.000},
    ("cu_morse_nvt", 0.66, 2): {"fetch_kib": 117702.1, "fetch_factor": 2.000,
                                "write_kib": 41769.9, "write_factor": 1.000},
    ("cu_morse_nvt", 0.72, 2): {"fetch_kib": 123122.7, "fetch_factor": 2.000,
                                "write_kib": 41832.4, "write_factor": 1.000},
    ("lj1m_nve", 0.3, 4): {"fetch_kib": 342936.1, "fetch_factor": 2.000,
                           "write_kib": 125831.8, "write_factor": 1.000},
}


def measured_traffic(workload: str, skin: float, tpa: int):
    meas = MEASURED_TRAFFIC.get((workload, round(skin, 3), tpa))
    if meas is None:
        return None
    return 1024.0 * (meas["fetch_kib"] * meas["fetch_factor"] + meas["write_kib"] * meas["write_factor"])


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_engine(state, model, skin, tpa, seed=1):
    kT = 1000.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    state.rng = seed
    md = tsa.nvt_langevin_init(state=state, model=model, kT=kT)
    eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, gamma=gamma, langevin=True, skin=skin, tpa=tpa,
                        seed=seed, chunk_steps=500)
    return md, eng


def time_fused_kernel(eng, n_launch=50, n_warm=5):
    """Average duration of the fused-step kernel as it runs inside the step:
    eager steps as the engine's run takes them (gated rebuild, fused kernel,
    and the reduction where it has a launch of its own), with hipEvents on
    the kernel's stream around each fused launch.  The rebuild queued ahead
    of every launch keeps the host ahead of the GPU, so each event pair
    brackets the kernel alone, in the cache state of the real step."""
    st = torch.cuda.current_stream()
    eng._open()
    pairs = []
    for k in range(n_warm + n_launch):
        q = eng.cur
        eng._rebuild(q)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        eng._fused(q, end_kick=True, do_next=True)
        b.record(st)
        if not eng._merge_reduce:
            eng._reduce(q)
        eng.cur = 1 - q
        if k >= n_warm:
            pairs.append((a, b))
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in pairs) / len(pairs)


def plain_gpu(state, model, steps, warmup):
    kT = 1000.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    st = state.clone()
    st.rng = 3
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    for _ in range(warmup):
        md = tsa.nvt_langevin_step(md, model, dt=dt, kT=kT, gamma=gamma)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        md = tsa.nvt_langevin_step(md, model, dt=dt, kT=kT, gamma=gamma)
    torch.cuda.synchronize()
    return st.n_atoms * steps / (time.perf_counter() - t0)


def _torch_pairs(pos, cell_col, system_idx, n_sys, rc):
    """Full neighbor list (i, j, S) with |r_j - r_i + S.L| < rc by torch ops:
    the reference's linked-cell idea (torch_nl.py:156-215, transforms.py:
    843-1006) restated for a GPU batch: wrap, bin into cells of edge >= rc,
    sort by (system, bin), expand the 27-bin stencil of every atom into
    candidates, keep those within rc.  Needs every cell width >= rc."""
    dev = pos.device
    row = cell_col.transpose(1, 2)  # row vectors a_k
    inv = torch.linalg.inv(row)
    frac = torch.bmm(pos.unsqueeze(1), inv[system_idx]).squeeze(1)
    wrap = torch.floor(frac)
    frac = frac - wrap
    width = 1.0 / torch.linalg.norm(inv, dim=1)  # plane spacings [S,3]
    nb = torch.clamp(torch.floor(width / rc), min=1).to(torch.int64)  # [S,3]
    ib = torch.clamp((frac * nb[system_idx]).to(torch.int64), max=nb[system_idx] - 1)
    nbin_s = nb.prod(1)
    off = torch.cumsum(nbin_s, 0) - nbin_s
    nbn = nb[system_idx]
    key = off[system_idx] + (ib[:, 0] * nbn[:, 1] + ib[:, 1]) * nbn[:, 2] + ib[:, 2]
    order = torch.argsort(key, stable=True)
    skey = key[order]
    total_bins = int(nbin_s.sum())
    bins = torch.arange(total_bins + 1, device=dev)
    bstart = torch.searchsorted(skey, bins)
    d = torch.tensor([-1, 0, 1], device=dev)
    sten = torch.stack(torch.meshgrid(d, d, d, indexing="ij"), -1).reshape(27, 3)
    out_i, out_j, out_s = [], [], []
    n = pos.shape[0]
    for a0 in range(0, n, 32768):  # atom chunks bound the candidate tensors
        a1 = min(n, a0 + 32768)
        ia = torch.arange(a0, a1, device=dev)
        sa = system_idx[a0:a1]
        cb = ib[a0:a1].unsqueeze(1) + sten.unsqueeze(0)  # [n,27,3]
        nbs = nb[sa].unsqueeze(1)
        shift = torch.div(cb, nbs, rounding_mode="floor")
        cbw = cb - shift * nbs
        nkey = off[sa].unsqueeze(1) + (cbw[..., 0] * nbs[..., 1] + cbw[..., 1]) * nbs[..., 2] + cbw[..., 2]
        lo, hi = bstart[nkey], bstart[nkey + 1]
        cnt = (hi - lo).reshape(-1)
        src = torch.repeat_interleave(torch.arange(cnt.numel(), device=dev), cnt)
        first = torch.cumsum(cnt, 0) - cnt
        k = lo.reshape(-1)[src] + (torch.arange(src.numel(), device=dev) - first[src])
        i = ia[torch.div(src, 27, rounding_mode="floor")]
        j = order[k]
        s_img = shift.reshape(-1, 3)[src].to(pos.dtype) + wrap[i] - wrap[j]  # unwrapped shifts
        si = system_idx[i]
        sl = (s_img[:, 0:1] * row[si, 0] + s_img[:, 1:2] * row[si, 1] + s_img[:, 2:3] * row[si, 2])
        dr = pos[j] - pos[i] + sl
        keep = ((dr * dr).sum(1) < rc * rc) & ~((i == j) & (s_img == 0).all(1))
        out_i.append(i[keep])
        out_j.append(j[keep])
        out_s.append(s_img[keep])
    return torch.cat(out_i), torch.cat(out_j), torch.cat(out_s)


def plain_torch(state, steps, warmup):
    """The reference pipeline restated in PyTorch ops on the GPU (a torch
    linked-cell list rebuilt every step, autograd forces of the Morse pair
    energy, BAOAB in torch ops; no HIP kernels, no graph capture)."""
    kT = 1000.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    dev = state.device
    x = state.positions.clone()
    m = state.masses.unsqueeze(-1)
    idx, cell, n_sys = state.system_idx, state.cell, state.n_systems
    D, a, r0 = MORSE["epsilon"], MORSE["alpha"], MORSE["sigma"]
    g = torch.Generator(device=dev).manual_seed(7)

    def forces(xx):
        i, j, s_img = _torch_pairs(xx.detach(), cell, idx, n_sys, 6.0)
        xr = xx.detach().requires_grad_(True)
        row = cell.transpose(1, 2)
        si = idx[i]
        dr = xr[j] - xr[i] + (s_img[:, 0:1] * row[si, 0] + s_img[:, 1:2] * row[si, 1]
                              + s_img[:, 2:3] * row[si, 2])
        r = dr.norm(dim=1)
        e = 0.5 * (D * (1.0 - torch.exp(-a * (r - r0))) ** 2 - D).sum()
        (f,) = torch.autograd.grad(e, xr)
        return -f

    p = torch.randn(x.shape, generator=g, device=dev, dtype=x.dtype) * torch.sqrt(m * kT)
    f = forces(x)
    c1 = math.exp(-gamma * dt)
    c2 = math.sqrt(kT * (1 - c1 * c1))

    def step(x, p, f):
        p = p + 0.5 * dt * f
        x = x + 0.5 * dt * p / m
        p = c1 * p + c2 * torch.sqrt(m) * torch.randn(x.shape, generator=g, device=dev, dtype=x.dtype)
        x = x + 0.5 * dt * p / m
        f = forces(x)
        return x, p + 0.5 * dt * f, f

    for _ in range(warmup):
        x, p, f = step(x, p, f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        x, p, f = step(x, p, f)
    torch.cuda.synchronize()
    return state.n_atoms * steps / (time.perf_counter() - t0)


def cpu_baseline(n_sys=32, min_seconds=10.0, max_steps=200, seed=0):
    """FP64 CPU oracle (reference algorithm: linked-cell list rebuilt every
    step, autograd half-list forces, BAOAB) on the first ``n_sys`` systems of
    the same workload, stepped until ``min_seconds`` of CPU work have run (a
    bounded sample of the workload)."""
    from oracle import md as omd
    from oracle import models as om
    from oracle import nl as onl

    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    torch.set_num_threads(cores)
    st = tsa.workloads.cu_batch(512, device="cpu", seed=seed)
    cnt = torch.bincount(st.system_idx)
    n_atoms = int(cnt[:n_sys].sum())
    pos = st.positions[:n_atoms].clone()
    idx = st.system_idx[:n_atoms].clone()
    cell = st.cell[:n_sys].clone()
    m = st.masses[:n_atoms].clone()
    z = st.atomic_numbers[:n_atoms].clone()
    pbc = st.pbc
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cell, pbc, z, idx, fn, 6.0, compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    kT = 1000.0 * omd.KB_EV
    dt = 0.002 * omd.TIME_UNIT
    gamma = 10.0 / omd.TIME_UNIT
    g = torch.Generator().manual_seed(0)
    p = omd.initialize_momenta(m, idx, torch.tensor(kT, dtype=torch.float64), torch.randn(n_atoms, 3, generator=g, dtype=torch.float64))
    e, f = force(pos)  # warm-up / init
    x = pos
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps and (steps < 2 or time.perf_counter() - t0 < min_seconds):
        noise = torch.randn(n_atoms, 3, generator=g, dtype=torch.float64)
        x, p, f, e = omd.langevin_step(x, p, f, m, idx, force, dt, kT, gamma, noise)
        steps += 1
    el = time.perf_counter() - t0
    return {
        "value": n_atoms * steps / el,
        "unit": "atom-steps/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"first {n_sys} systems ({n_atoms} atoms) of the 512-system batch, {steps} "
        f"BAOAB steps, FP64 CPU oracle (linked-cell list every step, autograd forces), "
        f"{el:.1f} s, cpu={platform.processor() or platform.machine()}",
    }


def bench_cu_morse(args) -> dict:
    """512-system Cu Morse NVT Langevin (the headline workload)."""
    dev = torch.device("cuda:0")
    state = tsa.workloads.cu_batch(args.n_systems, device=dev, seed=0)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    n = state.n_atoms
    log(f"workload: {args.n_systems} Cu systems, {n} atoms")

    from torch_sim_amd import autotune

    if args.skin is None or args.tpa is None:
        cfg = autotune.tune_md(state, model, log=log)
        skin = args.skin if args.skin is not None else cfg["skin"]
        tpa = args.tpa if args.tpa is not None else cfg["tpa"]
    else:
        skin, tpa = args.skin, args.tpa
    log(f"engine config: skin={skin} tpa={tpa}")

    md, eng = make_engine(state, model, skin, tpa)
    eng.run(args.warmup)
    torch.cuda.synchronize()
    reb0 = eng.n_rebuilds
    t0 = time.perf_counter()
    eng.run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n * args.steps / el
    n_reb = eng.n_rebuilds - reb0
    log(f"batched: {value:.4e} atom-steps/s ({1e3 * el / args.steps:.4f} ms/step), rebuilds {n_reb}")

    k_mean = eng.mean_neighbors
    ms_kernel = time_fused_kernel(eng)
    bytes_launch = eng.algorithmic_bytes_per_atom_step() * n
    achieved = bytes_launch / (ms_kernel * 1e-3) / 1e9
    log(f"fused kernel: {ms_kernel * 1e3:.1f} us, k_list {k_mean:.1f}, {achieved:.0f} GB/s")

    traffic = measured_traffic("cu_morse_nvt", skin, tpa) if args.n_systems == 512 else None

    plain = plain_hip = None
    if not args.no_plain:
        plain = plain_torch(state, args.plain_steps, 1)
        log(f"plain per-step GPU path (PyTorch ops): {plain:.4e} atom-steps/s")
        plain_hip = plain_gpu(state, model, args.plain_steps, 1)
        log(f"per-step HIP kernels, no graph, list every step: {plain_hip:.4e} atom-steps/s")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_systems)
        log(f"cpu baseline: {cpu['value']:.4e} atom-steps/s")

    out = {
        "metric": "MD atom-steps/s at 1 MI355X, batched graph-captured Morse NVT Langevin",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pair/f64-accum",
        "data": "synthetic (seeded Cu fcc batch, Maxwell-Boltzmann momenta)",
        "config": {
            "workload": "cu_morse_nvt_langevin_512sys",
            "n_systems": args.n_systems,
            "n_atoms": n,
            "atoms_per_system": "32-2048",
            "cutoff_A": 6.0,
            "skin_A": skin,
            "tpa": tpa,
            "dt_fs": 2.0,
            "T_K": 1000.0,
            "gamma_per_ps": 10.0,
            "list_rebuilds_in_timed_steps": n_reb,
            "mean_list_neighbors": round(k_mean, 2),
        },
        "plain_gpu": None if plain is None else {
            "value": plain, "unit": "atom-steps/s", "steps": args.plain_steps,
            "path": "reference pipeline in PyTorch ops on the GPU: torch linked-cell list every "
                    "step, autograd Morse forces, BAOAB in torch ops, no graph"},
        "batched_vs_plain": None if plain is None else value / plain,
        "per_step_hip": None if plain_hip is None else {
            "value": plain_hip, "unit": "atom-steps/s", "steps": args.plain_steps,
            "path": "nvt_langevin_step + model(state): HIP kernels per step, list rebuilt every "
                    "step, no graph capture"},
        "batched_vs_per_step_hip": None if plain_hip is None else value / plain_hip,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_md_fused (tsx_md_fused_step)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "bytes_per_launch": bytes_launch,
            "kernel_ms": ms_kernel,
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    return out


def _line(metric, value, unit, steps, warmup, ms, config, **extra) -> dict:
    out = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-pair/f64-accum", "data": "synthetic (seeded)",
        "config": config,
    }
    out.update(extra)
    return out


def bench_lj1m(args) -> dict:
    """1,048,576-atom LJ fluid, NVE, reduced units (rho* 0.8442, T* 0.72, rc 2.5,
    skin 0.3, dt 0.005 tau)."""
    dev = torch.device("cuda:0")
    state = tsa.workloads.lj_fluid(64, device=dev, rho=0.8442, jitter=0.05, seed=0)
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev)
    state.rng = 0
    md = tsa.nve_init(state=state, model=model, kT=0.72)
    tpa = args.tpa or 4
    eng = tsa.BatchedMD(model, md, dt=0.005, langevin=False, skin=args.skin or 0.3, tpa=tpa,
                        seed=1, chunk_steps=250)
    n = state.n_atoms
    eng.run(args.warmup)
    torch.cuda.synchronize()
    reb0 = eng.n_rebuilds
    t0 = time.perf_counter()
    eng.run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n * args.steps / el
    ms_kernel = time_fused_kernel(eng)
    bytes_launch = eng.algorithmic_bytes_per_atom_step() * n
    achieved = bytes_launch / (ms_kernel * 1e-3) / 1e9
    log(f"lj1m: {value:.4e} atom-steps/s, fused kernel {ms_kernel * 1e3:.1f} us, {achieved:.0f} GB/s")
    return _line(
        "MD atom-steps/s at 1 MI355X, 1M-atom LJ NVE", value, "atom-steps/s", args.steps,
        args.warmup, 1e3 * el / args.steps,
        {"workload": "lj_fluid_1m_nve", "n_atoms": n, "rho": 0.8442, "T": 0.72, "cutoff": 2.5,
         "skin": args.skin or 0.3, "tpa": tpa, "dt_tau": 0.005,
         "list_rebuilds_in_timed_steps": eng.n_rebuilds - reb0,
         "mean_list_neighbors": round(eng.mean_neighbors, 2)},
        roofline={"bound": "hbm", "kernel": "k_md_fused (tsx_md_fused_step)", "achieved": achieved,
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                  "bytes_per_launch": bytes_launch, "kernel_ms": ms_kernel,
                  "traffic": measured_traffic("lj1m_nve", args.skin or 0.3, tpa)},
    )


def bench_ar256(args) -> dict:
    """Ar fcc 4x4x4 (256 atoms) LJ NVE, 50 K, dt 2 fs, 100 steps; the FP64 CPU
    oracle runs the same 100 steps as the baseline."""
    from oracle import md as omd
    from oracle import models as om
    from oracle import nl as onl

    dev = torch.device("cuda:0")
    state = tsa.workloads.ar_fcc(device=dev)
    model = tsa.LennardJonesModel(sigma=3.405, epsilon=0.0104, cutoff=8.5125, device=dev)
    kT = 50.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    state.rng = 0
    md = tsa.nve_init(state=state, model=model, kT=kT)
    eng = tsa.BatchedMD(model, md, dt=dt, langevin=False, skin=1.0, tpa=8, seed=1,
                        chunk_steps=100)
    eng.run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = 256 * args.steps / el
    cpu = None
    if not args.no_cpu:
        cst = tsa.workloads.ar_fcc(device="cpu")
        fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 3.405, 0.0104)  # noqa: E731

        def force(x):
            o = om.pair_forward(x, cst.cell, cst.pbc, cst.atomic_numbers, cst.system_idx, fn,
                                8.5125, compute_stress=False, nl=onl.linked_cell_nl)
            return o["energy"], o["forces"]

        torch.set_num_threads(os.cpu_count() or 1)
        g = torch.Generator().manual_seed(0)
        p = omd.initialize_momenta(cst.masses, cst.system_idx, torch.tensor(50.0 * omd.KB_EV),
                                   torch.randn(256, 3, generator=g, dtype=torch.float64))
        x = cst.positions.clone()
        e, f = force(x)
        x, p, f, e = omd.nve_step(x, p, f, cst.masses, force, 0.002 * omd.TIME_UNIT)  # warm-up
        tc = time.perf_counter()
        for _ in range(100):
            x, p, f, e = omd.nve_step(x, p, f, cst.masses, force, 0.002 * omd.TIME_UNIT)
        elc = time.perf_counter() - tc
        cpu = {"value": 256 * 100 / elc, "unit": "atom-steps/s", "cores": torch.get_num_threads(),
               "kind": "port", "sample": f"full 100 NVE steps, FP64 CPU oracle, {elc:.2f} s, "
               f"cpu={platform.processor() or platform.machine()}"}
    log(f"ar256: {value:.4e} atom-steps/s")
    return _line("MD atom-steps/s at 1 MI355X, Ar-256 LJ NVE", value, "atom-steps/s", args.steps,
                 args.warmup, 1e3 * el / args.steps,
                 {"workload": "ar_fcc_256_nve", "n_atoms": 256, "T_K": 50.0, "dt_fs": 2.0,
                  "cutoff_A": 8.5125}, cpu_baseline=cpu)


def bench_fire(args) -> dict:
    """1024 binary soft-sphere systems (64-1024 atoms, phi 0.64, triclinic
    strain <= 5 %), ASE-FIRE with the Frechet cell filter: FIRE steps of the
    whole batch, and how many systems reach fmax < 1e-4 (atoms and cell)."""
    dev = torch.device("cuda:0")
    state = tsa.workloads.soft_sphere_batch(args.n_systems or 1024, n_min=64, n_max=1024,
                                            device=dev, seed=0)
    sig = torch.tensor([[1.0, 1.2], [1.2, 1.4]], dtype=torch.float64)
    skin = args.skin if args.skin is not None else 0.3
    model = tsa.SoftSphereMultiModel(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=sig,
                                     device=dev, compute_stress=True, skin=skin)
    fs = tsa.fire_init(state, model, cell_filter="frechet")
    for _ in range(args.warmup):
        tsa.fire_step(fs, model)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tsa.fire_step(fs, model)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    n = state.n_atoms
    conv = int(((fs.fmax[:, 0] < 1e-4) & (fs.fmax[:, 1] < 1e-4)).sum())
    value = n * args.steps / el
    log(f"fire: {value:.4e} atom-steps/s, {conv} systems converged after {args.warmup + args.steps} steps")
    return _line("FIRE atom-steps/s at 1 MI355X, 1024-system soft-sphere FIRE + Frechet", value,
                 "atom-steps/s", args.steps, args.warmup, 1e3 * el / args.steps,
                 {"workload": "soft_sphere_fire_frechet_1024", "n_systems": state.n_systems,
                  "n_atoms": n, "phi": 0.64, "fmax": 1e-4, "skin": skin,
                  "full_list_builds": model._vl.full_builds if model._vl else None,
                  "converged_after_steps": conv, "steps_taken": args.warmup + args.steps})


def bench_fit(args) -> dict:
    """256 binary LJ systems (128-512 atoms, rho* 0.8, Kob-Andersen sigma /
    epsilon): gradient of force-MSE (targets from parameters scaled by 1.03 /
    0.95) + final potential energy after 200 NVE steps (dt 0.002 tau,
    checkpoint every 20)."""
    from torch_sim_amd import autograd as tad

    dev = torch.device("cuda:0")
    state = tsa.workloads.binary_lj_batch(args.n_systems or 256, n_min=128, n_max=512, rho=0.8,
                                          device=dev, seed=0)
    sig0 = torch.tensor([[1.0, 0.8], [0.8, 0.88]], dtype=torch.float64)
    eps0 = torch.tensor([[1.0, 1.5], [1.5, 0.5]], dtype=torch.float64)
    model = tsa.MultiLennardJonesModel(torch.tensor([1, 2]), sig0, eps0, device=dev, cutoff=2.5)
    skin = args.skin if args.skin is not None else 0.3
    terms = tad.PairTerms(model, skin=skin)
    geo = tad.geometry(terms, state)
    zero = torch.zeros(2, 2, dtype=torch.float64, device=dev)
    with torch.no_grad():
        target = tad.pair_forces(state.positions, state.cell,
                                 tad.pair_params(sig0.to(dev) * 1.03, eps0.to(dev) * 0.95, zero), geo)
    g = torch.Generator(device=dev).manual_seed(0)
    p0 = torch.randn(state.positions.shape, generator=g, device=dev, dtype=torch.float64)
    n = state.n_atoms
    n_md = args.fit_md_steps

    def grad_eval(with_md: bool):
        sig = sig0.to(dev).clone().requires_grad_(True)
        eps = eps0.to(dev).clone().requires_grad_(True)
        params = tad.pair_params(sig, eps, zero)
        f = tad.pair_forces(state.positions, state.cell, params, geo)
        loss = ((f - target) ** 2).mean()
        if with_md:
            xn, _ = tad.nve_trajectory(state.positions, p0, params, geo, state.masses, state.cell,
                                       dt=0.002, n_steps=n_md, checkpoint_every=20)
            loss = loss + tad.pair_energy(xn, state.cell, params, geo).sum() / n
        loss.backward()
        return sig.grad, eps.grad

    def timed(with_md, k, w):
        for _ in range(w):
            grad_eval(with_md)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            grad_eval(with_md)
        torch.cuda.synchronize()
        return k / (time.perf_counter() - t0)

    fm = timed(False, 200, 10)  # ~0.5 ms each: enough of them to be above host jitter
    full = timed(True, args.steps, min(args.warmup, 1))
    log(f"fit: force-matching {fm:.2f} grad-evals/s, full loss ({n_md} NVE steps) {full:.3f} grad-evals/s")
    return _line("grad-evals/s at 1 MI355X, 256-system binary-LJ fit (force MSE + E after NVE)",
                 full, "grad-evals/s", args.steps, min(args.warmup, 1), 1e3 / full,
                 {"workload": "binary_lj_fit_256", "n_systems": state.n_systems, "n_atoms": n,
                  "nve_steps": n_md, "checkpoint_every": 20, "dt_tau": 0.002, "skin": skin,
                  "force_matching_only_grad_evals_per_s": fm})


def calibrate_copy(mib: int) -> None:
    """Three device-to-device copies of ``mib`` MiB: each reads and writes
    exactly that many bytes, so a byte counter read under rocprofv3 --pmc
    divided by the expected KiB is the factor applied to the counters of the
    measured kernels."""
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x3", flush=True)


WORKLOADS = {
    "cu_morse_nvt": (bench_cu_morse, 1000, 50),
    "lj1m_nve": (bench_lj1m, 500, 20),
    "ar256_nve": (bench_ar256, 100, 10),
    "fire_ss_1024": (bench_fire, 100, 10),
    "lj_fit_256": (bench_fit, 3, 1),
}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cu_morse_nvt")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--n-systems", type=int, default=None)
    ap.add_argument("--skin", type=float, default=None)
    ap.add_argument("--tpa", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-systems", type=int, default=32)
    ap.add_argument("--fit-md-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=int, default=None, metavar="MIB",
                    help="only run three device copies of MIB MiB (a kernel of known bytes "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy is not None:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; runs on more than one GPU are not part of this build "
              "(independent systems shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)
    fn, steps, warmup = WORKLOADS[args.workload]
    args.steps = steps if args.steps is None else args.steps
    args.warmup = warmup if args.warmup is None else args.warmup
    if args.workload == "cu_morse_nvt" and args.n_systems is None:
        args.n_systems = 512
    print(json.dumps(fn(args)), flush=True)


if __name__ == "__main__":
    main()
