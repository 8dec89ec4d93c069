"""Benchmark: MD atom-steps/s on one MI355X, batched (graph-captured) vs plain.

Headline workload: 512 periodic Cu fcc systems with random repeats (32-2048
atoms each, 402,324 atoms), positions jittered N(0, 0.05 A), Morse
D = 0.3429 eV, alpha = 1.3588 1/A, r0 = 2.866 A, cutoff 6 A, Langevin NVT
T = 1000 K, gamma = 10 / ps, dt = 2 fs.  FP32 pair arithmetic, FP64
accumulation / integration, synthetic seeded data.

One JSON line:
  value          the fused, graph-replayed engine over exactly --steps timed
                 steps, after max(--warmup, 200) untimed ones (the rebuild
                 rate settles at about 40 systems per step after ~150 steps
                 from the lattice start; the window before that is faster)
  steady_1000    1000 more timed steps of the same run, with rebuilds / step
  run_1000       a fresh engine timed over the workload's own 1000-step run
                 from the initial state (list build and capture included)
  plain_gpu      the reference pipeline in PyTorch ops on the GPU (the
                 denominator of the >= 5x target); per_step_hip beside it
  roofline       the fused-step kernel: the survey's algorithmic bytes
                 (184 + 4 k_list per atom-step) over its measured duration;
                 this design's own extra bytes reported apart
                 (overhead_bytes_per_atom) and the counter-measured traffic
                 with its source
  cpu_baseline   the FP64 CPU oracle on the host cores: a truncated run of
                 the full 512-system batch (value) and the full Ar-256 NVE run
  legs           the other configurations: 1M-atom LJ NVE, soft-sphere FIRE +
                 Frechet relaxed to fmax 1e-4 with in-place refill from a
                 queue, and the binary-LJ differentiable fit (grad-evals/s)
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import torch_sim_amd as tsa  # noqa: E402

HBM_PEAK_GBS = 8000.0
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)
WARMUP_FLOOR = 200

# Algorithmic bytes per atom-step of the fused step: the integrator reads x,
# p, F (3 x 24) and m (8) and writes x, p (2 x 24) and a packed FP32 copy of
# the position (16); the pair pass reads its own packed position (16), one
# 4-byte list word per stored neighbour and writes F (24).  The engine moves
# 84 more bytes per atom-step of its own design: the bin anchor (32), the
# displacement reference of the skin test (32), row order / system / row
# length (12) and the per-atom energy read back by the reduction (8).
ALGO_BYTES_BASE = 184.0
OVERHEAD_BYTES = 84.0

# HBM bytes per dispatch of the fused-step kernel from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE in KiB per dispatch, averaged over the dispatches),
# each scaled by what the same counter reads for a 256 MiB device copy
# (bench.py --calibrate-copy 256: FETCH_SIZE x 2.000, WRITE_SIZE x 1.000).
# Keyed by (workload, skin, tpa); "source" names the summary with the commands.
MEASURED_TRAFFIC = {
    ("cu_morse_nvt", 0.48, 2): {"fetch_kib": 102975.7, "fetch_factor": 2.000,
                                "write_kib": 42704.6, "write_factor": 1.000,
                                "source": "profiles/k_md_fused__cu_morse_nvt_512sys.md"},
    ("cu_morse_nvt", 0.6, 2): {"fetch_kib": 109373.6, "fetch_factor": 2.000,
                               "write_kib": 42558.4, "write_factor": 1.000,
                               "source": "profiles/k_md_fused__cu_morse_nvt_512sys.md"},
    ("lj1m_nve", 0.3, 4): {"fetch_kib": 319485.7, "fetch_factor": 2.000,
                           "write_kib": 126905.9, "write_factor": 1.000,
                           "source": "profiles/k_md_fused__lj_fluid_1m_nve.md"},
}


def measured_traffic(workload: str, skin: float, tpa: int):
    meas = MEASURED_TRAFFIC.get((workload, round(skin, 3), tpa))
    if meas is None:
        return None, None
    b = 1024.0 * (meas["fetch_kib"] * meas["fetch_factor"]
                  + meas["write_kib"] * meas["write_factor"])
    return b, meas["source"]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """CPU model, physical cores and the threads this process may use (the
    affinity mask, capped by OMP_NUM_THREADS when set)."""
    model, sockets, cores_per = None, 1, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Socket(s)":
                sockets = int(v)
            elif k.strip() == "Core(s) per socket":
                cores_per = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() else avail
    phys = sockets * cores_per if cores_per else None
    if phys:
        threads = min(threads, phys)
    return {"model": model, "physical_cores": phys, "cpus_available": avail,
            "threads": threads, "omp_num_threads": omp}


def make_engine(state, model, skin, tpa, bins=2, seed=1):
    kT = 1000.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    state.rng = seed
    md = tsa.nvt_langevin_init(state=state, model=model, kT=kT)
    eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, gamma=gamma, langevin=True, skin=skin, tpa=tpa,
                        seed=seed, chunk_steps=500, bins_per_rlist=bins)
    return md, eng


def timed_run(eng, steps: int):
    torch.cuda.synchronize()
    r0 = eng.n_rebuilds
    t0 = time.perf_counter()
    eng.run(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, eng.n_rebuilds - r0


def time_fused_kernel(eng, n_launch=100, n_warm=5):
    """Average duration of the fused-step kernel: HIP events on the replay
    stream around one replay of a graph of ``n_launch`` back-to-back fused
    launches (positions, momenta and the displacement checks advance as in
    the run; the lists are not rebuilt in between, so each launch does the
    work of the step it follows).  An event pair around a single eager launch
    adds about 10 us of marker and dispatch latency to a kernel of about
    80 us; inside one graph replay the launches follow each other as in the
    step's own graph, and the mean agrees with the profiler's kernel trace.
    The lists go stale on the way: this is the last thing done with ``eng``."""
    assert n_launch % 2 == 0  # the graph returns to the buffer it starts from
    if not eng._is_open:
        eng._open()
        eng._is_open = True
    for _ in range(n_warm):
        q = eng.cur
        eng._rebuild(q)
        eng._fused(q, end_kick=True, do_next=True)
        if not eng._merge_reduce:
            eng._reduce(q)
        eng.cur = 1 - q
    torch.cuda.synchronize()
    q0 = eng.cur
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for k in range(n_launch):
            eng._fused((q0 + k) % 2, end_kick=True, do_next=True)
    g.replay()
    st = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    g.replay()
    b.record(st)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / n_launch


def roofline(eng, n_atoms, workload, skin, tpa) -> dict:
    """The fused kernel against HBM peak with the survey's byte model."""
    k = eng.mean_neighbors
    ms = time_fused_kernel(eng)
    algo = (ALGO_BYTES_BASE + 4.0 * k) * n_atoms
    achieved = algo / (ms * 1e-3) / 1e9
    traffic, source = measured_traffic(workload, skin, tpa)
    log(f"fused kernel: {ms * 1e3:.1f} us, k_list {k:.1f}, {achieved:.0f} GB/s algorithmic "
        f"({achieved / HBM_PEAK_GBS:.3f} of peak)")
    return {
        "bound": "hbm", "kernel": "k_md_fused (tsx_md_fused_step)", "achieved": achieved,
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "algorithmic_bytes_per_atom": ALGO_BYTES_BASE + 4.0 * k,
        "bytes_per_launch": algo, "mean_list_neighbors": k, "kernel_ms": ms,
        "overhead_bytes_per_atom": OVERHEAD_BYTES,
        "with_overhead_frac": (algo + OVERHEAD_BYTES * n_atoms) / (ms * 1e-3) / 1e9
        / HBM_PEAK_GBS,
        "traffic_over_algorithmic": None if traffic is None else traffic / algo,
        "traffic_source": source,
    }


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



def _oracle_force_chunked(st, fn, cutoff, chunk=128):
    """Oracle forces / energies of a batch in groups of ``chunk`` systems
    (systems are independent; this bounds the oracle's candidate tensors)."""
    from oracle import models as om
    from oracle import nl as onl

    counts = torch.bincount(st.system_idx, minlength=st.n_systems).tolist()
    starts = [0]
    for c in counts:
        starts.append(starts[-1] + c)

    def force(x):
        es, fs = [], []
        for s0 in range(0, st.n_systems, chunk):
            s1 = min(st.n_systems, s0 + chunk)
            a, b = starts[s0], starts[s1]
            o = om.pair_forward(x[a:b], st.cell[s0:s1], st.pbc, st.atomic_numbers[a:b],
                                st.system_idx[a:b] - s0, fn, cutoff, compute_stress=False,
                                nl=onl.linked_cell_nl)
            es.append(o["energy"])
            fs.append(o["forces"])
        return torch.cat(es), torch.cat(fs)

    return force


def cpu_baseline(cfg2_steps: int = 2) -> dict:
    """The FP64 CPU oracle (the reference algorithm: linked-cell list rebuilt
    every step, autograd half-list forces, BAOAB / velocity Verlet) on the
    host cores: ``cfg2_steps`` BAOAB steps of the full 512-system Cu batch
    after one untimed force call (a truncated run of the headline workload:
    ``value``), and the full Ar-256 NVE run of 100 steps after one untimed
    step."""
    from oracle import md as omd
    from oracle import models as om

    cpu = host_cpu()
    torch.set_num_threads(cpu["threads"])
    st = tsa.workloads.cu_batch(512, device="cpu", seed=0)
    n = st.n_atoms
    force = _oracle_force_chunked(st, lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE), 6.0)
    kT = 1000.0 * omd.KB_EV
    dt = 0.002 * omd.TIME_UNIT
    gamma = 10.0 / omd.TIME_UNIT
    g = torch.Generator().manual_seed(0)
    p = omd.initialize_momenta(st.masses, st.system_idx, torch.tensor(kT, dtype=torch.float64),
                               torch.randn(n, 3, generator=g, dtype=torch.float64))
    t0 = time.perf_counter()
    e, f = force(st.positions)
    t_first = time.perf_counter() - t0
    x = st.positions
    t0 = time.perf_counter()
    for _ in range(cfg2_steps):
        noise = torch.randn(n, 3, generator=g, dtype=torch.float64)
        x, p, f, e = omd.langevin_step(x, p, f, st.masses, st.system_idx, force, dt, kT, gamma,
                                       noise)
    el = time.perf_counter() - t0
    value = n * cfg2_steps / el
    log(f"cpu baseline (512-system batch, {cfg2_steps} steps): {value:.4e} atom-steps/s")

    ar = tsa.workloads.ar_fcc(device="cpu")
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 3.405, 0.0104)  # noqa: E731
    ar_force = _oracle_force_chunked(ar, fn, 8.5125)
    g = torch.Generator().manual_seed(0)
    pa = omd.initialize_momenta(ar.masses, ar.system_idx, torch.tensor(50.0 * omd.KB_EV),
                                torch.randn(256, 3, generator=g, dtype=torch.float64))
    xa = ar.positions.clone()
    ea, fa = ar_force(xa)
    dta = 0.002 * omd.TIME_UNIT
    xa, pa, fa, ea = omd.nve_step(xa, pa, fa, ar.masses, ar_force, dta)  # untimed
    t0 = time.perf_counter()
    for _ in range(100):
        xa, pa, fa, ea = omd.nve_step(xa, pa, fa, ar.masses, ar_force, dta)
    ela = time.perf_counter() - t0
    log(f"cpu baseline (Ar-256, 100 NVE steps): {256 * 100 / ela:.4e} atom-steps/s")
    return {
        "value": value,
        "unit": "atom-steps/s",
        "cores": cpu["threads"],
        "kind": "port",
        "sample": f"truncated run of the headline workload: the full 512-system Cu batch "
                  f"({n} atoms), {cfg2_steps} BAOAB steps after one untimed force call "
                  f"({t_first:.1f} s), FP64 CPU oracle (linked-cell list every step, autograd "
                  f"half-list forces), {el:.1f} s on {cpu['threads']} threads",
        "cpu_model": cpu["model"],
        "physical_cores": cpu["physical_cores"],
        "cpus_available": cpu["cpus_available"],
        "ar256_nve_full": {"value": 256 * 100 / ela, "unit": "atom-steps/s", "steps": 100,
                           "seconds": ela,
                           "sample": "Ar fcc 4x4x4 (256 atoms), LJ rc 8.5125 A, 50 K, dt 2 fs, "
                                     "100 velocity-Verlet steps after one untimed step"},
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
        bins = args.bins if args.bins is not None else cfg.get("bins_per_rlist", 2)
    else:
        skin, tpa, bins = args.skin, args.tpa, args.bins or 2
    log(f"engine config: skin={skin} tpa={tpa} bins_per_rlist={bins}")

    md, eng = make_engine(state, model, skin, tpa, bins)
    warm = max(args.warmup, WARMUP_FLOOR)
    eng.run(warm)
    el, n_reb = timed_run(eng, args.steps)
    value = n * args.steps / el
    log(f"batched: {value:.4e} atom-steps/s ({1e3 * el / args.steps:.4f} ms/step), "
        f"rebuilds {n_reb} ({n_reb / args.steps:.1f} per step)")
    el_s, reb_s = timed_run(eng, 1000)
    steady = n * 1000 / el_s
    log(f"steady 1000 steps: {steady:.4e} atom-steps/s, {reb_s / 1000:.1f} rebuilds per step")
    roof = roofline(eng, n, "cu_morse_nvt", skin, tpa)
    traffic_ok = args.n_systems == 512 and bins == 2
    if not traffic_ok:
        roof["traffic"] = roof["traffic_over_algorithmic"] = roof["traffic_source"] = None
    _, eng2 = make_engine(state, model, skin, tpa, bins)
    el_r, reb_r = timed_run(eng2, 1000)
    run1000 = n * 1000 / el_r
    log(f"1000-step run from the initial state: {run1000:.4e} atom-steps/s")
    del eng2

    plain = plain_hip = None
    if not args.no_plain:
        plain = plain_torch(state, args.plain_steps, 1)
        log(f"plain per-step GPU path (PyTorch ops): {plain:.4e} atom-steps/s")
        plain_hip = plain_gpu(state, model, args.plain_steps, 1)
        log(f"per-step HIP kernels, no graph, list every step: {plain_hip:.4e} atom-steps/s")
    return {
        "metric": "MD atom-steps/s at 1 MI355X, batched graph-captured Morse NVT Langevin",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_effective": warm,
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
            "bins_per_rlist": bins,
            "dt_fs": 2.0,
            "T_K": 1000.0,
            "gamma_per_ps": 10.0,
            "list_rebuilds_in_timed_steps": n_reb,
            "rebuilds_per_step": n_reb / args.steps,
        },
        "steady_1000": {"value": steady, "unit": "atom-steps/s", "steps": 1000,
                        "ms_per_step": el_s, "rebuilds_per_step": reb_s / 1000},
        "run_1000": {"value": run1000, "unit": "atom-steps/s", "steps": 1000,
                     "ms_per_step": el_r, "rebuilds_per_step": reb_r / 1000,
                     "what": "a fresh engine over the workload's 1000-step run from the initial "
                             "state (first list build and graph capture included)"},
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
        "roofline": roof,
    }


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
    skin 0.3, dt 0.005 tau), 500 timed steps."""
    dev = torch.device("cuda:0")
    state = tsa.workloads.lj_fluid(64, device=dev, rho=0.8442, jitter=0.05, seed=0)
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev)
    state.rng = 0
    md = tsa.nve_init(state=state, model=model, kT=0.72)
    tpa, skin = args.tpa or 4, args.skin or 0.3
    eng = tsa.BatchedMD(model, md, dt=0.005, langevin=False, skin=skin, tpa=tpa, seed=1,
                        chunk_steps=250)
    n = state.n_atoms
    eng.run(args.warmup)
    el, n_reb = timed_run(eng, args.steps)
    value = n * args.steps / el
    log(f"lj1m: {value:.4e} atom-steps/s")
    return _line(
        "MD atom-steps/s at 1 MI355X, 1M-atom LJ NVE", value, "atom-steps/s", args.steps,
        args.warmup, 1e3 * el / args.steps,
        {"workload": "lj_fluid_1m_nve", "n_atoms": n, "rho": 0.8442, "T": 0.72, "cutoff": 2.5,
         "skin": skin, "tpa": tpa, "dt_tau": 0.005, "list_rebuilds_in_timed_steps": n_reb},
        roofline=roofline(eng, n, "lj1m_nve", skin, tpa),
    )


def bench_ar256(args) -> dict:
    """Ar fcc 4x4x4 (256 atoms) LJ NVE, 50 K, dt 2 fs, 100 steps."""
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
    el, _ = timed_run(eng, args.steps)
    value = 256 * args.steps / el
    log(f"ar256: {value:.4e} atom-steps/s")
    return _line("MD atom-steps/s at 1 MI355X, Ar-256 LJ NVE", value, "atom-steps/s", args.steps,
                 args.warmup, 1e3 * el / args.steps,
                 {"workload": "ar_fcc_256_nve", "n_atoms": 256, "T_K": 50.0, "dt_fs": 2.0,
                  "cutoff_A": 8.5125})


def bench_fire(args) -> dict:
    """Soft-sphere FIRE + Frechet relaxations to fmax 1e-4 (atoms and cell)
    with in-place refill: a queue of ``--fire-queue`` binary soft-sphere
    systems (64-1024 atoms, phi 0.64, triclinic strain <= 5 %) fed to
    ``optimize`` with an in-flight budget of the first 1024 systems' atoms,
    convergence checked every 5 steps; every system ends converged or at
    ``--fire-max-steps``."""
    dev = torch.device("cuda:0")
    n_queue = args.fire_queue
    queue = tsa.workloads.soft_sphere_batch(n_queue, n_min=64, n_max=1024, device=dev, seed=0)
    sig = torch.tensor([[1.0, 1.2], [1.2, 1.4]], dtype=torch.float64)
    skin = args.skin if args.skin is not None else 0.3
    model = tsa.SoftSphereMultiModel(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=sig,
                                     device=dev, compute_stress=True, skin=skin)
    counts = torch.bincount(queue.system_idx, minlength=n_queue)
    budget = int(counts[: min(1024, n_queue)].sum())
    work: dict = {}
    conv_fn = tsa.generate_force_convergence_fn(1e-4, include_cell_forces=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = tsa.optimize(queue, model, optimizer=tsa.Optimizer.fire, convergence_fn=conv_fn,
                       max_steps=args.fire_max_steps, steps_between_swaps=5,
                       autobatcher=tsa.InFlightAutoBatcher(max_atoms=budget),
                       init_kwargs={"cell_filter": "frechet"}, stats=work)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    conv = int(((out.fmax[:, 0] < 1e-4) & (out.fmax[:, 1] < 1e-4)).sum())
    value = work["atom_steps"] / el
    log(f"fire: {conv}/{n_queue} relaxed in {el:.2f} s ({n_queue / el:.1f} systems/s), "
        f"{work['batch_steps']} batch steps, {work['refills']} refills, {value:.4e} atom-steps/s")
    return _line("FIRE atom-steps/s at 1 MI355X, soft-sphere FIRE + Frechet to fmax 1e-4 with "
                 "in-place refill", value, "atom-steps/s", work["batch_steps"], 0,
                 1e3 * el / max(1, work["batch_steps"]),
                 {"workload": "soft_sphere_fire_frechet_refill", "queue_systems": n_queue,
                  "queue_atoms": queue.n_atoms, "in_flight_atom_budget": budget, "phi": 0.64,
                  "fmax": 1e-4, "steps_between_swaps": 5, "max_steps": args.fire_max_steps,
                  "skin": skin},
                 systems_relaxed=conv, systems_per_s=n_queue / el, seconds=el,
                 refills=work["refills"], swap_checks=work["swaps"],
                 full_list_builds=model._vl.full_builds if model._vl else None)


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
                  "nve_steps": n_md, "checkpoint_every": 20, "dt_tau": 0.002, "skin": skin},
                 force_matching_only_grad_evals_per_s=fm)


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


# workload -> (function, default steps, default warmup)
WORKLOADS = {
    "cu_morse_nvt": (bench_cu_morse, 1000, 50),
    "lj1m_nve": (bench_lj1m, 500, 20),
    "ar256_nve": (bench_ar256, 100, 10),
    "fire_ss_refill": (bench_fire, 0, 0),
    "lj_fit_256": (bench_fit, 3, 1),
}
# the legs the headline line carries: (workload, steps, warmup)
LEGS = (("lj1m_nve", 500, 20), ("fire_ss_refill", 0, 0), ("lj_fit_256", 3, 1))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cu_morse_nvt")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--n-systems", type=int, default=None)
    ap.add_argument("--skin", type=float, default=None)
    ap.add_argument("--tpa", type=int, default=None)
    ap.add_argument("--bins", type=int, default=None, help="cell-list bins per list radius")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="headline workload only (no LJ-1M / FIRE / fit legs)")
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="BAOAB steps of the full batch in the CPU baseline")
    ap.add_argument("--fit-md-steps", type=int, default=200)
    ap.add_argument("--fire-queue", type=int, default=4096)
    ap.add_argument("--fire-max-steps", type=int, default=10000)
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
    out = fn(args)
    if args.workload == "cu_morse_nvt":
        if not args.no_legs:
            legs = {}
            for name, steps, warmup in LEGS:
                sub = argparse.Namespace(**vars(args))
                sub.steps, sub.warmup, sub.skin, sub.tpa, sub.n_systems = steps, warmup, None, None, None
                log(f"leg {name}")
                legs[name] = WORKLOADS[name][0](sub)
                torch.cuda.empty_cache()
            out["legs"] = legs
        out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
