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
# WRITE_SIZE x 1.000).  Keyed by (n_systems, skin, tpa); the summary with its commands is
# profiles/k_md_fused__cu_morse_nvt_512sys.md.
MEASURED_TRAFFIC = {
    (512, 0.72, 2): {"fetch_kib": 107098.6, "fetch_factor": 2.000, "write_kib": 36122.8,
                     "write_factor": 1.000},
}


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


def time_fused_kernel(eng, n_launch=50):
    """Average duration of the fused-step kernel: hipEvents around each eager
    launch on the stream it is launched on."""
    st = torch.cuda.current_stream()
    durs = []
    for _ in range(n_launch):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        eng._fused(eng.cur, end_kick=True, do_next=True)
        b.record(st)
        eng.cur = 1 - eng.cur
        durs.append((a, b))
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in durs]
    ms.sort()
    return sum(ms[: len(ms) - 5]) / (len(ms) - 5)  # drop the 5 slowest (clock ramp)


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


def cpu_baseline(n_sys=12, steps=2, seed=0):
    """FP64 CPU oracle (reference algorithm: linked-cell list rebuilt every
    step, autograd half-list forces, BAOAB) on the first ``n_sys`` systems of
    the same workload."""
    from oracle import md as omd
    from oracle import models as om
    from oracle import nl as onl

    cores = os.cpu_count() or 1
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
    for _ in range(steps):
        noise = torch.randn(n_atoms, 3, generator=g, dtype=torch.float64)
        x, p, f, e = omd.langevin_step(x, p, f, m, idx, force, dt, kT, gamma, noise)
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


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n-systems", type=int, default=512)
    ap.add_argument("--skin", type=float, default=None)
    ap.add_argument("--tpa", type=int, default=None)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-systems", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; runs on more than one GPU are not part of this build "
              "(independent systems shard across GPUs without communication).", file=sys.stderr)
        sys.exit(2)
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

    traffic = None
    meas = MEASURED_TRAFFIC.get((args.n_systems, round(skin, 3), tpa))
    if meas is not None:
        traffic = 1024.0 * (meas["fetch_kib"] * meas["fetch_factor"] + meas["write_kib"] * meas["write_factor"])

    plain = None
    if not args.no_plain:
        plain = plain_gpu(state, model, args.plain_steps, 1)
        log(f"plain per-step GPU path: {plain:.4e} atom-steps/s")
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
        "plain_gpu": None if plain is None else {"value": plain, "unit": "atom-steps/s",
                                                  "steps": args.plain_steps,
                                                  "path": "nvt_langevin_step + model(state), list rebuilt every step"},
        "batched_vs_plain": None if plain is None else value / plain,
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
