"""GPU parity of the integrators and the fused graph engine.

Trajectories are compared with the Langevin noise supplied as a tensor, so
the GPU path and the FP64 oracle see the same random stream.

Tolerance (stated, derived from the FP32 pair arithmetic): the per-pair
force carries a relative error of ~2e-7 (FP32 exp / sqrt on FP64
displacements), i.e. up to ~5e-6 eV/A per atom summed over 78 neighbours.
Over 20 steps of dt = 0.196 (2 fs, mass 63.5 amu) that bounds the momentum
difference by ~2e-5 x 20 and the position difference by ~1e-6 A; the test
allows positions 2e-5 A, momenta 5e-4 amu A / t, forces 2e-5 of max |F| and
energies 1e-6 relative.
"""

import pytest
import torch

import torch_sim_amd as tsa
from oracle import md as omd, models as om, rng as orng

pytestmark = pytest.mark.gpu
F64 = torch.float64
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)
POS_TOL = 2e-5
MOM_TOL = 5e-4


def _morse_force_fn(st_cpu):
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, st_cpu.cell, st_cpu.pbc, st_cpu.atomic_numbers, st_cpu.system_idx, fn, 6.0,
                            compute_stress=False)
        return o["energy"], o["forces"]

    return force


def _setup(dev, n_sys=3, seed=21, kelvin=1000.0):
    st = tsa.workloads.cu_batch(n_sys, device=dev, seed=seed)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = seed
    kT = kelvin * tsa.MetalUnits.temperature
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    return st, model, md, kT


def _noise(n, steps, seed=0):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn(n, 3, generator=g, dtype=F64) for _ in range(steps)]


def _close(a, b, atol):
    d = (a.cpu() - b).abs().max().item()
    print(f"max deviation {d:.3e} (tolerance {atol:.1e})", flush=True)
    assert d <= atol, d


def test_per_step_langevin_matches_oracle(dev):
    st, model, md, kT = _setup(dev)
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    noise = _noise(st.n_atoms, 20)
    cpu = st.to(device="cpu")
    force = _morse_force_fn(cpu)
    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    m, idx = cpu.masses, cpu.system_idx
    for k in range(20):
        md = tsa.nvt_langevin_step(md, model, dt=dt, kT=kT, gamma=gamma, noise=noise[k].to(dev))
        x, p, f, e = omd.langevin_step(x, p, f, m, idx, force, dt, kT, gamma, noise[k])
    _close(md.positions, x, POS_TOL)
    _close(md.momenta, p, MOM_TOL)
    _close(md.forces, f, 2e-5 * f.abs().max().item())
    torch.testing.assert_close(md.energy.cpu(), e, rtol=1e-6, atol=0)


def test_per_step_nve_matches_oracle(dev):
    st, model, md, kT = _setup(dev, seed=22, kelvin=300.0)
    dt = 0.002 * tsa.MetalUnits.time
    cpu = st.to(device="cpu")
    force = _morse_force_fn(cpu)
    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    for _ in range(20):
        md = tsa.nve_step(md, model, dt=dt)
        x, p, f, e = omd.nve_step(x, p, f, cpu.masses, force, dt)
    _close(md.positions, x, POS_TOL)
    _close(md.momenta, p, MOM_TOL)


@pytest.mark.parametrize("langevin", [True, False])
def test_engine_matches_oracle_with_supplied_noise(dev, langevin):
    st, model, md, kT = _setup(dev, seed=23)
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    steps = 20
    noise = _noise(st.n_atoms, steps, seed=5)
    eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, gamma=gamma, langevin=langevin, skin=0.6, tpa=4)
    eng.run(steps, noise=[z.to(dev) for z in noise])
    cpu = st.to(device="cpu")
    force = _morse_force_fn(cpu)
    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    for k in range(steps):
        if langevin:
            x, p, f, e = omd.langevin_step(x, p, f, cpu.masses, cpu.system_idx, force, dt, kT, gamma, noise[k])
        else:
            x, p, f, e = omd.nve_step(x, p, f, cpu.masses, force, dt)
    out = eng.state()
    _close(out.positions, x, POS_TOL)
    _close(out.momenta, p, MOM_TOL)
    torch.testing.assert_close(out.energy.cpu(), e, rtol=1e-6, atol=0)


@pytest.mark.parametrize("tpa", [1, 2, 4])
def test_engine_internal_noise_is_the_philox_stream(dev, tpa):
    """The fused step's own Langevin noise (with two or more lanes per atom,
    lanes 0 and 1 each draw one Philox block) is the oracle's Philox4x32-10 +
    Box-Muller stream for (seed, step, atom): supplying the oracle's normals
    gives the same trajectory up to the last bits of log / sin / cos (a wrong
    normal would move a momentum by about 0.5)."""
    st, model, md, kT = _setup(dev, n_sys=2, seed=29)
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    seed, steps = 987654321, 3
    out = []
    for supplied in (False, True):
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, gamma=gamma, langevin=True, skin=0.6, tpa=tpa,
                            seed=seed)
        eng.initialize()
        ctr0 = int(eng.step_ctr.item())
        noise = None
        if supplied:
            noise = [torch.from_numpy(orng.langevin_normals(seed, ctr0 + k, st.n_atoms)).to(dev)
                     for k in range(steps)]
        print("tpa", tpa, "supplied noise", supplied, "first step counter", ctr0, flush=True)
        eng.run(steps, noise=noise)
        out.append(eng.state())
    _close(out[0].momenta, out[1].momenta.cpu(), 1e-9)
    _close(out[0].positions, out[1].positions.cpu(), 1e-9)
    assert (out[0].momenta - md.momenta).abs().max().item() > 0.01  # the noise did act


def test_rebuild_launch_reduction_equals_own_launch(dev):
    """With systems of at most 2048 atoms the per-system reduction of a step
    (energies, gate clearing, step counter) runs inside the next rebuild's
    launch: trajectories, energies and the step counter are bitwise those of
    the engine that launches the reduction on its own, with and without graph
    replay and across a chunk boundary."""
    st, model, md, kT = _setup(dev, n_sys=4, seed=33)
    dt = 0.002 * tsa.MetalUnits.time
    for use_graph in (True, False):
        runs = []
        for merge in (True, False):
            eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.3, seed=11,
                                use_graph=use_graph, chunk_steps=8)
            assert eng._merge_reduce
            eng._merge_reduce = merge
            print("graph", use_graph, "merged reduction", merge, flush=True)
            eng.run(21)
            eng.run(4)
            runs.append((eng.state(), int(eng.step_ctr.item()), int(eng.n_rebuilds)))
        (a, ca, ra), (b, cb, rb) = runs
        assert ra == rb and ra > 0
        assert ca == cb
        assert torch.equal(a.positions, b.positions)
        assert torch.equal(a.momenta, b.momenta)
        assert torch.equal(a.energy, b.energy)


def test_autotuner_picks_from_its_grid_and_caches(dev, tmp_path, monkeypatch):
    """tune_md times every (tpa, skin) pair of its grid, re-times the two
    fastest at every bin size and returns one of those; the choice is cached
    per workload shape, temperature, grid, library and GPU, in memory and in
    the JSON file, so the same call again runs no trial while another grid
    or temperature is another key."""
    import json
    from torch_sim_amd import autotune
    monkeypatch.setenv("TSX_AUTOTUNE_CACHE", str(tmp_path / "at.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    st, model, md, kT = _setup(dev, n_sys=4, seed=41)
    lines = []
    grid = dict(tpas=(2, 4), skins=(0.3, 0.6), bins=(1, 2))
    cfg = autotune.tune_md(st, model, warmup_steps=5, trial_steps=10, log=lines.append, **grid)
    print(lines, cfg, flush=True)
    assert (cfg["skin"], cfg["tpa"]) in {(0.3, 2), (0.3, 4), (0.6, 2), (0.6, 4)}
    assert cfg["bins_per_rlist"] in (1, 2)
    assert len(lines) == 8 and sum("re-timed" in ln for ln in lines) == 4
    assert cfg["rate"] > 0
    with open(tmp_path / "at.json") as fh:
        assert list(json.load(fh).values()) == [cfg]
    autotune._MEM.clear()
    again = []
    assert autotune.tune_md(st, model, log=again.append, **grid) == cfg and not again
    k = autotune._key(st, model, 1000.0, grid["tpas"], grid["skins"], grid["bins"])
    assert k != autotune._key(st, model, 1000.0, (2, 4, 8), grid["skins"], grid["bins"])
    assert k != autotune._key(st, model, 300.0, grid["tpas"], grid["skins"], grid["bins"])


def test_engine_graph_replay_equals_eager(dev):
    st, model, md, kT = _setup(dev, seed=24)
    dt = 0.002 * tsa.MetalUnits.time
    runs = []
    for use_graph in (True, False):
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.6, seed=7, use_graph=use_graph,
                            chunk_steps=16)
        eng.run(41)
        runs.append(eng.state())
    assert torch.equal(runs[0].positions, runs[1].positions)
    assert torch.equal(runs[0].momenta, runs[1].momenta)
    assert torch.equal(runs[0].energy, runs[1].energy)


def test_engine_rebuilds_track_per_step_path(dev):
    """Hot Cu (2500 K) over 150 steps with a small skin forces many rebuilds;
    the skin-list engine must follow the rebuild-every-step path."""
    st, model, md, _ = _setup(dev, n_sys=2, seed=25, kelvin=2500.0)
    kT = 2500.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    steps = 150
    noise = [z.to(dev) for z in _noise(st.n_atoms, steps, seed=9)]
    eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.2, tpa=2)
    eng.run(steps, noise=noise)
    assert eng.n_rebuilds >= 3
    ref = md
    ref_state = tsa.MDState(positions=md.positions.clone(), masses=md.masses, cell=md.cell, pbc=md.pbc,
                            atomic_numbers=md.atomic_numbers, system_idx=md.system_idx,
                            momenta=md.momenta.clone(), energy=md.energy.clone(), forces=md.forces.clone())
    for k in range(steps):
        ref_state = tsa.nvt_langevin_step(ref_state, model, dt=dt, kT=kT, noise=noise[k])
    out = eng.state()
    assert (out.positions - ref_state.positions).abs().max().item() < 1e-6
    del ref


def test_engine_batch_equals_single_bitwise_with_rebuilds(dev):
    """Per-system rebuild gates: each system's list is rebuilt on its own
    schedule, so a batched run is bit-identical to running each system alone,
    rebuilds included (reference: test_integrators.py:1104-1154)."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=26)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 3
    md = tsa.nve_init(state=st, model=model, kT=1500.0 * tsa.MetalUnits.temperature)
    dt = 0.002 * tsa.MetalUnits.time
    eng = tsa.BatchedMD(model, md, dt=dt, langevin=False, skin=0.2)
    eng.run(40)
    assert eng.n_rebuilds > 0
    full = eng.state()
    a = 0
    total = 0
    for k, single in enumerate(md.split()):
        e1 = tsa.BatchedMD(model, single, dt=dt, langevin=False, skin=0.2)
        e1.run(40)
        total += e1.n_rebuilds
        s1 = e1.state()
        n = single.n_atoms
        assert torch.equal(s1.positions, full.positions[a : a + n])
        assert torch.equal(s1.momenta, full.momenta[a : a + n])
        assert torch.equal(s1.energy[0], full.energy[k])
        a += n
    assert total == eng.n_rebuilds


def test_engine_fill_kernels_give_the_same_trajectory(dev):
    """Hot Cu (2000 K) with a small skin rebuilds many systems' lists; the
    per-bin, per-atom and one-wave-per-bin fills (4 or 8 centres per pass)
    write the same rows, so the trajectories and energies are bitwise the
    same, and so are consecutive runs against one run of their total length
    (the engine leaves the last step open)."""
    st = tsa.workloads.cu_batch(5, device=dev, seed=28)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 6
    kT = 2000.0 * tsa.MetalUnits.temperature
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    dt = 0.002 * tsa.MetalUnits.time
    runs = []
    for mode in (1, 2, 3, 4, 5):
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.2, seed=9)
        eng.nl.fill_mode = mode
        print("fill mode", mode, flush=True)
        eng.run(60)
        eng.run(1)
        eng.run(59)
        runs.append((eng.state(), eng.n_rebuilds))
    one = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.2, seed=9)
    one.run(120)
    runs.append((one.state(), one.n_rebuilds))
    assert runs[0][1] > 20
    for out, nreb in runs[1:]:
        assert nreb == runs[0][1]
        assert torch.equal(out.positions, runs[0][0].positions)
        assert torch.equal(out.momenta, runs[0][0].momenta)
        assert torch.equal(out.energy, runs[0][0].energy)


def test_engine_streamed_list_words_change_nothing(dev):
    """Non-temporal loads of the list words (the engine's choice for systems
    above 16,384 atoms) give bitwise the trajectory of plain loads."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=29)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 7
    kT = 1500.0 * tsa.MetalUnits.temperature
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    dt = 0.002 * tsa.MetalUnits.time
    runs = []
    for streamed in (False, True):
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.3, seed=4)
        assert not eng.stream_words  # small systems: plain loads by default
        eng.stream_words = streamed
        eng.run(50)
        runs.append(eng.state())
    assert torch.equal(runs[0].positions, runs[1].positions)
    assert torch.equal(runs[0].momenta, runs[1].momenta)
    assert torch.equal(runs[0].energy, runs[1].energy)


def test_engine_orthorhombic_path_equals_general_bitwise(dev):
    """Cubic cells take the fused kernel's orthorhombic stencil shift (one
    product per component); forcing the general three-term form must give the
    same bits, rebuilds included, graph replay on."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=27)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 5
    kT = 1500.0 * tsa.MetalUnits.temperature
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    dt = 0.002 * tsa.MetalUnits.time
    runs = []
    for ortho in (True, False):
        eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, langevin=True, skin=0.2, seed=7)
        assert eng.orthorhombic  # cubic cells
        eng.orthorhombic = ortho
        eng.run(60)
        assert eng.n_rebuilds > 0
        runs.append(eng.state())
    assert torch.equal(runs[0].positions, runs[1].positions)
    assert torch.equal(runs[0].momenta, runs[1].momenta)
    assert torch.equal(runs[0].energy, runs[1].energy)


def test_engine_nve_energy_conservation(dev):
    st = tsa.workloads.ar_fcc(dev, reps=(4, 4, 4))
    model = tsa.LennardJonesModel(sigma=3.405, epsilon=0.0104, cutoff=8.5125, device=dev)
    st.rng = 0
    md = tsa.nve_init(state=st, model=model, kT=50.0 * tsa.MetalUnits.temperature)
    e0 = (md.energy + tsa.calc_kinetic_energy(masses=md.masses, momenta=md.momenta,
                                              system_idx=md.system_idx)).item()
    eng = tsa.BatchedMD(model, md, dt=0.002 * tsa.MetalUnits.time, langevin=False, skin=0.5)
    eng.run(100)
    out = eng.state()
    e1 = (out.energy + tsa.calc_kinetic_energy(masses=out.masses, momenta=out.momenta,
                                               system_idx=out.system_idx)).item()
    # truncated (not shifted) cutoff: bounded jumps only (reference quirk)
    assert abs(e1 - e0) / st.n_atoms < 5e-4


def test_integrate_engine_and_plain_paths_run(dev):
    st = tsa.workloads.cu_batch(2, device=dev, seed=27)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    for engine in (True, False):
        out = tsa.integrate(st, model, integrator=tsa.Integrator.nvt_langevin, n_steps=12,
                            temperature=1000.0, timestep=0.002, engine=engine, seed=1)
        assert torch.isfinite(out.positions).all() and out.energy.shape == (2,)
        t = out.calc_temperature()
        assert (t > 100).all() and (t < 5000).all()


class _VirialMD(tsa.BatchedMD):
    """The engine with the fused step's per-atom virial output switched on
    (the C-ABI's w_atom; the engine itself passes NULL)."""

    def _fused(self, q, *, end_kick, do_next, noise=None):
        from torch_sim_amd._lib import call, ptr, stream

        if not hasattr(self, "w_atom"):
            self.w_atom = torch.zeros((self.n_atoms, 6), dtype=torch.float64,
                                      device=self.pos[0].device)
        nl, m = self.nl, self.model
        call(
            "tsx_md_fused_step",
            m.kind, ptr(m._params), m.n_species, ptr(self.pos[q]), ptr(self.pos[1 - q]),
            ptr(self.xref), ptr(self.cell), ptr(self.sysidx.sys_start), ptr(self.sysidx.atom_sys),
            ptr(nl.nbr), ptr(nl.count), ptr(nl.order), self.n_atoms, nl.jbits, nl.tpa, nl.ld,
            m.cutoff,
            ptr(self.p), ptr(self.m), ptr(self.f), ptr(self.e_atom), ptr(self.w_atom), self.dt,
            int(self.langevin), ptr(self.kT), ptr(self.gamma), ptr(noise), self.seed,
            ptr(self.step_ctr), int(end_kick), int(do_next), 0.5 * self.skin,
            ptr(self._gate(1 - q)), ptr(self.rel[q]), ptr(self.rel[1 - q]), ptr(self.anchor),
            ptr(self.ustep), int(self.orthorhombic), int(self.stream_words), stream(),
        )


def test_fused_virial_variant_keeps_the_trajectory(dev):
    """The fused step's virial variant (w_atom given) runs the same force
    arithmetic as the engine's variant: 40 Langevin steps with list rebuilds
    give bitwise the same positions, momenta and energies, and the per-atom
    virials come out finite and nonzero."""
    st, model, md, kT = _setup(dev, n_sys=3, seed=41)
    dt = 0.002 * tsa.MetalUnits.time
    runs = []
    for cls in (tsa.BatchedMD, _VirialMD):
        eng = cls(model, md, dt=dt, kT=kT, langevin=True, skin=0.3, seed=5, use_graph=False)
        print("engine", cls.__name__, flush=True)
        eng.run(40)
        runs.append((eng.state(), eng.n_rebuilds, getattr(eng, "w_atom", None)))
    (a, na, _), (b, nb, w) = runs
    assert na == nb and na > 0
    assert torch.equal(a.positions, b.positions)
    assert torch.equal(a.momenta, b.momenta)
    assert torch.equal(a.energy, b.energy)
    assert torch.isfinite(w).all() and w.abs().max().item() > 0
