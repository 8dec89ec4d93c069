"""GPU parity at full and above-2048-atom sizes.

Systems above 2048 atoms take other code than the batched Cu tests: the
grid-wide bin / scan / scatter / sort kernels, the chunked per-system
reductions (``k_seg_finish``), the per-atom fill on one large box, and the
engine's reduction launch of its own.  These tests run them against the FP64
oracle:

* a 16,384-atom LJ fluid: pair set bit-exact, forces / energy, 10 NVE steps;
* a mixed batch of two systems above 2048 atoms and two small ones, with
  both fills: pair set, forces / energy / stress, 10 supplied-noise
  Langevin steps;
* the full 512-system Cu Morse batch (402,324 atoms) of the headline
  workload: one model call against the oracle;
* the 1,048,576-atom LJ fluid: the pair rows of 256 sampled atoms against a
  brute-force minimum-image search, their forces against the oracle's pair
  function with autograd, sum F = 0, graph == eager bitwise, and bounded NVE
  energy drift;
* ``integrate`` against the oracle's BAOAB fed the Philox normals, which
  pins the K / ps / (1/ps) conversions and the engine hand-off;
* ``initialize_momenta`` against the oracle fed the generator's normals;
* a freed ``system_idx`` whose memory is handed to a new layout.

Tolerances are those of the other GPU tests: forces 1e-5 of max |F|,
energies 1e-6 relative, positions 2e-5 A and momenta 5e-4 after the stated
steps (FP32 pair arithmetic).
"""

import gc
import math

import pytest
import torch

import torch_sim_amd as tsa
from oracle import md as omd, models as om, nl as onl, rng as orng

pytestmark = pytest.mark.gpu
F64 = torch.float64
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)


def _say(*a):
    print(*a, flush=True)


def _oracle_forward(st_cpu, fn, cutoff, chunk=64, **kw):
    """Oracle forward in groups of ``chunk`` systems (systems are
    independent; this bounds the oracle's candidate tensors)."""
    counts = torch.bincount(st_cpu.system_idx, minlength=st_cpu.n_systems).tolist()
    outs, a = [], 0
    for s0 in range(0, st_cpu.n_systems, chunk):
        s1 = min(st_cpu.n_systems, s0 + chunk)
        b = a + sum(counts[s0:s1])
        o = om.pair_forward(st_cpu.positions[a:b], st_cpu.cell[s0:s1], st_cpu.pbc,
                            st_cpu.atomic_numbers[a:b], st_cpu.system_idx[a:b] - s0, fn, cutoff,
                            nl=onl.linked_cell_nl, **kw)
        outs.append(o)
        a = b
    return {k: torch.cat([o[k] for o in outs]) for k in outs[0]}


def _check_forward(out, ref, stress=False):
    fr = ref["forces"]
    df = (out["forces"].cpu() - fr).abs().max().item()
    _say(f"force deviation {df:.3e} of max |F| {fr.abs().max().item():.3e}")
    assert df <= 1e-5 * fr.abs().max().item() + 1e-9
    de = ((out["energy"].cpu() - ref["energy"]).abs() / ref["energy"].abs()).max().item()
    _say(f"energy relative deviation {de:.3e}")
    assert de <= 1e-6
    if stress:
        sr = ref["stress"]
        ds = (out["stress"].cpu() - sr).abs().max().item()
        assert ds <= 1e-5 * sr.abs().max().item() + 1e-9, ds


def _sorted_set(mapping, sysmap, shifts):
    return onl.sort_pairs(mapping.cpu(), sysmap.cpu(), shifts.cpu())


def _assert_same_set(a, b):
    assert a[0].shape == b[0].shape, (a[0].shape, b[0].shape)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert torch.equal(a[2].to(F64), b[2].to(F64))


# -------------------------------------------------------- 16k-atom LJ fluid
@pytest.fixture(scope="module")
def lj16k(dev):
    st = tsa.workloads.lj_fluid(16, device=dev, seed=3)
    assert st.n_atoms == 16384
    return st


def test_lj16k_pair_set_exact(dev, lj16k):
    st = lj16k
    got = _sorted_set(*tsa.torchsim_nl(st.positions, st.cell.mT.contiguous(), st.pbc, 2.5,
                                       st.system_idx))
    cpu = st.to(device="cpu")
    ref = onl.sort_pairs(*onl.linked_cell_nl(cpu.positions, cpu.cell.mT, cpu.pbc, 2.5,
                                             cpu.system_idx))
    _say("pairs", got[0].shape[1])
    _assert_same_set(got, ref)


@pytest.mark.parametrize("tpa", [1, 4])
def test_lj16k_model_matches_oracle(dev, lj16k, tpa):
    st = lj16k
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev,
                                  compute_stress=True, tpa=tpa)
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 1.0, 1.0)  # noqa: E731
    _check_forward(model(st), _oracle_forward(st.to(device="cpu"), fn, 2.5), stress=True)


def test_lj16k_engine_nve_matches_oracle(dev, lj16k):
    st = lj16k.clone()
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev)
    st.rng = 4
    md = tsa.nve_init(state=st, model=model, kT=0.72)
    steps, dt = 10, 0.005
    eng = tsa.BatchedMD(model, md, dt=dt, langevin=False, skin=0.3, tpa=4)
    assert not eng._merge_reduce  # one system above 2048 atoms: reduction launch of its own
    eng.run(steps)
    out = eng.state()
    cpu = st.to(device="cpu")
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 1.0, 1.0)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cpu.cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 2.5,
                            compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    for _ in range(steps):
        x, p, f, e = omd.nve_step(x, p, f, cpu.masses, force, dt)
    dx = (out.positions.cpu() - x).abs().max().item()
    dp = (out.momenta.cpu() - p).abs().max().item()
    _say(f"positions {dx:.3e}, momenta {dp:.3e}")
    assert dx <= 2e-5 and dp <= 5e-4
    torch.testing.assert_close(out.energy.cpu(), e, rtol=1e-6, atol=0)


def test_binary_lj_engine_nve_matches_oracle(dev):
    """The fused step's several-species variant (per-pair constants in an LDS
    table, each atom's species from the bits of its frame record) on a batch
    of binary (Kob-Andersen) LJ liquids: 10 NVE steps against the oracle's
    velocity Verlet with its species-lookup pair function."""
    st = tsa.workloads.binary_lj_batch(6, n_min=128, n_max=512, rho=0.8, device=dev, seed=5)
    sig = torch.tensor([[1.0, 0.8], [0.8, 0.88]], dtype=F64)
    eps = torch.tensor([[1.0, 1.5], [1.5, 0.5]], dtype=F64)
    model = tsa.MultiLennardJonesModel(torch.tensor([1, 2]), sig, eps, device=dev, cutoff=2.5)
    st.rng = 6
    md = tsa.nve_init(state=st, model=model, kT=0.5)
    steps, dt = 10, 0.002
    eng = tsa.BatchedMD(model, md, dt=dt, langevin=False, skin=0.3, tpa=2)
    eng.run(steps)
    out = eng.state()
    cpu = st.to(device="cpu")
    fn = om.MultiPairFn([1, 2], sig, eps, kind="lj")

    def force(x):
        o = om.pair_forward(x, cpu.cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 2.5,
                            compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    for _ in range(steps):
        x, p, f, e = omd.nve_step(x, p, f, cpu.masses, force, dt)
    dx = (out.positions.cpu() - x).abs().max().item()
    dp = (out.momenta.cpu() - p).abs().max().item()
    _say(f"binary LJ: positions {dx:.3e}, momenta {dp:.3e}")
    assert dx <= 2e-5 and dp <= 5e-4
    torch.testing.assert_close(out.energy.cpu(), e, rtol=1e-6, atol=0)


# ------------------------------------------- mixed batch above 2048 atoms
@pytest.fixture(scope="module")
def mixed(dev):
    st = tsa.workloads.cu_batch(4, device=dev, seed=8,
                                reps=[[9, 9, 9], [2, 2, 2], [10, 9, 8], [3, 4, 5]])
    assert torch.bincount(st.system_idx).tolist() == [2916, 32, 2880, 240]
    return st


@pytest.mark.parametrize("fill_mode", [1, 2])
def test_mixed_batch_pair_set_exact_both_fills(dev, mixed, fill_mode):
    from torch_sim_amd.neighbors import NeighborList, SystemIndex

    st = mixed
    si = SystemIndex(st.system_idx, st.n_systems)
    nl = NeighborList(6.0, tpa=2, fill_mode=fill_mode, bins_per_rlist=2)
    nl.build(st.positions.contiguous(), st.cell.contiguous(), st.pbc, si)
    got = _sorted_set(*nl.pairs(si, F64))
    cpu = st.to(device="cpu")
    ref = onl.sort_pairs(*onl.linked_cell_nl(cpu.positions, cpu.cell.mT, cpu.pbc, 6.0,
                                             cpu.system_idx))
    _say("fill mode", fill_mode, "pairs", got[0].shape[1])
    _assert_same_set(got, ref)


def test_mixed_batch_model_matches_oracle(dev, mixed):
    st = mixed
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev, compute_stress=True,
                           per_atom_energies=True)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    out = model(st)
    ref = _oracle_forward(st.to(device="cpu"), fn, 6.0, per_atom=True)
    _check_forward(out, ref, stress=True)
    torch.testing.assert_close(out["energies"].cpu(), ref["energies"], rtol=1e-5, atol=1e-9)


def test_mixed_batch_engine_langevin_matches_oracle(dev, mixed):
    st = mixed.clone()
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 9
    kT = 1000.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    gamma = 10.0 / tsa.MetalUnits.time
    md = tsa.nvt_langevin_init(state=st, model=model, kT=kT)
    steps = 10
    g = torch.Generator().manual_seed(10)
    noise = [torch.randn(st.n_atoms, 3, generator=g, dtype=F64) for _ in range(steps)]
    eng = tsa.BatchedMD(model, md, dt=dt, kT=kT, gamma=gamma, langevin=True, skin=0.6, tpa=2)
    assert not eng._merge_reduce
    eng.run(steps, noise=[z.to(dev) for z in noise])
    out = eng.state()
    cpu = st.to(device="cpu")
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cpu.cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 6.0,
                            compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    x, p = md.positions.cpu().clone(), md.momenta.cpu().clone()
    e, f = force(x)
    for k in range(steps):
        x, p, f, e = omd.langevin_step(x, p, f, cpu.masses, cpu.system_idx, force, dt, kT, gamma,
                                       noise[k])
    dx = (out.positions.cpu() - x).abs().max().item()
    dp = (out.momenta.cpu() - p).abs().max().item()
    _say(f"positions {dx:.3e}, momenta {dp:.3e}")
    assert dx <= 2e-5 and dp <= 5e-4
    torch.testing.assert_close(out.energy.cpu(), e, rtol=1e-6, atol=0)


# -------------------------------------- the full 512-system Cu Morse batch
def test_full_cu_512_batch_model_matches_oracle(dev):
    """The headline workload's batch (512 systems, 402,324 atoms, jbits 11):
    one model call, forces, energies and stress against the FP64 oracle."""
    st = tsa.workloads.cu_batch(512, device=dev, seed=0)
    assert st.n_atoms == 402324
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev, compute_stress=True, tpa=2)
    out = model(st)
    torch.cuda.synchronize()
    _say("gpu forward done; oracle on", st.n_atoms, "atoms")
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    ref = _oracle_forward(st.to(device="cpu"), fn, 6.0, chunk=64)
    _check_forward(out, ref, stress=True)


# ------------------------------------------------ the 1M-atom LJ fluid
@pytest.fixture(scope="module")
def lj1m(dev):
    st = tsa.workloads.lj_fluid(64, device=dev, rho=0.8442, jitter=0.05, seed=0)
    assert st.n_atoms == 1048576
    return st


def _min_image_rows(pos_cpu, box, rows, rc):
    """Brute force over all atoms, minimum image (the box is far wider than
    2 rc, so each neighbour has one image): sorted neighbour indices and the
    displacement r_j - r_i of each sampled atom."""
    out = []
    for i in rows.tolist():
        d = pos_cpu - pos_cpu[i]
        d = d - box * torch.round(d / box)
        r2 = (d * d).sum(1)
        r2[i] = float("inf")
        j = torch.nonzero(r2 < rc * rc).flatten()
        out.append((j, d[j]))
    return out


def test_lj1m_sampled_rows_and_forces_match_oracle(dev, lj1m):
    st = lj1m
    n = st.n_atoms
    box = float(st.cell[0, 0, 0])
    g = torch.Generator().manual_seed(1)
    rows = torch.randperm(n, generator=g)[:256].sort().values
    mapping, _, shifts = tsa.torchsim_nl(st.positions, st.cell.mT.contiguous(), st.pbc, 2.5,
                                         st.system_idx)
    assert mapping.shape[1] % 2 == 0
    sel = torch.isin(mapping[0], rows.to(dev))
    got_i, got_j = mapping[0, sel].cpu(), mapping[1, sel].cpu()
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev, tpa=4,
                                  per_atom_energies=True)
    out = model(st)
    f_all = out["forces"]
    fsum = f_all.sum(0).abs().max().item()
    fmax = f_all.abs().max().item()
    _say(f"pairs {mapping.shape[1]}, |sum F| {fsum:.3e}, max |F| {fmax:.3e}")
    assert fsum <= 1e-6 * fmax
    f_rows = f_all[rows.to(dev)].cpu()
    e_rows = out["energies"][rows.to(dev)].cpu()
    del mapping, shifts, sel, out, f_all
    pos = st.positions.cpu()
    ref = _min_image_rows(pos, box, rows, 2.5)
    for k, i in enumerate(rows.tolist()):
        j_ref, d_ref = ref[k]
        mine = got_j[got_i == i].sort().values
        assert torch.equal(mine, j_ref), i
        # force on i from the oracle's pair function by autograd:
        # F_i = -dE_i/dx_i with E_i = sum_j V(|x_j - x_i + image|)
        xi = pos[i].clone().requires_grad_(True)
        dr = (pos[j_ref] + (d_ref - (pos[j_ref] - pos[i]))) - xi
        e = om.lennard_jones_pair(dr.norm(dim=1), None, None, 1.0, 1.0).sum()
        (gi,) = torch.autograd.grad(e, xi)
        assert (f_rows[k] + gi).abs().max().item() <= 1e-5 * fmax + 1e-9
        # per-atom energy carries the full energy of each of the atom's pairs
        assert abs(e_rows[k].item() - e.item()) <= 1e-6 * abs(e.item()) + 1e-9


def test_lj1m_engine_graph_equals_eager_and_energy_bounded(dev, lj1m):
    st = lj1m.clone()
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev)
    st.rng = 0
    md = tsa.nve_init(state=st, model=model, kT=0.72)
    runs = []
    for use_graph in (True, False):
        eng = tsa.BatchedMD(model, md, dt=0.005, langevin=False, skin=0.3, tpa=4, seed=1,
                            use_graph=use_graph, chunk_steps=50)
        eng.run(30)
        runs.append(eng.state())
        _say("graph", use_graph, "rebuilds", eng.n_rebuilds)
    assert torch.equal(runs[0].positions, runs[1].positions)
    assert torch.equal(runs[0].momenta, runs[1].momenta)
    assert torch.equal(runs[0].energy, runs[1].energy)
    del runs
    gc.collect()

    def pairs_within_rc(x):
        mapping, _, _ = tsa.torchsim_nl(x, st.cell.mT.contiguous(), st.pbc, 2.5, st.system_idx)
        return mapping.shape[1] // 2

    # The potential is truncated, not shifted (pair_potential.py:398-400), so
    # E + KE jumps by -V(rc) = 0.0163 whenever a pair crosses rc; the jittered
    # lattice (54 pairs per atom inside rc) disorders towards the liquid's 55.3
    # and E + KE falls by about 0.01 per atom in 100 steps.  The same forces
    # come from the shifted potential V - V(rc), whose energy
    # E - P V(rc) + KE (P = pairs inside rc) is conserved up to the integrator
    # error: that is what is bounded here.
    v_rc = 4.0 * (2.5 ** -12 - 2.5 ** -6)
    ke0 = tsa.calc_kinetic_energy(masses=md.masses, momenta=md.momenta, system_idx=md.system_idx)
    e0 = (md.energy + ke0).item()
    h0 = e0 - pairs_within_rc(md.positions) * v_rc
    eng = tsa.BatchedMD(model, md, dt=0.005, langevin=False, skin=0.3, tpa=4, seed=1)
    eng.run(100)
    out = eng.state()
    ke1 = tsa.calc_kinetic_energy(masses=out.masses, momenta=out.momenta,
                                  system_idx=out.system_idx)
    e1 = (out.energy + ke1).item()
    h1 = e1 - pairs_within_rc(out.positions) * v_rc
    n = st.n_atoms
    drift = abs(h1 - h0) / n
    _say(f"E + KE per atom {e0 / n:.6f} -> {e1 / n:.6f}; shifted-potential energy per atom "
         f"{h0 / n:.6f} -> {h1 / n:.6f}, drift {drift:.2e}")
    assert drift < 2e-4
    assert out.forces.sum(0).abs().max().item() <= 1e-6 * out.forces.abs().max().item()


# ---------------------------------------------------------------- integrate
def test_integrate_langevin_matches_oracle_with_philox_normals(dev):
    """integrate(temperature=K, timestep=ps, gamma=1/ps, seed=s) through the
    engine equals the oracle's BAOAB with the oracle's own unit constants,
    fed the Philox normals of (s, step, atom): the K -> eV, ps -> internal
    and 1/ps -> internal conversions and the engine hand-off are pinned."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=31)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    seed, steps, T, dt_ps, gamma_ps = 123456789, 12, 1000.0, 0.002, 10.0
    st.rng = 17
    out = tsa.integrate(st, model, integrator=tsa.Integrator.nvt_langevin, n_steps=steps,
                        temperature=T, timestep=dt_ps, gamma=gamma_ps, seed=seed)
    # the same initial momenta: the state's generator seeded the same way
    st0 = tsa.workloads.cu_batch(3, device=dev, seed=31)
    st0.rng = 17
    kT = T * omd.KB_EV
    md0 = tsa.nvt_langevin_init(state=st0, model=model, kT=kT)
    cpu = st0.to(device="cpu")
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cpu.cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 6.0,
                            compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    dt = dt_ps * omd.TIME_UNIT
    gamma = gamma_ps / omd.TIME_UNIT
    x, p = md0.positions.cpu().clone(), md0.momenta.cpu().clone()
    e, f = force(x)
    for k in range(steps):
        z = torch.from_numpy(orng.langevin_normals(seed, 1 + k, st0.n_atoms))
        x, p, f, e = omd.langevin_step(x, p, f, cpu.masses, cpu.system_idx, force, dt, kT, gamma, z)
    dx = (out.positions.cpu() - x).abs().max().item()
    dp = (out.momenta.cpu() - p).abs().max().item()
    _say(f"positions {dx:.3e}, momenta {dp:.3e}")
    assert dx <= 2e-5 and dp <= 5e-4
    assert (out.momenta.cpu() - md0.momenta.cpu()).abs().max().item() > 0.05  # the noise acted
    torch.testing.assert_close(out.energy.cpu(), e, rtol=1e-6, atol=0)


def test_integrate_nve_matches_oracle(dev):
    st = tsa.workloads.cu_batch(2, device=dev, seed=32)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 3
    out = tsa.integrate(st, model, integrator="nve", n_steps=15, temperature=600.0,
                        timestep=0.001)
    st0 = tsa.workloads.cu_batch(2, device=dev, seed=32)
    st0.rng = 3
    md0 = tsa.nve_init(state=st0, model=model, kT=600.0 * omd.KB_EV)
    cpu = st0.to(device="cpu")
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cpu.cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 6.0,
                            compute_stress=False, nl=onl.linked_cell_nl)
        return o["energy"], o["forces"]

    x, p = md0.positions.cpu().clone(), md0.momenta.cpu().clone()
    e, f = force(x)
    for _ in range(15):
        x, p, f, e = omd.nve_step(x, p, f, cpu.masses, force, 0.001 * omd.TIME_UNIT)
    assert (out.positions.cpu() - x).abs().max().item() <= 2e-5
    assert (out.momenta.cpu() - p).abs().max().item() <= 5e-4


# ------------------------------------------ quantities against the oracle
def test_calc_kT_and_kinetic_energy_match_oracle(dev):
    st = tsa.workloads.cu_batch(5, device=dev, seed=34)
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    st.rng = 6
    md = tsa.nvt_langevin_init(state=st, model=model, kT=800.0 * tsa.MetalUnits.temperature)
    p, m, idx = md.momenta, md.masses, md.system_idx
    kt = tsa.calc_kT(masses=m, momenta=p, system_idx=idx)
    ke = tsa.calc_kinetic_energy(masses=m, momenta=p, system_idx=idx)
    torch.testing.assert_close(kt.cpu(), omd.calc_kT(p.cpu(), m.cpu(), idx.cpu(), 5), rtol=1e-12,
                               atol=0)
    torch.testing.assert_close(ke.cpu(), omd.kinetic_energy(p.cpu(), m.cpu(), idx.cpu(), 5),
                               rtol=1e-12, atol=0)
    v = p / m.unsqueeze(-1)
    torch.testing.assert_close(tsa.calc_kT(masses=m, velocities=v, system_idx=idx), kt,
                               rtol=1e-12, atol=0)
    whole = tsa.calc_kT(masses=m, momenta=p)
    assert abs(whole.item() - (ke.sum() * 2 / (3 * st.n_atoms)).item()) < 1e-12 * whole.item()
    with pytest.raises(ValueError):
        tsa.calc_kT(masses=m, momenta=p, velocities=v)


def test_initialize_momenta_matches_oracle(dev):
    """Maxwell-Boltzmann momenta on the GPU against the oracle fed the same
    normals (the generator's stream redrawn): per-system kT, the per-system
    mean removed where a system has more than one atom (a one-atom system
    keeps its draw), a zero total momentum per system."""
    reps = [[2, 2, 2], [3, 2, 2], [2, 3, 3]]
    st = tsa.workloads.cu_batch(3, device=dev, reps=reps, seed=35)
    lone = tsa.SimState(positions=torch.zeros(1, 3, dtype=F64, device=dev),
                        masses=torch.full((1,), 63.546, dtype=F64, device=dev),
                        cell=st.cell[:1].clone(), pbc=st.pbc,
                        atomic_numbers=torch.full((1,), 29, dtype=torch.long, device=dev),
                        system_idx=torch.zeros(1, dtype=torch.long, device=dev))
    st = tsa.concatenate_states([st, lone])
    kT = torch.tensor([300.0, 800.0, 1500.0, 1000.0], dtype=F64, device=dev) \
        * tsa.MetalUnits.temperature
    g = torch.Generator(device=dev).manual_seed(1234)
    p = tsa.initialize_momenta(st.positions, st.masses, st.system_idx, kT, generator=g)
    g.manual_seed(1234)
    xi = torch.randn(st.positions.shape, device=dev, dtype=F64, generator=g)
    ref = omd.initialize_momenta(st.masses.cpu(), st.system_idx.cpu(), kT.cpu(), xi.cpu())
    torch.testing.assert_close(p.cpu(), ref, rtol=1e-13, atol=1e-15)
    tot = torch.zeros(4, 3, dtype=F64).index_add(0, st.system_idx.cpu(), p.cpu())
    assert tot[:3].abs().max().item() < 1e-10
    lone_p = (xi[-1] * torch.sqrt(st.masses[-1] * kT[3])).cpu()
    torch.testing.assert_close(p[-1].cpu(), lone_p, rtol=1e-15, atol=0)


# ------------------------------------------- caches against reused memory
def test_freed_system_idx_memory_reused_by_a_new_layout(dev):
    """A model keeps its system layout between calls.  When the state's
    ``system_idx`` is freed and the caching allocator hands the same block to
    a new ``system_idx`` with another per-system layout (same atom and
    system counts), the model must not reuse the old layout."""
    model = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    a_cpu = tsa.workloads.cu_batch(2, device="cpu", seed=35, reps=[[2, 2, 3], [3, 3, 2]])
    b_cpu = tsa.workloads.cu_batch(2, device="cpu", seed=36, reps=[[3, 3, 2], [2, 2, 3]])
    counts_a = torch.bincount(a_cpu.system_idx).tolist()
    assert torch.bincount(b_cpu.system_idx).tolist() == counts_a[::-1] != counts_a
    n = a_cpu.n_atoms
    # system_idx lives at the head of a buffer of a size no other tensor here
    # has, so that the allocator hands the freed block to the next buffer of
    # that size; both heads get the same in-place history (zero, copy)
    size = 400_000 + n

    def on_device(st_cpu):
        buf = torch.zeros(size, dtype=torch.int64, device=dev)
        buf[:n].copy_(st_cpu.system_idx)
        return tsa.SimState(positions=st_cpu.positions.to(dev), masses=st_cpu.masses.to(dev),
                            cell=st_cpu.cell.to(dev), pbc=True,
                            atomic_numbers=st_cpu.atomic_numbers.to(dev), system_idx=buf[:n])

    a = on_device(a_cpu)
    model(a)
    ptr_a, ver_a = a.system_idx.data_ptr(), a.system_idx._version
    del a
    gc.collect()
    b = on_device(b_cpu)
    if b.system_idx.data_ptr() != ptr_a:
        pytest.skip("the allocator did not hand the freed block back; nothing to test")
    assert b.system_idx._version == ver_a
    out = model(b)
    ref = om.pair_forward(b_cpu.positions, b_cpu.cell, b_cpu.pbc, b_cpu.atomic_numbers,
                          b_cpu.system_idx, fn, 6.0, nl=onl.linked_cell_nl)
    _check_forward(out, ref)
    assert math.isfinite(out["energy"].sum().item())
