"""Force-defined, generic and composite models on the GPU against the FP64
oracle: PairForcesModel / ParticleLifeModel (pair_potential.py:467-612,
particle_life.py), PairPotentialModel with a user pair function (the
Born-Mayer-Huggins-Tosi-Fumi example of pair_potential.py:35-61), SumModel
and SerialSumModel (interface.py:204-351).

Tolerances: the HIP particle-life kernel has FP32 pair arithmetic: forces
1e-5 of max |F|, stresses 1e-5 of their largest entry.  The generic path is
FP64 torch arithmetic over the HIP list's pairs (a different summation
order from the oracle's): 1e-10 relative.
"""

import functools

import pytest
import torch

import torch_sim_amd as tsa
from oracle import models as om
from oracle import nl as onl

pytestmark = pytest.mark.gpu
F64 = torch.float64
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)
BMHTF = dict(A=20.3548, B=3.1546, C=674.4793, D=837.0770, sigma=2.755)


def bmhtf_pair(dr, zi, zj, A, B, C, D, sigma):  # noqa: ARG001, N803
    """A exp(B (sigma - r)) - C / r^6 - D / r^8, zero at r <= 0: a user pair
    function that no HIP kernel knows."""
    e = A * torch.exp(B * (sigma - dr)) - C / dr.pow(6) - D / dr.pow(8)
    return torch.where(dr > 0, e, torch.zeros_like(e))


def _close(a, b, tag, rel):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    scale = max(b.abs().max().item(), 1e-300)
    err = (a - b).abs().max().item() / scale
    print(f"{tag}: rel err {err:.2e} (scale {scale:.3e})", flush=True)
    assert err <= rel, (tag, err)


def _oracle_forces(st, fn, cutoff, half=False):
    c = st.to(device="cpu")
    return om.pair_forces_forward(c.positions, c.cell, c.pbc, c.atomic_numbers, c.system_idx,
                                  fn, cutoff, half=half, per_atom=True, nl=onl.linked_cell_nl)


def _oracle_pairs(st, fn, cutoff, half=True):
    c = st.to(device="cpu")
    return om.pair_forward(c.positions, c.cell, c.pbc, c.atomic_numbers, c.system_idx, fn,
                           cutoff, half=half, per_atom=True, nl=onl.linked_cell_nl)


def _nacl(dev, reps=2, a=5.64, jitter=0.05, seed=3):
    """Rock-salt NaCl, reps^3 conventional cells, jittered."""
    base = torch.tensor([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=F64)
    na = base
    cl = base + torch.tensor([.5, 0, 0], dtype=F64)
    cells = torch.stack(torch.meshgrid(*[torch.arange(reps)] * 3, indexing="ij"), -1)
    cells = cells.reshape(-1, 1, 3).to(F64)
    frac = torch.cat([(na + cells).reshape(-1, 3), (cl + cells).reshape(-1, 3)]) / reps
    g = torch.Generator().manual_seed(seed)
    pos = frac * a * reps + jitter * torch.randn(frac.shape, generator=g, dtype=F64)
    n = pos.shape[0]
    z = torch.cat([torch.full((n // 2,), 11), torch.full((n // 2,), 17)])
    m = torch.cat([torch.full((n // 2,), 22.99), torch.full((n // 2,), 35.45)]).to(F64)
    return tsa.SimState(positions=pos.to(dev), masses=m.to(dev),
                        cell=(a * reps * torch.eye(3, dtype=F64)).unsqueeze(0).to(dev),
                        pbc=True, atomic_numbers=z.to(dev))


def test_particle_life_kernel_matches_the_oracle(dev):
    """HIP particle life (kind 3) on triclinic soft-sphere packings: forces,
    stress and per-atom stresses against pair_forces_forward; energies zero;
    every system's forces sum to zero (test_particle_life.py:62-80); the
    half-list model gives exactly half the forces and the same stresses."""
    st = tsa.workloads.soft_sphere_batch(3, n_min=64, n_max=128, device=dev, seed=8)
    kw = dict(sigma=1.0, epsilon=1.5, beta=0.3, cutoff=1.0, device=dev, compute_stress=True,
              per_atom_stresses=True)
    model = tsa.ParticleLifeModel(**kw)
    out = model(st)
    fn = functools.partial(om.particle_life_force, A=1.5, beta=0.3, sigma=1.0)
    ref = _oracle_forces(st, fn, 1.0)
    assert torch.equal(out["energy"], torch.zeros_like(out["energy"]))
    _close(out["forces"], ref["forces"], "forces", 1e-5)
    _close(out["stress"], ref["stress"], "stress", 1e-5)
    _close(out["stresses"], ref["stresses"], "stresses", 1e-5)
    scale = out["forces"].abs().max().item()
    for s in range(st.n_systems):
        tot = out["forces"][st.system_idx == s].sum(0).abs().max().item()
        assert tot < 1e-10 * scale, (s, tot)
    half = tsa.PairForcesModel(functools.partial(tsa.particle_life_pair_force, A=1.5, beta=0.3,
                                                 sigma=1.0), cutoff=1.0, device=dev,
                               compute_stress=True, per_atom_stresses=True,
                               reduce_to_half_list=True)(st)
    assert torch.equal(2.0 * half["forces"], out["forces"])
    assert torch.equal(half["stress"], out["stress"])
    _close(half["forces"], _oracle_forces(st, fn, 1.0, half=True)["forces"], "half forces",
           1e-5)


def test_generic_force_function_matches_the_oracle(dev):
    """A force function no kernel knows runs on the generic path: FP64, the
    HIP list's pairs, the reference's scatters; full and half lists."""
    st = tsa.workloads.soft_sphere_batch(2, n_min=64, n_max=96, device=dev, seed=9)

    def force_fn(r, zi, zj):  # a species-dependent, asymmetric force
        return (1.0 - r) * (1.0 + 0.5 * (zi == 1).to(r.dtype) - 0.25 * (zj == 2).to(r.dtype))

    for half in (False, True):
        model = tsa.PairForcesModel(force_fn, cutoff=1.2, device=dev, compute_stress=True,
                                    per_atom_stresses=True, reduce_to_half_list=half)
        out = model(st)
        ref = _oracle_forces(st, force_fn, 1.2, half=half)
        for k in ("forces", "stress", "stresses"):
            _close(out[k], ref[k], f"half={half} {k}", 1e-10)


def test_generic_pair_function_bmhtf_on_rock_salt(dev):
    """The reference's own example potential (pair_potential.py:35-61) on a
    jittered 64-ion NaCl box with a 10 A cutoff (several images of the
    11.3 A cell): energy, forces, stress and per-atom outputs against the
    oracle; the forces sum to zero (test_pair_potential.py:231-243)."""
    st = _nacl(dev)
    fn = functools.partial(bmhtf_pair, **BMHTF)
    model = tsa.PairPotentialModel(fn, cutoff=10.0, device=dev, compute_stress=True,
                                   per_atom_energies=True, per_atom_stresses=True)
    out = model(st)
    ref = _oracle_pairs(st, fn, 10.0, half=False)
    for k in ("energy", "forces", "stress", "energies", "stresses"):
        _close(out[k], ref[k], k, 1e-10)
    assert out["forces"].sum(0).abs().max().item() < 1e-10 * out["forces"].abs().max().item()
    half = tsa.PairPotentialModel(fn, cutoff=10.0, device=dev, compute_stress=True,
                                  reduce_to_half_list=True)(st)
    for k in ("energy", "forces", "stress"):
        _close(half[k], out[k], f"half {k}", 1e-12)


def test_generic_path_agrees_with_the_hip_kernel(dev):
    """Morse as an opaque callable (generic FP64 path) against MorseModel
    (HIP, FP32 pair arithmetic) on a Cu batch."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=64)
    kw = dict(cutoff=6.0, device=dev, compute_stress=True, per_atom_energies=True)
    gen = tsa.PairPotentialModel(lambda r, zi, zj: tsa.morse_pair(r, zi, zj, **MORSE), **kw)
    assert not gen.fused_kernels
    a, b = gen(st), tsa.MorseModel(**MORSE, **kw)(st)
    _close(b["forces"], a["forces"], "forces", 1e-5)
    _close(b["stress"], a["stress"], "stress", 1e-5)
    torch.testing.assert_close(b["energy"], a["energy"], rtol=1e-6, atol=0)
    torch.testing.assert_close(b["energies"], a["energies"], rtol=1e-6, atol=1e-9)


def test_generic_path_retain_graph(dev):
    """retain_graph=True on the generic path keeps the energy attached to
    the positions: dE/dx = -F."""
    st = _nacl(dev, seed=4)
    model = tsa.PairPotentialModel(functools.partial(bmhtf_pair, **BMHTF), cutoff=8.0,
                                   device=dev, retain_graph=True)
    pos = st.positions.clone().requires_grad_(True)
    state = tsa.SimState(positions=pos, masses=st.masses, cell=st.cell, pbc=st.pbc,
                         atomic_numbers=st.atomic_numbers, system_idx=st.system_idx)
    out = model(state)
    assert out["energy"].requires_grad
    out["energy"].sum().backward()
    _close(pos.grad, -out["forces"].detach(), "dE/dx = -F", 1e-10)


class _Producer(tsa.ModelInterface):
    """Energy 1 per system and a per-atom 'latent' output."""

    def __init__(self, dev):
        super().__init__()
        self._device, self._dtype = torch.device(dev), F64
        self._compute_stress = self._compute_forces = False

    def forward(self, state, **_kw):
        return {"energy": torch.ones(state.n_systems, dtype=F64, device=state.device),
                "latent": state.positions[:, 0] + 2.0}


class _Consumer(tsa.ModelInterface):
    """Energy = per-system sum of the earlier model's 'latent'."""

    def __init__(self, dev):
        super().__init__()
        self._device, self._dtype = torch.device(dev), F64
        self._compute_stress = self._compute_forces = False
        self.seen = None

    def forward(self, state, **_kw):
        self.seen = state.latent.clone()
        e = torch.zeros(state.n_systems, dtype=F64, device=state.device)
        return {"energy": e.index_add(0, state.system_idx, state.latent)}


def test_sum_and_serial_sum_models(dev):
    """SumModel adds energy, forces and stress of its children
    (test_sum_model.py:152-162); SerialSumModel equals it for independent
    children and hands earlier outputs to later children as state extras
    without touching the caller's state (test_sum_model.py:186-218)."""
    st = tsa.workloads.cu_batch(3, device=dev, seed=65)
    lj = tsa.LennardJonesModel(sigma=2.3, epsilon=0.2, cutoff=5.5, device=dev,
                               compute_stress=True)
    mo = tsa.MorseModel(**MORSE, cutoff=6.0, device=dev, compute_stress=True)
    a, b = lj(st), mo(st)
    out = tsa.SumModel(lj, mo)(st)
    for k in ("energy", "forces", "stress"):
        assert torch.equal(out[k], a[k] + b[k]), k
    ser = tsa.SerialSumModel(lj, mo)(st)
    for k in ("energy", "forces", "stress"):
        assert torch.equal(ser[k], out[k]), k
    cons = _Consumer(dev)
    got = tsa.SerialSumModel(_Producer(dev), cons)(st)
    latent = st.positions[:, 0] + 2.0
    assert torch.equal(cons.seen, latent) and torch.equal(got["latent"], latent)
    want = torch.ones(st.n_systems, dtype=F64, device=dev).index_add(0, st.system_idx, latent)
    # index_add sums with atomics in no fixed order: equal up to rounding
    torch.testing.assert_close(got["energy"], want, rtol=1e-13, atol=0)
    assert not st.has_extras("latent")


def test_integrate_runs_a_sum_model_step_by_step(dev):
    """integrate() with a model the fused engine cannot run takes the
    per-step path: the hand-written loop of nvt_langevin_step, to the last
    bits (the generic force model's index_add sums in no fixed order)."""
    st = tsa.workloads.cu_batch(2, device=dev, seed=66)
    model = tsa.SumModel(tsa.MorseModel(**MORSE, cutoff=6.0, device=dev),
                         tsa.PairForcesModel(lambda r, zi, zj: 0.01 * (6.0 - r), cutoff=6.0,
                                             device=dev))
    st.rng = 12
    out = tsa.integrate(st, model, integrator=tsa.Integrator.nvt_langevin, n_steps=6,
                        temperature=600.0, timestep=0.002)
    s2 = tsa.workloads.cu_batch(2, device=dev, seed=66)
    s2.rng = 12
    kt = 600.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    md = tsa.nvt_langevin_init(s2, model, kT=kt)
    for _ in range(6):
        md = tsa.nvt_langevin_step(md, model, dt=dt, kT=kt)
    dx = (out.positions - md.positions).abs().max().item()
    dp = (out.momenta - md.momenta).abs().max().item()
    print(f"integrate vs loop: dx {dx:.2e} dp {dp:.2e}", flush=True)
    assert dx < 1e-10 and dp < 1e-10
    with pytest.raises(TypeError, match="closed-form"):
        tsa.BatchedMD(tsa.PairPotentialModel(lambda r, zi, zj: r, cutoff=6.0, device=dev), md,
                      dt=dt)
