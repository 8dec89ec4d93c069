"""GPU parity of the pair models against the FP64 oracle.

Tolerances (FP32 pair arithmetic, FP64 accumulation):
  forces   max |dF| <= 1e-5 * max |F_ref| (plus 1e-9 absolute for F ~ 0)
  energy   relative 1e-6 per system
  stress   max |d sigma| <= 1e-5 * max |sigma_ref| + 1e-9
"""

import math

import pytest
import torch

import torch_sim_amd as tsa
from oracle import models as om

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _ref(state, fn, cutoff, per_atom=False):
    return om.pair_forward(
        state.positions.cpu(), state.cell.cpu(), state.pbc.cpu(), state.atomic_numbers.cpu(),
        state.system_idx.cpu(), fn, cutoff, per_atom=per_atom,
    )


def _check(out, ref, f_tol=1e-5, e_tol=1e-6):
    fr = ref["forces"]
    df = (out["forces"].cpu() - fr).abs().max().item()
    assert df <= f_tol * max(fr.abs().max().item(), 1e-3) + 1e-9, df
    er = ref["energy"]
    de = ((out["energy"].cpu() - er).abs() / er.abs().clamp_min(1e-12)).max().item()
    assert de <= e_tol, de
    if "stress" in out:
        sr = ref["stress"]
        ds = (out["stress"].cpu() - sr).abs().max().item()
        assert ds <= 1e-5 * sr.abs().max().item() + 1e-9, ds


def test_lj_argon_256(dev):
    st = tsa.workloads.ar_fcc(dev, jitter=0.05, seed=1)
    model = tsa.LennardJonesModel(sigma=3.405, epsilon=0.0104, cutoff=8.5125, device=dev, compute_stress=True)
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 3.405, 0.0104)  # noqa: E731
    _check(model(st), _ref(st, fn, 8.5125))


def test_morse_cu_batch(dev):
    st = tsa.workloads.cu_batch(6, device=dev, seed=11)
    model = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, compute_stress=True)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, 2.866, 0.3429, 1.3588)  # noqa: E731
    _check(model(st), _ref(st, fn, 6.0))


def test_soft_sphere_multi_triclinic_per_atom(dev):
    st = tsa.workloads.soft_sphere_batch(4, n_min=64, n_max=160, device=dev, seed=2)
    sig = torch.tensor([[1.0, 1.2], [1.2, 1.4]], dtype=F64)
    model = tsa.SoftSphereMultiModel(
        atomic_numbers=torch.tensor([1, 2]), sigma_matrix=sig, device=dev, compute_stress=True,
        per_atom_energies=True, per_atom_stresses=True,
    )
    fn = om.MultiPairFn([1, 2], sig, torch.ones(2, 2, dtype=F64), torch.full((2, 2), 2.0, dtype=F64))
    out = model(st)
    ref = _ref(st, fn, 1.4, per_atom=True)
    _check(out, ref)
    torch.testing.assert_close(out["energies"].cpu(), ref["energies"], rtol=1e-5, atol=1e-9)
    torch.testing.assert_close(out["stresses"].cpu(), ref["stresses"], rtol=1e-4, atol=1e-8)


def test_soft_sphere_general_alpha(dev):
    st = tsa.workloads.soft_sphere_batch(2, n_min=64, n_max=96, device=dev, seed=3)
    model = tsa.SoftSphereModel(sigma=1.3, epsilon=0.7, alpha=2.5, device=dev, compute_stress=True)
    fn = lambda d, zi, zj: om.soft_sphere_pair(d, zi, zj, 1.3, 0.7, 2.5)  # noqa: E731
    _check(model(st), _ref(st, fn, 1.3), f_tol=2e-5)


@pytest.mark.parametrize("a,rc,kind", [(5.26, 8.5125, "lj"), (3.61, 6.0, "morse")])
def test_perfect_lattice_closed_form(dev, a, rc, kind):
    shells = [(1, 12), (2, 6), (3, 24), (4, 12), (5, 24), (6, 8), (7, 48), (8, 6)]
    if kind == "lj":
        pair = lambda r: 4 * 0.0104 * ((3.405 / r) ** 12 - (3.405 / r) ** 6)  # noqa: E731
        model = tsa.LennardJonesModel(sigma=3.405, epsilon=0.0104, cutoff=rc, device=dev)
        st = tsa.workloads.ar_fcc(dev, reps=(3, 3, 3), a=a)
    else:
        pair = lambda r: 0.3429 * (1 - math.exp(-1.3588 * (r - 2.866))) ** 2 - 0.3429  # noqa: E731
        model = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=rc, device=dev)
        st = tsa.workloads.cu_batch(1, device=dev, jitter=0.0, reps=[[3, 3, 3]])
    e_shell = 0.5 * sum(n * pair(a * math.sqrt(k / 2)) for k, n in shells if a * math.sqrt(k / 2) < rc)
    out = model(st)
    assert abs(out["energy"].item() / st.n_atoms - e_shell) <= 1e-6 * abs(e_shell)
    assert out["forces"].abs().max().item() < 1e-6


@pytest.mark.parametrize("tpa", [1, 2, 4, 8])
def test_threads_per_atom_agree(dev, tpa):
    st = tsa.workloads.cu_batch(3, device=dev, seed=4)
    m1 = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, tpa=1, compute_stress=True)
    mk = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, tpa=tpa, compute_stress=True)
    a, b = m1(st), mk(st)
    for k in ("energy", "forces", "stress"):
        torch.testing.assert_close(a[k], b[k], rtol=1e-12, atol=1e-13)


def test_batch_equals_single_bitwise_and_no_mutation(dev):
    st = tsa.workloads.cu_batch(3, device=dev, seed=6)
    model = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, compute_stress=True)
    pos0, cell0 = st.positions.clone(), st.cell.clone()
    full = model(st)
    assert torch.equal(st.positions, pos0) and torch.equal(st.cell, cell0)
    a = 0
    for k, single in enumerate(st.split()):
        out = model(single)
        n = single.n_atoms
        assert torch.equal(out["energy"][0], full["energy"][k])
        assert torch.equal(out["forces"], full["forces"][a : a + n])
        assert torch.equal(out["stress"][0], full["stress"][k])
        a += n


def test_lattice_translation_invariance(dev):
    st = tsa.workloads.cu_batch(1, device=dev, seed=9, reps=[[2, 3, 2]])
    model = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, compute_stress=True)
    ref = model(st)
    moved = st.clone()
    moved.positions[0] = moved.positions[0] + 3 * moved.cell[0, :, 0]
    moved.positions[5] = moved.positions[5] - 2 * moved.cell[0, :, 2]
    out = model(moved)
    for k in ("energy", "forces", "stress"):
        torch.testing.assert_close(out[k], ref[k], rtol=1e-9, atol=1e-9)


def test_dtype_guard_and_float32(dev):
    st = tsa.workloads.cu_batch(2, device=dev, seed=1)
    model = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev)
    with pytest.raises(TypeError):
        tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev,
                       dtype=torch.float32)(st)
    m32 = tsa.MorseModel(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev, dtype=torch.float32)
    out32 = m32(st.to(dtype=torch.float32))
    out64 = model(st)
    assert out32["forces"].dtype == torch.float32
    torch.testing.assert_close(out32["energy"].double(), out64["energy"], rtol=1e-5, atol=1e-4)


def test_non_periodic_molecule(dev):
    g = torch.Generator().manual_seed(0)
    pos = torch.randn(12, 3, generator=g, dtype=F64) * 1.5
    st = tsa.SimState(positions=pos.to(dev), masses=torch.ones(12, dtype=F64, device=dev),
                      cell=torch.zeros(1, 3, 3, dtype=F64, device=dev), pbc=False,
                      atomic_numbers=torch.ones(12, dtype=torch.long, device=dev))
    model = tsa.LennardJonesModel(sigma=1.0, epsilon=1.0, cutoff=2.5, device=dev)
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 1.0, 1.0)  # noqa: E731
    ref = om.pair_forward(pos, torch.zeros(1, 3, 3, dtype=F64), torch.tensor([False] * 3),
                          torch.ones(12, dtype=torch.long), torch.zeros(12, dtype=torch.long), fn, 2.5,
                          compute_stress=False)
    out = model(st)
    _check(out, ref)


def test_verlet_skin_matches_rebuild_every_call(dev):
    """A model keeping a Verlet list with skin (gated per-system rebuilds,
    strain-aware validity check) gives the forces, energies and stresses of a
    list rebuilt at every call, through displacements and cell strains."""
    st = tsa.workloads.soft_sphere_batch(6, n_min=64, n_max=200, device=dev, seed=21)
    sig = torch.tensor([[1.0, 1.2], [1.2, 1.4]], dtype=F64)
    kw = dict(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=sig, device=dev, compute_stress=True)
    m_skin = tsa.SoftSphereMultiModel(**kw, skin=0.3)
    m_ref = tsa.SoftSphereMultiModel(**kw)
    g = torch.Generator(device=dev).manual_seed(0)
    x = st.clone()
    for k in range(8):
        if k:
            x.positions = x.positions + 0.04 * torch.randn(x.positions.shape, generator=g,
                                                           device=dev, dtype=F64)
            strain = torch.eye(3, dtype=F64, device=dev) + 0.01 * torch.randn(
                (x.n_systems, 3, 3), generator=g, device=dev, dtype=F64)
            x.set_cell(torch.bmm(strain, x.cell), scale_atoms=True)
        a, b = m_skin(x), m_ref(x)
        torch.testing.assert_close(a["forces"], b["forces"], rtol=1e-12, atol=1e-12)
        torch.testing.assert_close(a["energy"], b["energy"], rtol=1e-12, atol=1e-12)
        torch.testing.assert_close(a["stress"], b["stress"], rtol=1e-12, atol=1e-12)
    # one full build at the start, more only on capacity overflow
    assert m_skin._vl.updates == 8 and m_skin._vl.full_builds <= 3


def test_verlet_overflow_during_call_recomputes_forces(dev):
    """The build status of a skin model is read once, after its force kernels
    are queued.  A gated rebuild that overflows the rows (atoms pulled
    together, same cells) must make the call rebuild every system with more
    room and compute the forces again: they equal those of a fresh list."""
    st = tsa.workloads.cu_batch(2, device=dev, seed=31)
    kw = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588, cutoff=6.0, device=dev)
    m_skin = tsa.MorseModel(**kw, skin=0.5)
    m_skin(st)
    builds0 = m_skin._vl.full_builds
    x = st.clone()
    x.positions = 0.8 * st.positions  # about twice the neighbours per row
    a, b = m_skin(x), tsa.MorseModel(**kw)(x)
    assert m_skin._vl.full_builds > builds0  # the overflow path ran
    torch.testing.assert_close(a["forces"], b["forces"], rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(a["energy"], b["energy"], rtol=1e-12, atol=1e-12)
