"""Host-side checks of the model surface (no device work): recognised and
refused pair functions, neighbor-list functions, the pair functions' own
values, and the loud refusal of CPU tensors (there is no CPU path)."""

import functools

import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd import _lib
from torch_sim_amd.models import (GENERIC, LJ, MORSE, PARTICLE_LIFE, SOFT_SPHERE,
                                  resolve_force_fn, resolve_pair_fn)

F64 = torch.float64


def test_resolve_closed_forms_and_defaults():
    assert resolve_pair_fn(functools.partial(tsa.morse_pair, sigma=2.0)) == (MORSE, 2.0, 5.0, 5.0,
                                                                              None)
    assert resolve_pair_fn(tsa.lennard_jones_pair)[:3] == (LJ, 1.0, 1.0)
    k, s, e, a, z = resolve_pair_fn(functools.partial(tsa.soft_sphere_pair, alpha=3.0))
    assert (k, s, e, a, z) == (SOFT_SPHERE, 1.0, 1.0, 3.0, None)
    fn = tsa.MultiSoftSpherePairFn(torch.tensor([18, 36]), torch.tensor([[1.0, 2.0], [2.0, 3.0]]),
                                   torch.ones(2, 2))
    k, s, e, a, z = resolve_pair_fn(fn)
    assert k == SOFT_SPHERE and s.shape == (2, 2) and z.tolist() == [18, 36]


@pytest.mark.parametrize("other", [
    lambda d, zi, zj: d,
    functools.partial(tsa.morse_pair, 1.0),  # dr bound positionally: a user callable
])
def test_other_pair_functions_take_the_generic_path(other):
    """Callables that are not a recognised closed form run on the generic
    path (HIP list, torch pair arithmetic); the fused engines refuse them."""
    with pytest.raises(TypeError):
        resolve_pair_fn(other)
    model = tsa.PairPotentialModel(other, cutoff=2.0, device="cpu")
    assert model.kind == GENERIC and not model.fused_kernels
    assert tsa.MorseModel(device="cpu").fused_kernels


def test_a_keyword_the_pair_function_does_not_take_is_refused():
    with pytest.raises(TypeError, match="alpha"):
        tsa.PairPotentialModel(functools.partial(tsa.lennard_jones_pair, sigma=1.0, alpha=2.0),
                               cutoff=2.0, device="cpu")
    with pytest.raises(TypeError, match="gamma"):
        tsa.PairForcesModel(functools.partial(tsa.particle_life_pair_force, gamma=1.0),
                            cutoff=1.0, device="cpu")


def test_force_functions_route_and_particle_life_values():
    """particle_life_pair_force against test_particle_life.py:13-35 and its
    closed form: r/beta - 1 inside beta, the tent A (1 - |2r - 1 - beta| /
    (1 - beta)) up to sigma (A at its peak r = (1 + beta) / 2), zero beyond."""
    z = torch.ones(4, dtype=torch.long)
    f = tsa.particle_life_pair_force
    r = torch.tensor([0.1, 0.2, 1.0, 1.5], dtype=F64)
    out = f(r, z, z, A=1.0, beta=0.3, sigma=1.0)
    assert (out[:2] < 0).all() and (out[2:] == 0).all()
    torch.testing.assert_close(out[:2], r[:2] / 0.3 - 1.0, rtol=0, atol=1e-15)
    r6 = torch.tensor([0.6], dtype=F64)
    torch.testing.assert_close(f(r6, z[:1], z[:1], A=3.0, beta=0.3, sigma=1.0),
                               3.0 * f(r6, z[:1], z[:1], A=1.0, beta=0.3, sigma=1.0))
    peak = torch.tensor([0.65], dtype=F64)
    assert abs(f(peak, z[:1], z[:1], A=2.5, beta=0.3, sigma=1.0).item() - 2.5) < 1e-15
    assert abs(f(torch.tensor([0.3], dtype=F64), z[:1], z[:1], A=2.0, beta=0.3,
                 sigma=1.0).item()) < 1e-15  # the tent starts at zero
    m = tsa.ParticleLifeModel(sigma=1.2, epsilon=2.0, beta=0.25, device="cpu")
    assert m.kind == PARTICLE_LIFE and m.cutoff == 3.0 and not m.fused_kernels
    assert resolve_force_fn(m.force_fn) == (PARTICLE_LIFE, 1.2, 2.0, 0.25)
    assert resolve_force_fn(lambda d, zi, zj: d)[0] == GENERIC


def test_sum_model_construction_rules():
    """interface.py:227-250: two or more children, one device and dtype;
    compute_stress / compute_forces are the children's conjunction."""
    a = tsa.MorseModel(device="cpu", compute_stress=True)
    b = tsa.LennardJonesModel(device="cpu", compute_stress=True)
    with pytest.raises(ValueError, match="at least two"):
        tsa.SumModel(a)
    with pytest.raises(ValueError, match="Dtype mismatch"):
        tsa.SumModel(a, tsa.LennardJonesModel(device="cpu", dtype=torch.float32))
    with pytest.raises(ValueError, match="Device mismatch"):
        tsa.SumModel(a, tsa.LennardJonesModel(device="meta"))
    sm = tsa.SumModel(a, b)
    assert sm.compute_stress and sm.compute_forces and sm.retain_graph is False
    assert not tsa.SumModel(a, tsa.LennardJonesModel(device="cpu")).compute_stress
    sm.retain_graph = True
    assert a.retain_graph and b.retain_graph and sm.retain_graph
    with pytest.raises(NotImplementedError):  # the HIP models fix compute_stress
        sm.compute_stress = True


def test_state_extras_follow_the_reference_rules():
    """store_model_extras (state.py:376-404): per-atom by leading n_atoms,
    per-system by n_systems, canonical keys and scalars skipped; extras read
    as attributes, survive clone / split / concatenate / select."""
    st = tsa.workloads.cu_batch(2, device="cpu", seed=3, reps=[(2, 2, 2), (2, 2, 3)])
    n, s = st.n_atoms, st.n_systems
    out = {"energy": torch.zeros(s), "forces": torch.zeros(n, 3), "latent": torch.arange(n),
           "charge": torch.ones(s, 2), "scalar": torch.tensor(1.0), "odd": torch.zeros(7)}
    st.store_model_extras(out)
    assert st.has_extras("latent") and st.has_extras("charge")
    assert not st.has_extras("scalar") and not st.has_extras("odd")
    assert not st.has_extras("energy") and not hasattr(st, "odd")
    assert torch.equal(st.latent, torch.arange(n))
    st.latent = torch.arange(n) * 2  # a write updates the extra
    assert torch.equal(st.atom_extras["latent"], torch.arange(n) * 2)
    c = st.clone()
    c.atom_extras["latent"].zero_()
    assert st.latent.sum() > 0
    parts = st.split()
    assert parts[1].latent.shape[0] == parts[1].n_atoms and parts[1].charge.shape == (1, 2)
    back = tsa.concatenate_states(parts)
    assert torch.equal(back.latent, st.latent) and torch.equal(back.charge, st.charge)
    sel = tsa.select_systems(st, [1])
    assert torch.equal(sel.latent, parts[1].latent)


def test_custom_neighbor_list_fn_is_refused():
    with pytest.raises(NotImplementedError, match="torchsim_nl"):
        tsa.MorseModel(cutoff=6.0, device="cpu", neighbor_list_fn=lambda **kw: None)


def test_pair_functions_match_the_reference_values():
    """test_soft_sphere.py:36-51, test_lennard_jones.py:26-31, test_morse.py:22-27."""
    z = torch.ones(1, dtype=torch.long)
    r = torch.tensor([0.5], dtype=F64)
    assert tsa.soft_sphere_pair(r, z, z).item() == 0.125
    r_min = torch.tensor([2 ** (1 / 6)], dtype=F64)
    assert abs(tsa.lennard_jones_pair(r_min, z, z).item() + 1.0) < 1e-14
    assert abs(tsa.morse_pair(torch.tensor([1.0], dtype=F64), z, z).item() + 5.0) < 1e-14
    fn = tsa.MultiSoftSpherePairFn(torch.tensor([18, 36]), torch.tensor([[1.0, 2.0], [2.0, 3.0]]),
                                   torch.ones(2, 2))
    e = fn(torch.full((2,), 0.5), torch.tensor([18, 18]), torch.tensor([18, 36]))
    assert e.tolist() == [0.125, 0.28125]
    with pytest.raises(ValueError, match="sigma_matrix"):
        tsa.MultiSoftSpherePairFn(torch.tensor([18, 36]), torch.ones(3, 3), torch.ones(2, 2))


def test_property_setters_raise_and_cpu_state_is_refused():
    model = tsa.MorseModel(cutoff=6.0, device="cpu")
    for name in ("device", "dtype", "compute_stress", "compute_forces"):
        with pytest.raises(NotImplementedError):
            setattr(model, name, getattr(model, name))
    st = tsa.workloads.cu_batch(1, device="cpu", seed=1)
    with pytest.raises(_lib.TsxError, match="ROCm GPU"):
        model(st)
    with pytest.raises(TypeError, match="dtype"):
        model(st.to(dtype=torch.float32))
