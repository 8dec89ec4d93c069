"""Host-side checks of the model surface (no device work): recognised and
refused pair functions, neighbor-list functions, the pair functions' own
values, and the loud refusal of CPU tensors (there is no CPU path)."""

import functools

import pytest
import torch

import torch_sim_amd as tsa
from torch_sim_amd import _lib
from torch_sim_amd.models import LJ, MORSE, SOFT_SPHERE, resolve_pair_fn

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


@pytest.mark.parametrize("bad", [
    lambda d, zi, zj: d,
    functools.partial(tsa.morse_pair, 1.0),  # dr bound positionally
    functools.partial(tsa.lennard_jones_pair, sigma=1.0, alpha=2.0),  # LJ has no alpha
])
def test_other_pair_functions_are_refused(bad):
    with pytest.raises(TypeError):
        tsa.PairPotentialModel(bad, cutoff=2.0, device="cpu")


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
