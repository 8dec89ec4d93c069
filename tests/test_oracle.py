"""Pin the CPU oracle against the reference's known answers, closed forms,
scipy and brute force (CPU only)."""

import itertools
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import fire as ofire
from oracle import mat33, md as omd, models as om, nl as onl

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
F64 = torch.float64


def _t(x, dtype=F64):
    return torch.tensor(x, dtype=dtype)


def _single(pos, cell_row, pbc=True):
    pos = _t(pos)
    n = pos.shape[0]
    return pos, _t(cell_row).view(1, 3, 3), torch.tensor([[pbc] * 3]), torch.zeros(n, dtype=torch.long)


def _cacrp2o7():
    g = GOLD["cacrp2o7"]
    return _single(g["positions"], g["cell_row"])


def fcc(a, reps):
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    g = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"), -1).reshape(-1, 1, 3)
    return torch.tensor(((g + basis[None]).reshape(-1, 3) * a))


def shell_energy(pair, a, rc):
    """Closed-form fcc lattice energy per atom: 0.5 sum over shells n_k V(r_k)."""
    shells = [(1, 12), (2, 6), (3, 24), (4, 12), (5, 24), (6, 8), (7, 48), (8, 6)]
    return 0.5 * sum(n * pair(a * math.sqrt(k / 2)) for k, n in shells if a * math.sqrt(k / 2) < rc)


# ------------------------------------------------------------------ wrap
def test_wrap_reconstructs_positions():
    g = GOLD["wrap_batched"]
    cell = _t(g["cell_row"])
    pos = _t(g["positions"])
    idx = torch.tensor(g["system_idx"])
    wrapped, shifts = onl.wrap_positions(pos, cell, idx, torch.tensor(g["pbc"]))
    rec = wrapped + (shifts.unsqueeze(-1) * cell[idx]).sum(dim=1)
    torch.testing.assert_close(rec, pos, atol=1e-10, rtol=0)
    assert (shifts[0] != 0).any()


def test_wrap_non_periodic_axis():
    g = GOLD["wrap_non_periodic_axis"]
    w, s = onl.wrap_positions(_t(g["positions"]), _t(g["cell_row"]), torch.zeros(1, dtype=torch.long), torch.tensor(g["pbc"]))
    assert w.tolist() == g["expected_wrapped"]
    assert s.tolist() == g["expected_shifts"]


def test_wrap_singular_cell_unchanged():
    g = GOLD["wrap_singular_cell"]
    w, s = onl.wrap_positions(_t(g["positions"]), torch.zeros(1, 3, 3, dtype=F64), torch.zeros(1, dtype=torch.long), torch.tensor([[True] * 3]))
    assert w.tolist() == g["expected_wrapped"] and (s == 0).all()


@pytest.mark.parametrize("reps,count", GOLD["cell_shift_counts"]["cases"])
def test_shift_set_sizes(reps, count):
    assert math.prod(2 * r + 1 for r in reps) == count


def test_num_cell_repeats_closed_form():
    # cubic cell of side L: ceil(rc / L) on periodic axes, 0 elsewhere
    cell = torch.eye(3, dtype=F64).unsqueeze(0) * 4.0
    n = onl.num_cell_repeats(9.0, cell, torch.tensor([[True, False, True]]))
    assert n.tolist() == [[3, 0, 3]]


# ---------------------------------------------------------------- pairs
def test_ar_fcc_48_pairs():
    g = GOLD["ar_fcc_pairs"]
    pos, cell, pbc, idx = _single(fcc(g["a"], (1, 1, 1)).tolist(), (np.eye(3) * g["a"]).tolist())
    for fn in (onl.brute_force_nl, onl.linked_cell_nl):
        m, _, _ = fn(pos, cell, pbc, g["cutoff"], idx)
        assert m.shape[1] == g["n_pairs"]


@pytest.mark.parametrize("a,rc", [(5.26, 8.5125), (3.61, 6.0)])
def test_fcc_78_neighbour_shells(a, rc):
    pos, cell, pbc, idx = _single(fcc(a, (2, 2, 2)).tolist(), (np.eye(3) * 2 * a).tolist())
    m, _, _ = onl.brute_force_nl(pos, cell, pbc, rc, idx)
    assert m.shape[1] == 78 * pos.shape[0]


def _sorted_mic(pos, cell, m, s, sh):
    cs = torch.einsum("jn,jnm->jm", sh, cell[s])
    return torch.sort((pos[m[1]] - pos[m[0]] + cs).norm(dim=1)).values


@pytest.mark.parametrize("cutoff", [2.0, 5.0, 7.0])
@pytest.mark.parametrize("self_interaction", [False, True])
def test_brute_force_equals_linked_cell_triclinic(cutoff, self_interaction):
    pos, cell, pbc, idx = _cacrp2o7()
    a = onl.brute_force_nl(pos, cell, pbc, cutoff, idx, self_interaction)
    b = onl.sort_pairs(*onl.linked_cell_nl(pos, cell, pbc, cutoff, idx, self_interaction))
    assert torch.equal(a[0], b[0]) and torch.equal(a[2], b[2])


@pytest.mark.parametrize("cutoff", [2.0, 5.0, 7.0])
def test_lattice_image_invariance(cutoff):
    pos, cell, pbc, idx = _cacrp2o7()
    n = pos.shape[0]
    ar = torch.arange(n)
    ints = torch.stack([(ar % 3) - 1, (ar % 5) - 2, (ar % 4) - 2], 1).to(F64)
    shifted = pos + ints @ cell[0]
    a = onl.brute_force_nl(pos, cell, pbc, cutoff, idx)
    b = onl.brute_force_nl(shifted, cell, pbc, cutoff, idx)
    assert a[0].shape == b[0].shape
    torch.testing.assert_close(_sorted_mic(pos, cell, *a), _sorted_mic(shifted, cell, *b), atol=1e-9, rtol=0)


def test_nl_edge_cases_and_empty():
    g = GOLD["nl_edge_case"]
    for pbc in (True, False):
        for si in (True, False):
            pos, cell, p, idx = _single(g["positions"], (np.eye(3) * g["cell_diag"]).tolist(), pbc)
            m, s, _ = onl.brute_force_nl(pos, cell, p, g["cutoff"], idx, si)
            assert m.shape[1] > 0 and (s == 0).all()
    g = GOLD["nl_empty"]
    pos, cell, p, idx = _single(g["positions"], (np.eye(3) * g["cell_diag"]).tolist(), False)
    m, s, sh = onl.brute_force_nl(pos, cell, p, g["cutoff"], idx)
    assert m.shape == (2, 0) and s.numel() == 0 and sh.shape == (0, 3)


def test_full_to_half_list_kat():
    g = GOLD["half_list"]
    m, s, sh = onl.full_to_half_list(torch.tensor(g["mapping"]), torch.tensor(g["system_mapping"]), _t(g["shifts"]))
    assert m.shape[1] == 2
    assert s.tolist() == g["expected_system_mapping"]
    assert sh.tolist() == g["expected_shifts"]


def test_half_list_self_image_rule():
    # i == j pairs keep the lexicographically positive shift
    m = torch.tensor([[0, 0, 0, 0], [0, 0, 0, 0]])
    sh = _t([[1, 0, 0], [-1, 0, 0], [0, -1, 1], [0, 1, -1]])
    hm, _, hs = onl.full_to_half_list(m, torch.zeros(4, dtype=torch.long), sh)
    assert hs.tolist() == [[1, 0, 0], [0, 1, -1]]


# ------------------------------------------------------------- models
def test_pair_closed_forms():
    g = GOLD["pair_closed_forms"]
    z = torch.ones(1, dtype=torch.long)
    c = g["lj_min_energy"]
    assert abs(om.lennard_jones_pair(_t([c["r"]]), z, z, c["sigma"], c["epsilon"]).item() - c["energy"]) < 1e-12
    c = g["morse_min_energy"]
    assert abs(om.morse_pair(_t([c["r"]]), z, z, c["sigma"], c["epsilon"], c["alpha"]).item() - c["energy"]) < 1e-12
    c = g["soft_sphere_half"]
    assert om.soft_sphere_pair(_t([c["r"]]), z, z, c["sigma"], c["epsilon"], c["alpha"]).item() == c["energy"]
    assert om.soft_sphere_pair(_t([1.0, 1.5]), z.expand(2), z.expand(2)).tolist() == [0.0, 0.0]
    assert om.lennard_jones_pair(_t([0.0]), z, z).item() == 0.0


@pytest.mark.parametrize(
    "pair,a,rc",
    [
        (lambda r: 4 * 0.0104 * ((3.405 / r) ** 12 - (3.405 / r) ** 6), 5.26, 8.5125),
        (lambda r: 0.3429 * (1 - math.exp(-1.3588 * (r - 2.866))) ** 2 - 0.3429, 3.61, 6.0),
    ],
)
def test_perfect_lattice_energy_matches_shell_sum(pair, a, rc):
    pos = fcc(a, (3, 3, 3))
    n = pos.shape[0]
    cell = (torch.eye(3, dtype=F64) * 3 * a).unsqueeze(0)
    fn = (lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 3.405, 0.0104)) if a == 5.26 else (
        lambda d, zi, zj: om.morse_pair(d, zi, zj, 2.866, 0.3429, 1.3588))
    out = om.pair_forward(pos, cell, torch.tensor([True] * 3), torch.ones(n, dtype=torch.long), torch.zeros(n, dtype=torch.long), fn, rc)
    assert abs(out["energy"].item() / n - shell_energy(pair, a, rc)) < 1e-12
    assert out["forces"].abs().max() < 1e-12


def _jittered_system(seed=0):
    g = torch.Generator().manual_seed(seed)
    pos, cell, pbc, idx = _cacrp2o7()
    pos = pos + 0.05 * torch.randn(pos.shape, generator=g, dtype=F64)
    return pos, cell, pbc, idx


def test_forces_and_stress_are_gradients():
    pos, cell_row, pbc, idx = _jittered_system()
    cell = cell_row.transpose(1, 2).contiguous()
    n = pos.shape[0]
    z = torch.ones(n, dtype=torch.long)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, 2.0, 0.5, 1.5)  # noqa: E731
    out = om.pair_forward(pos, cell, pbc[0], z, idx, fn, 5.0)
    # F = -dE/dx by central differences
    h = 1e-5
    for i, k in [(0, 0), (3, 1), (11, 2)]:
        p1, p2 = pos.clone(), pos.clone()
        p1[i, k] += h
        p2[i, k] -= h
        e1 = om.pair_forward(p1, cell, pbc[0], z, idx, fn, 5.0)["energy"]
        e2 = om.pair_forward(p2, cell, pbc[0], z, idx, fn, 5.0)["energy"]
        assert abs(-(e1 - e2).item() / (2 * h) - out["forces"][i, k].item()) < 1e-6
    # stress = (1/V) dE/d(strain) for a homogeneous deformation
    vol = torch.det(cell[0]).abs()
    for a, b in [(0, 0), (1, 2)]:
        eps = torch.zeros(3, 3, dtype=F64)
        eps[a, b] = eps[b, a] = h
        def e_at(sgn):
            d = torch.eye(3, dtype=F64) + sgn * eps
            return om.pair_forward(pos @ d.T, (d @ cell[0]).unsqueeze(0), pbc[0], z, idx, fn, 5.0)["energy"].item()
        deriv = (e_at(1) - e_at(-1)) / (2 * h)
        factor = 1.0 if a == b else 2.0
        assert abs(deriv / vol.item() / factor - out["stress"][0, a, b].item()) < 1e-6
    assert out["forces"].sum(0).abs().max() < 1e-12
    torch.testing.assert_close(out["stress"][0], out["stress"][0].T)


def test_half_equals_full():
    pos, cell_row, pbc, idx = _jittered_system(1)
    cell = cell_row.transpose(1, 2).contiguous()
    z = torch.ones(pos.shape[0], dtype=torch.long)
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 2.0, 0.3)  # noqa: E731
    a = om.pair_forward(pos, cell, pbc[0], z, idx, fn, 5.0, half=True, per_atom=True)
    b = om.pair_forward(pos, cell, pbc[0], z, idx, fn, 5.0, half=False, per_atom=True)
    for k in ("energy", "forces", "stress", "energies", "stresses"):
        torch.testing.assert_close(a[k], b[k], rtol=1e-10, atol=1e-12)


# ----------------------------------------------------------------- md
def test_baoab_one_step_closed_form():
    m = _t([2.0])
    x0, p0 = _t([[0.1, 0.2, 0.3]]), _t([[0.5, -0.2, 0.1]])
    xi = _t([[0.3, -1.1, 0.7]])
    dt, kT, gamma = 0.2, 0.05, 0.3
    zero = lambda x: (torch.zeros(1, dtype=F64), torch.zeros_like(x))  # noqa: E731
    x1, p1, _, _ = omd.langevin_step(x0, p0, torch.zeros_like(x0), m, torch.zeros(1, dtype=torch.long), zero, dt, kT, gamma, xi)
    c1 = math.exp(-gamma * dt)
    c2 = math.sqrt(kT * (1 - c1 * c1))
    p_exp = c1 * p0 + c2 * math.sqrt(2.0) * xi
    x_exp = x0 + p0 / 2.0 * dt / 2 + p_exp / 2.0 * dt / 2
    torch.testing.assert_close(p1, p_exp, rtol=1e-14, atol=0)
    torch.testing.assert_close(x1, x_exp, rtol=1e-14, atol=1e-16)


def test_initialize_momenta_removes_com():
    g = torch.Generator().manual_seed(0)
    idx = torch.tensor([0, 0, 0, 1, 2, 2])
    m = _t([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    noise = torch.randn(6, 3, generator=g, dtype=F64)
    p = omd.initialize_momenta(m, idx, _t([0.1, 0.2, 0.3]), noise)
    assert p[:3].sum(0).abs().max() < 1e-14 and p[4:].sum(0).abs().max() < 1e-14
    torch.testing.assert_close(p[3], noise[3] * math.sqrt(4.0 * 0.2))


def test_nve_second_order_energy_conservation():
    """Velocity Verlet on a cutoff-shifted LJ crystal conserves energy with
    an error that scales as dt^2 (the reference's truncated potential adds
    jumps at the cutoff, so the shift isolates the integrator)."""
    pos = fcc(5.26, (3, 3, 3))
    n = pos.shape[0]
    cell = (torch.eye(3, dtype=F64) * 3 * 5.26).unsqueeze(0)
    z, idx = torch.ones(n, dtype=torch.long), torch.zeros(n, dtype=torch.long)
    m = torch.full((n,), 39.948, dtype=F64)
    vc = 4 * 0.0104 * ((1 / 2.5) ** 12 - (1 / 2.5) ** 6)
    fn = lambda d, zi, zj: om.lennard_jones_pair(d, zi, zj, 3.405, 0.0104) - vc * (d < 8.5125)  # noqa: E731

    def force(x):
        o = om.pair_forward(x, cell, torch.tensor([True] * 3), z, idx, fn, 8.5125, compute_stress=False)
        return o["energy"], o["forces"]

    drifts = []
    for dt_fs in (2.0, 1.0):
        g = torch.Generator().manual_seed(0)
        p = omd.initialize_momenta(m, idx, _t(50.0 * omd.KB_EV), torch.randn(n, 3, generator=g, dtype=F64))
        e, f = force(pos)
        x = pos
        dt = dt_fs * 1e-3 * omd.TIME_UNIT
        e0 = (e + omd.kinetic_energy(p, m, idx, 1)).item()
        for _ in range(int(100 / dt_fs)):
            x, p, f, e = omd.nve_step(x, p, f, m, force, dt)
        drifts.append(abs((e + omd.kinetic_energy(p, m, idx, 1)).item() - e0) / n)
    assert drifts[0] < 1e-6
    assert 2.5 < drifts[0] / drifts[1] < 6.0


# ---------------------------------------------------------------- mat33
def test_frechet_adjoint_identity():
    """sum_ij L(A, E_mn)_ij G_ij == L(A^T, G)_mn : one Frechet derivative per
    system gives the cell forces the reference computes with nine."""
    rng = np.random.default_rng(0)
    a = torch.tensor(rng.normal(0, 0.05, (4, 3, 3)))
    g = torch.tensor(rng.normal(0, 1.0, (4, 3, 3)))
    nine = mat33.frechet_cell_forces(a, g)
    _, one = mat33.expm_frechet(a.transpose(1, 2).contiguous(), g)
    torch.testing.assert_close(nine, one, rtol=1e-12, atol=1e-14)


def test_expm_logm_roundtrip():
    rng = np.random.default_rng(1)
    a = torch.tensor(rng.normal(0, 0.1, (3, 3, 3)))
    torch.testing.assert_close(mat33.logm(mat33.expm(a)), a, rtol=1e-10, atol=1e-12)


def test_fire_first_step_uses_raw_forces():
    """ASE-FIRE: all-NaN velocities -> zero, no mixing, v = F dt (fire.py:308-315)."""
    pos = _t([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0]])
    cell = (torch.eye(3, dtype=F64) * 10).unsqueeze(0)
    idx = torch.zeros(2, dtype=torch.long)
    z = torch.ones(2, dtype=torch.long)
    fn = lambda d, zi, zj: om.soft_sphere_pair(d, zi, zj, 1.5, 1.0, 2.0)  # noqa: E731
    model = lambda x, c: om.pair_forward(x, c, torch.tensor([True] * 3), z, idx, fn, 1.5)  # noqa: E731
    st = ofire.fire_init(pos, cell, torch.ones(2, dtype=F64), idx, model)
    f0 = st["forces"].clone()
    st = ofire.ase_fire_step(st, model)
    torch.testing.assert_close(st["velocities"], f0 * 0.1)
