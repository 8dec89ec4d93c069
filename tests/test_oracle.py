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
from oracle import mat33, md as omd, models as om, nl as onl, rng as orng

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


# ---------------------------------------------------------------- FIRE closed forms
def _const_force_model(force, stress=None):
    """E = -sum F.x: constant forces, optional constant stress [S,3,3]."""
    force = _t(force)

    def fn(x, cell):
        out = {"forces": force.clone(), "energy": -(force * x).sum().view(1)}
        out["stress"] = torch.zeros(cell.shape[0], 3, 3, dtype=F64) if stress is None else stress.clone()
        return out

    return fn


def _fire_state(force, v, n_pos, dt=0.1, alpha=0.1, flavor_vv=False):
    pos = torch.zeros(len(force), 3, dtype=F64)
    cell = (torch.eye(3, dtype=F64) * 10).unsqueeze(0)
    idx = torch.zeros(len(force), dtype=torch.long)
    model = _const_force_model(force)
    st = ofire.fire_init(pos, cell, torch.ones(len(force), dtype=F64), idx, model, dt_start=dt,
                         alpha_start=alpha)
    st["velocities"] = _t(v)
    st["n_pos"][:] = n_pos
    return st, model


def test_ase_fire_negative_power_branch():
    """P < 0 (test_optimizers.py:600-662): dt *= f_dec, alpha = alpha_start,
    n_pos = 0, velocities zeroed then v = dt_new F."""
    f = [[0.3, -0.2, 0.1], [-0.1, 0.4, 0.2]]
    st, model = _fire_state(f, (-0.1 * _t(f)).tolist(), n_pos=4, alpha=0.07)
    st = ofire.ase_fire_step(st, model, max_step=10.0, alpha_start=0.1)
    assert st["dt"].item() == pytest.approx(0.05, abs=1e-15)
    assert st["alpha"].item() == pytest.approx(0.1, abs=1e-15)
    assert st["n_pos"].item() == 0
    torch.testing.assert_close(st["velocities"], 0.05 * _t(f), rtol=0, atol=1e-15)
    torch.testing.assert_close(st["positions"], 0.05 * 0.05 * _t(f), rtol=0, atol=1e-15)


def test_ase_fire_mixing_known_answer():
    """One atom, v = (1,0,0), F = (2,2,0): P = 2 > 0, n_pos <= n_min so dt and
    alpha stay; |v| F/|F| = (1,1,0)/sqrt(2); v' = 0.9 v + 0.1 (1,1,0)/sqrt(2);
    then v += dt F and x += dt v."""
    st, model = _fire_state([[2.0, 2.0, 0.0]], [[1.0, 0.0, 0.0]], n_pos=2)
    st = ofire.ase_fire_step(st, model, max_step=10.0)
    r = 1.0 / math.sqrt(2.0)
    v = _t([[0.9 + 0.1 * r + 0.2, 0.1 * r + 0.2, 0.0]])
    assert st["dt"].item() == 0.1 and st["alpha"].item() == 0.1 and st["n_pos"].item() == 3
    torch.testing.assert_close(st["velocities"], v, rtol=0, atol=1e-15)
    torch.testing.assert_close(st["positions"], 0.1 * v, rtol=0, atol=1e-15)
    # two atoms: the norms are per system (|v| = 1, |F| = sqrt 2), not per atom
    st, model = _fire_state([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]], n_pos=0)
    st = ofire.ase_fire_step(st, model, max_step=10.0)
    v = _t([[0.9 + 0.1 * r + 0.1, 0.0, 0.0], [0.0, 0.1 * r + 0.1, 0.0]])
    torch.testing.assert_close(st["velocities"], v, rtol=0, atol=1e-15)


def test_fire_mixing_leaves_parallel_velocity_unchanged():
    """v = c F with c > 0 across a system: mixing with the system-wide norms
    is the identity for any alpha (a per-atom norm would change it)."""
    f = _t([[0.5, 0.0, 0.0], [0.0, 3.0, 0.0], [0.0, 0.0, 0.25]])
    st, model = _fire_state(f.tolist(), (0.7 * f).tolist(), n_pos=9, dt=0.2, alpha=0.3)
    st = ofire.ase_fire_step(st, model, max_step=100.0)
    dt = 0.2 * 1.1
    assert st["dt"].item() == pytest.approx(dt, rel=1e-15)
    assert st["alpha"].item() == pytest.approx(0.3 * 0.99, rel=1e-15)
    torch.testing.assert_close(st["velocities"], 0.7 * f + dt * f, rtol=1e-14, atol=0)


@pytest.mark.parametrize("flavor", ["ase", "vv"])
def test_fire_n_pos_ordering_differs_between_flavors(flavor):
    """n_pos == n_min with P > 0: ASE tests n_pos > n_min before the increment
    (fire.py:342-345, no dt growth); VV increments first (fire.py:244-246,
    dt grows by f_inc)."""
    f = [[1.0, 0.0, 0.0]]
    st, model = _fire_state(f, [[1.0, 0.0, 0.0]], n_pos=5)
    if flavor == "ase":
        st = ofire.ase_fire_step(st, model)
        assert st["dt"].item() == 0.1
    else:
        st = ofire.vv_fire_step(st, model)
        assert st["dt"].item() == pytest.approx(0.11, rel=1e-15)
    assert st["n_pos"].item() == 6


def test_fire_dt_max_and_max_step_clamp():
    """dt growth is capped at dt_max; a step longer than max_step is scaled to
    exactly max_step along its own direction (per-system norm)."""
    f = [[30.0, 40.0, 0.0], [0.0, 0.0, 0.0]]
    st, model = _fire_state(f, [[3.0, 4.0, 0.0], [0.0, 0.0, 0.0]], n_pos=10, dt=0.95)
    st = ofire.ase_fire_step(st, model, dt_max=1.0, max_step=0.2)
    assert st["dt"].item() == 1.0
    d = st["positions"]
    assert d[0].norm().item() == pytest.approx(0.2, rel=1e-14)
    torch.testing.assert_close(d[0] / d[0].norm(), _t([0.6, 0.8, 0.0]), rtol=0, atol=1e-15)
    assert d[1].abs().max().item() == 0.0


def test_vv_fire_negative_power_zeroes_velocities():
    """VV flavour, P < 0 after the second half kick: velocities are zeroed,
    dt *= f_dec, alpha = alpha_start (fire.py:240-262)."""
    f = [[1.0, 0.0, 0.0]]
    st, model = _fire_state(f, [[-5.0, 0.0, 0.0]], n_pos=3, alpha=0.05)
    st = ofire.vv_fire_step(st, model)
    assert st["dt"].item() == 0.05 and st["alpha"].item() == 0.1 and st["n_pos"].item() == 0
    assert st["velocities"].abs().max().item() == 0.0
    # x = dt (v + dt F / 2) with the old dt
    torch.testing.assert_close(st["positions"], _t([[0.1 * (-5.0 + 0.05), 0.0, 0.0]]), rtol=0, atol=1e-15)


def test_cell_filter_init_closed_forms():
    """Isotropic stress p on a cubic cell of N atoms: virial = -V p I.
    Frechet: F = I, log F = 0, so cell forces = virial / N; unit filter:
    cell positions = N I, cell forces = virial / N.  Hydrostatic projection
    keeps the mean diagonal, constant volume removes it."""
    n, a, p = 3, 4.0, 0.02
    stress = (p * torch.eye(3, dtype=F64)).unsqueeze(0)
    pos = torch.zeros(n, 3, dtype=F64)
    cell = (torch.eye(3, dtype=F64) * a).unsqueeze(0)
    idx = torch.zeros(n, dtype=torch.long)
    model = _const_force_model(torch.zeros(n, 3).tolist(), stress)
    want = -(a ** 3) * p * torch.eye(3, dtype=F64) / n
    for kw in ({"frechet": True}, {"unit": True}):
        st = ofire.fire_init(pos, cell, torch.ones(n, dtype=F64), idx, model, **kw)
        torch.testing.assert_close(st["cell_forces"][0], want, rtol=1e-15, atol=0)
    assert torch.equal(st["cell_positions"][0], n * torch.eye(3, dtype=F64))
    v = _t([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]]])
    h = ofire._apply_constraints(v, {"hydrostatic_strain": True})
    torch.testing.assert_close(h[0], 5.0 * torch.eye(3, dtype=F64))
    c = ofire._apply_constraints(v, {"constant_volume": True})
    assert torch.diagonal(c[0]).sum().item() == pytest.approx(0.0, abs=1e-14)
    torch.testing.assert_close(c[0] - v[0], -5.0 * torch.eye(3, dtype=F64))


def test_frechet_first_step_scales_cell_by_scalar_exp():
    """Zero atomic forces, isotropic stress: the first ASE step moves the cell
    log-strain by dt^2 * c with c the (isotropic) cell force, so the cell and
    every position scale by exp(dt^2 c / N)."""
    n, a, p, dt = 4, 5.0, -0.01, 0.1
    stress = (p * torch.eye(3, dtype=F64)).unsqueeze(0)
    pos = _t([[0.5, 1.0, 1.5], [2.0, 0.2, 3.0], [4.0, 4.5, 0.1], [1.0, 3.0, 2.0]])
    cell = (torch.eye(3, dtype=F64) * a).unsqueeze(0)
    idx = torch.zeros(n, dtype=torch.long)
    model = _const_force_model(torch.zeros(n, 3).tolist(), stress)
    st = ofire.fire_init(pos, cell, torch.ones(n, dtype=F64), idx, model, frechet=True, dt_start=dt)
    c = -(a ** 3) * p / n  # isotropic cell force
    st = ofire.ase_fire_step(st, model, max_step=10.0)
    s = math.exp(dt * dt * c / n)
    torch.testing.assert_close(st["cell"][0], a * s * torch.eye(3, dtype=F64), rtol=1e-14, atol=0)
    torch.testing.assert_close(st["positions"], pos * s, rtol=1e-14, atol=1e-15)
    torch.testing.assert_close(st["cell_positions"][0], dt * dt * c * torch.eye(3, dtype=F64),
                               rtol=1e-12, atol=1e-16)


# ---------------------------------------------------------------- gradients
KA_SIG = torch.tensor([[1.0, 0.8], [0.8, 0.88]], dtype=F64)
KA_EPS = torch.tensor([[1.0, 1.5], [1.5, 0.5]], dtype=F64)


def _binary_lj(seed=0, reps=3, rho=0.8):
    a = (4.0 / rho) ** (1.0 / 3.0)
    g = torch.Generator().manual_seed(seed)
    pos = fcc(a, [reps] * 3)
    pos = pos + 0.05 * torch.randn(pos.shape, generator=g, dtype=F64)
    n = pos.shape[0]
    cell = (torch.eye(3, dtype=F64) * a * reps).unsqueeze(0)
    return pos, cell, torch.tensor([1, 2] * (n // 2)), torch.zeros(n, dtype=torch.long)


def _ka_forward(pos, cell, z, idx, s, e, create_graph):
    fn = om.MultiPairFn([1, 2], s, e, kind="lennard_jones")
    return om.pair_forward(pos, cell, torch.tensor([True] * 3), z, idx, fn, 2.5,
                           compute_stress=False, create_graph=create_graph, nl=onl.linked_cell_nl)


def _fd_check(loss_fn, gs, ge, h=1e-6, rel=1e-6):
    """Central differences of a scalar loss(sigma, eps) against autograd; an
    off-diagonal parameter is shared, so it moves both entries."""
    for a, b in [(0, 0), (0, 1), (1, 1)]:
        dp = torch.zeros(2, 2, dtype=F64)
        dp[a, b] = dp[b, a] = h
        for g, fd in ((gs, (loss_fn(KA_SIG + dp, KA_EPS) - loss_fn(KA_SIG - dp, KA_EPS)) / (2 * h)),
                      (ge, (loss_fn(KA_SIG, KA_EPS + dp) - loss_fn(KA_SIG, KA_EPS - dp)) / (2 * h))):
            got = (g[a, b] + g[b, a]).item() if a != b else g[a, b].item()
            assert abs(got - fd) <= rel * max(abs(fd), 1e-3), (a, b, got, fd)


def test_force_matching_autograd_matches_finite_differences():
    """The gradient oracle (autograd with create_graph through the forces)
    against FP64 central differences of the force-matching loss."""
    pos, cell, z, idx = _binary_lj()
    target = _ka_forward(pos, cell, z, idx, KA_SIG * 1.03, KA_EPS * 0.95, False)["forces"]

    def loss(s, e, cg=False):
        x = pos.clone().requires_grad_(cg)  # the force graph runs through the positions
        return ((_ka_forward(x, cell, z, idx, s, e, cg)["forces"] - target) ** 2).mean()

    s, e = KA_SIG.clone().requires_grad_(True), KA_EPS.clone().requires_grad_(True)
    gs, ge = torch.autograd.grad(loss(s, e, True), [s, e])
    _fd_check(lambda a, b: loss(a, b).item(), gs, ge)


def test_nve_adjoint_oracle_matches_finite_differences():
    """Autograd through 8 velocity-Verlet steps of the oracle: dE(x_T)/dtheta
    and dE(x_T)/dx0 against central differences."""
    pos, cell, z, idx = _binary_lj(seed=1, reps=2, rho=0.5)
    m = torch.ones(pos.shape[0], dtype=F64)
    g = torch.Generator().manual_seed(3)
    p0 = 0.5 * torch.randn(pos.shape, generator=g, dtype=F64)

    def final_energy(x0, s, e, cg):
        def force_fn(x):
            out = _ka_forward(x, cell, z, idx, s, e, cg)
            return out["energy"], out["forces"]

        x, p = x0, p0
        _, f = force_fn(x)
        for _ in range(8):
            x, p, f, _ = omd.nve_step(x, p, f, m, force_fn, 0.004)
        return force_fn(x)[0].sum()

    x0 = pos.clone().requires_grad_(True)
    s, e = KA_SIG.clone().requires_grad_(True), KA_EPS.clone().requires_grad_(True)
    gx, gs, ge = torch.autograd.grad(final_energy(x0, s, e, True), [x0, s, e])
    _fd_check(lambda a, b: final_energy(pos, a, b, False).item(), gs, ge)
    h = 1e-6
    for (i, c) in [(0, 0), (5, 2), (17, 1)]:
        d = torch.zeros_like(pos)
        d[i, c] = h
        fd = (final_energy(pos + d, KA_SIG, KA_EPS, False) - final_energy(pos - d, KA_SIG, KA_EPS, False)).item() / (2 * h)
        assert abs(gx[i, c].item() - fd) <= 1e-6 * max(abs(fd), 1e-3), (i, c, gx[i, c].item(), fd)


# ------------------------------------------------------------- Langevin noise
@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, out):
    """Random123's Philox4x32-10 known-answer vectors."""
    got = orng.philox4x32_10(ctr, key)
    assert tuple(int(x) for x in got) == out


def test_langevin_normals_box_muller_and_moments():
    """z0, z1 lie on the circle of radius sqrt(-2 log u1) at angle 2 pi u2 of
    block 0; z2 is block 1's cosine normal; about 4e5 draws have the first
    four moments of N(0, 1) and differ between steps, seeds and components."""
    n = 1 << 16
    z = orng.langevin_normals(seed=12345, step=7, n_atoms=n)
    i = np.arange(n, dtype=np.uint64)
    r0 = orng.philox4x32_10((i, 7, 0, 0), (12345, 0))
    r1 = orng.philox4x32_10((i, 7, 0, 1), (12345, 0))
    rad = np.sqrt(-2.0 * np.log(orng.u01_53(r0[0], r0[1])))
    np.testing.assert_allclose(np.hypot(z[:, 0], z[:, 1]), rad, rtol=1e-13)
    ang = np.arctan2(z[:, 1], z[:, 0]) % (2 * np.pi)
    d = np.abs(ang - 2 * np.pi * orng.u01_53(r0[2], r0[3]))
    assert np.minimum(d, 2 * np.pi - d).max() < 1e-9
    rad1 = np.sqrt(-2.0 * np.log(orng.u01_53(r1[0], r1[1])))
    np.testing.assert_allclose(np.abs(z[:, 2]), rad1 * np.abs(np.cos(2 * np.pi * orng.u01_53(r1[2], r1[3]))),
                               rtol=1e-12, atol=1e-15)
    allz = np.concatenate([z, orng.langevin_normals(seed=12345, step=8, n_atoms=n)]).ravel()
    assert abs(allz.mean()) < 0.01 and abs(allz.var() - 1.0) < 0.01
    assert abs((allz ** 3).mean()) < 0.03 and abs((allz ** 4).mean() - 3.0) < 0.05
    other = orng.langevin_normals(seed=12346, step=7, n_atoms=n)
    assert abs(np.corrcoef(z[:, 0], other[:, 0])[0, 1]) < 0.02
    assert abs(np.corrcoef(z[:, 0], z[:, 2])[0, 1]) < 0.02
    # the step counter's high word reaches the generator
    hi = orng.langevin_normals(seed=12345, step=7 + (1 << 32), n_atoms=8)
    assert not np.array_equal(hi, z[:8])
