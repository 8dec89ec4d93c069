"""Closed-form pins of the oracle pieces that the GPU tests lean on and that
no reference golden vector covers directly: the species-pair lookup, the
per-atom energies and stresses, and the kinetic-energy / temperature
reductions (CPU only).

Every expected value here is computed from the mathematics of the case (a
pair formula evaluated by hand, an fcc shell sum, a constructed momentum
set), never by calling the oracle function under test, so that a plausible
mistake in the oracle -- a transposed species index, a dropped factor 0.5, a
wrong sign, a per-atom weight of 1 instead of 1/2 -- fails one of them.
"""

import math

import pytest
import torch

from oracle import md as omd, models as om

F64 = torch.float64


# ----------------------------------------------------- species-pair lookup
def _mss(sigma, epsilon=None, alpha=None, kind="soft_sphere", z=(18, 36)):
    n = len(z)
    eps = torch.ones(n, n, dtype=F64) if epsilon is None else torch.tensor(epsilon, dtype=F64)
    alp = None if alpha is None else torch.tensor(alpha, dtype=F64)
    return om.MultiPairFn(list(z), torch.tensor(sigma, dtype=F64), eps, alp, kind=kind)


def test_multi_soft_sphere_species_lookup_exact_values():
    """tests/models/test_soft_sphere.py:84-104 with the values written out:
    V = eps/alpha (1 - r/sigma)^alpha at r = 0.5 with sigma_AA = 1,
    sigma_AB = 2, sigma_BB = 3, eps = 1, alpha = 2."""
    fn = _mss([[1.0, 2.0], [2.0, 3.0]], alpha=[[2.0, 2.0], [2.0, 2.0]])
    dr = torch.full((4,), 0.5, dtype=F64)
    zi = torch.tensor([18, 18, 36, 36])
    zj = torch.tensor([18, 36, 18, 36])
    e = fn(dr, zi, zj)
    assert e.tolist() == [0.125, 0.28125, 0.28125, 0.5 * (1.0 - 0.5 / 3.0) ** 2]


def test_multi_soft_sphere_species_order_follows_atomic_numbers_list():
    """Rows follow the order of ``atomic_numbers`` (soft_sphere.py:132-140):
    listing Kr first makes sigma[0, 0] the Kr-Kr value."""
    fn = _mss([[3.0, 2.0], [2.0, 1.0]], z=(36, 18), alpha=[[2.0, 2.0], [2.0, 2.0]])
    dr = torch.full((2,), 0.5, dtype=F64)
    e = fn(dr, torch.tensor([18, 36]), torch.tensor([18, 36]))
    assert e.tolist() == [0.125, 0.5 * (1.0 - 0.5 / 3.0) ** 2]


def test_multi_soft_sphere_alpha_and_epsilon_matrices_are_used():
    """eps_AB = 3, alpha_AB = 3 at r = 0.5, sigma_AB = 1: 3/3 (1/2)^3 = 0.125;
    eps_AA = 2, alpha_AA = 4: 2/4 (1/2)^4 = 0.03125."""
    fn = _mss([[1.0, 1.0], [1.0, 1.0]], epsilon=[[2.0, 3.0], [3.0, 1.0]],
              alpha=[[4.0, 3.0], [3.0, 2.0]])
    dr = torch.full((2,), 0.5, dtype=F64)
    e = fn(dr, torch.tensor([18, 18]), torch.tensor([36, 18]))
    assert e.tolist() == [0.125, 0.03125]


def test_multi_lennard_jones_minimum_is_minus_epsilon_of_the_pair():
    """V_ab(2^(1/6) sigma_ab) = -eps_ab for each species pair, and zero force
    there (the LJ branch of the lookup, Kob-Andersen parameters)."""
    sig = [[1.0, 0.8], [0.8, 0.88]]
    eps = [[1.0, 1.5], [1.5, 0.5]]
    fn = _mss(sig, epsilon=eps, kind="lennard_jones", z=(1, 2))
    for a, b in [(1, 1), (1, 2), (2, 1), (2, 2)]:
        s, e = sig[a - 1][b - 1], eps[a - 1][b - 1]
        r = torch.tensor([2.0 ** (1.0 / 6.0) * s], dtype=F64, requires_grad=True)
        v = fn(r, torch.tensor([a]), torch.tensor([b]))
        assert abs(v.item() + e) < 1e-14
        (g,) = torch.autograd.grad(v.sum(), r)
        assert abs(g.item()) < 1e-12


def test_multi_species_dimers_energy_and_force_closed_form():
    """pair_forward on isolated dimers of each species pair (one system per
    dimer, non-periodic): E = V_ab(r) and F_j = -V_ab'(r) r_hat, written out
    from the LJ formula with the pair's own (sigma, epsilon)."""
    sig = [[1.0, 0.8], [0.8, 0.88]]
    eps = [[1.0, 1.5], [1.5, 0.5]]
    fn = _mss(sig, epsilon=eps, kind="lennard_jones", z=(7, 9))
    pairs = [(7, 7, 1.05), (7, 9, 0.93), (9, 7, 1.21), (9, 9, 0.97)]
    pos, zs, cells = [], [], []
    u = torch.tensor([0.6, 0.0, 0.8], dtype=F64)
    for k, (za, zb, r) in enumerate(pairs):
        base = torch.tensor([2.0, 3.0, 1.0], dtype=F64)
        pos += [base, base + r * u]
        zs += [za, zb]
        cells.append(torch.eye(3, dtype=F64) * 20.0)
    pos = torch.stack(pos)
    z = torch.tensor(zs)
    idx = torch.arange(len(pairs)).repeat_interleave(2)
    out = om.pair_forward(pos, torch.stack(cells), torch.tensor([False] * 3), z, idx, fn, 2.5,
                          per_atom=True)
    for k, (za, zb, r) in enumerate(pairs):
        s, e = sig[za // 9][zb // 9], eps[za // 9][zb // 9]
        v = 4 * e * ((s / r) ** 12 - (s / r) ** 6)
        dv = 4 * e * (-12 * s**12 / r**13 + 6 * s**6 / r**7)
        assert abs(out["energy"][k].item() - v) < 1e-13
        fj = -dv * u
        torch.testing.assert_close(out["forces"][2 * k + 1], fj, rtol=1e-12, atol=1e-13)
        torch.testing.assert_close(out["forces"][2 * k], -fj, rtol=1e-12, atol=1e-13)
        # each atom of the pair carries the full pair energy (reference quirk)
        assert abs(out["energies"][2 * k].item() - v) < 1e-13
        assert abs(out["energies"][2 * k + 1].item() - v) < 1e-13


# --------------------------------------------------- per-atom energy / stress
SHELLS = [(1, 12), (2, 6), (3, 24), (4, 12), (5, 24), (6, 8), (7, 48), (8, 6)]
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)


def _fcc(a, reps):
    basis = torch.tensor([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], dtype=F64)
    g = torch.stack(torch.meshgrid(*[torch.arange(r, dtype=F64) for r in reps], indexing="ij"),
                    -1).reshape(-1, 1, 3)
    return ((g + basis[None]).reshape(-1, 3)) * a


def _morse_v(r):
    s, e, a = MORSE["sigma"], MORSE["epsilon"], MORSE["alpha"]
    return e * (1 - math.exp(-a * (r - s))) ** 2 - e


def _morse_dv(r):
    s, e, a = MORSE["sigma"], MORSE["epsilon"], MORSE["alpha"]
    x = math.exp(-a * (r - s))
    return 2 * e * a * x * (1 - x)


@pytest.mark.parametrize("a", [3.61, 3.50])
def test_perfect_lattice_per_atom_energy_and_stress_closed_form(a):
    """Cu fcc 3x3x3 (108 atoms, rc = 6): every atom's energy is the full shell
    sum sum_s n_s V(r_s) (per-atom energies count each pair at both ends);
    the system stress is isotropic, N / (6 V) sum_s n_s r_s V'(r_s), and
    each atom carries 1/N of it.  a = 3.50 compresses the crystal so that the
    pressure term is large (a sign error would flip it)."""
    rc = 6.0
    pos = _fcc(a, (3, 3, 3))
    n = pos.shape[0]
    cell = (torch.eye(3, dtype=F64) * 3 * a).unsqueeze(0)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    out = om.pair_forward(pos, cell, torch.tensor([True] * 3), torch.ones(n, dtype=torch.long),
                          torch.zeros(n, dtype=torch.long), fn, rc, per_atom=True)
    shells = [(a * math.sqrt(k / 2), c) for k, c in SHELLS if a * math.sqrt(k / 2) < rc]
    e_atom = sum(c * _morse_v(r) for r, c in shells)
    assert (out["energies"] - e_atom).abs().max() < 1e-12
    assert abs(out["energies"].sum().item() - 2 * out["energy"].item()) < 1e-10
    vol = (3 * a) ** 3
    p = n / (6 * vol) * sum(c * r * _morse_dv(r) for r, c in shells)
    torch.testing.assert_close(out["stress"][0], p * torch.eye(3, dtype=F64), rtol=1e-11, atol=1e-15)
    expect_atom = (p / n) * torch.eye(3, dtype=F64)
    assert (out["stresses"] - expect_atom).abs().max() < 1e-14 + 1e-11 * abs(p / n)


def test_per_atom_outputs_sum_to_system_values_on_a_jittered_batch():
    """Two jittered triclinic systems: sum of per-atom energies = 2 E and
    sum of per-atom stresses = the system stress, per system (the weights
    1 and 1/2 of pair_potential.py:411-417 and 253-264)."""
    g = torch.Generator().manual_seed(4)
    pos1 = _fcc(3.61, (2, 2, 3)) + 0.08 * torch.randn(48, 3, generator=g, dtype=F64)
    pos2 = _fcc(3.61, (3, 2, 2)) + 0.08 * torch.randn(48, 3, generator=g, dtype=F64)
    c1 = torch.tensor([[7.22, 0.3, 0.0], [0.0, 7.22, 0.2], [0.1, 0.0, 10.83]], dtype=F64)
    c2 = torch.tensor([[10.83, 0.0, -0.2], [0.4, 7.22, 0.0], [0.0, 0.0, 7.22]], dtype=F64)
    pos = torch.cat([pos1, pos2])
    idx = torch.arange(2).repeat_interleave(48)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    out = om.pair_forward(pos, torch.stack([c1, c2]), torch.tensor([True] * 3),
                          torch.ones(96, dtype=torch.long), idx, fn, 6.0, per_atom=True)
    for s in range(2):
        sel = idx == s
        assert abs(out["energies"][sel].sum().item() - 2 * out["energy"][s].item()) < 1e-11
        torch.testing.assert_close(out["stresses"][sel].sum(0), out["stress"][s], rtol=1e-11,
                                   atol=1e-14)


# ------------------------------------------------------------ calc_kT / KE
def test_kinetic_energy_and_kT_closed_form():
    """Momenta p_i = sqrt(m_i) u_i with chosen u: per-system
    KE = 0.5 sum |u_i|^2 and kT = sum |u_i|^2 / (3 n) (quantities.py:16-66,
    104-139, dof = 3 n)."""
    u = torch.tensor([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0],
                      [1.0, 1.0, 1.0], [2.0, -1.0, 2.0]], dtype=F64)
    m = torch.tensor([1.0, 4.0, 9.0, 2.0, 0.5], dtype=F64)
    idx = torch.tensor([0, 0, 0, 1, 1])
    p = torch.sqrt(m).unsqueeze(-1) * u
    ke = omd.kinetic_energy(p, m, idx, 2)
    torch.testing.assert_close(ke, torch.tensor([0.5 * 14.0, 0.5 * 12.0], dtype=F64),
                               rtol=1e-15, atol=0)
    kt = omd.calc_kT(p, m, idx, 2)
    torch.testing.assert_close(kt, torch.tensor([14.0 / 9.0, 12.0 / 6.0], dtype=F64),
                               rtol=1e-15, atol=0)


def test_initialized_momenta_have_the_requested_temperature():
    """Maxwell-Boltzmann momenta at kT with the centre-of-mass momentum
    removed: E[kT_measured] = kT (n - 1) / n, checked on 20,000 atoms of
    unequal mass at 0.3 % (about 10 standard errors)."""
    g = torch.Generator().manual_seed(7)
    n = 20000
    m = 1.0 + 60.0 * torch.rand(n, generator=g, dtype=F64)
    idx = torch.zeros(n, dtype=torch.long)
    kT = torch.tensor(0.0862, dtype=F64)
    p = omd.initialize_momenta(m, idx, kT, torch.randn(n, 3, generator=g, dtype=F64))
    kt = omd.calc_kT(p, m, idx, 1).item()
    assert abs(kt / (0.0862 * (n - 1) / n) - 1.0) < 3e-3


# ----------------------------------------------------------- NPT (GJF)
def test_gjf_particle_update_samples_the_harmonic_distribution_at_large_dt():
    """Gronbech-Jensen and Farago (Mol. Phys. 111, 983 (2013)): their
    Langevin integrator gives the exact configurational distribution of a
    harmonic well at any stable time step, <k x^2> = kT.  20,000
    oscillators (omega = 1) at dt = 0.5 with friction 0.5: a wrong damping
    factor a or b moves <k x^2> by tens of percent."""
    from oracle import md as omd

    g = torch.Generator().manual_seed(11)
    n, k, kT, dt, alpha = 20000, 1.0, 1.0, 0.5, 0.5
    m = torch.ones(n, dtype=F64)
    a_t = torch.full((n,), alpha, dtype=F64)
    x = torch.zeros(n, 3, dtype=F64)
    v = torch.zeros(n, 3, dtype=F64)
    f = -k * x
    acc, n_acc = 0.0, 0
    for step in range(600):
        beta = math.sqrt(2 * alpha * kT * dt) * torch.randn(n, 3, generator=g, dtype=F64)
        x = omd.gjf_particle_position(x, v, f, m, dt, a_t, beta, torch.ones(n, dtype=F64))
        f_new = -k * x
        v = omd.gjf_particle_velocity(v, f, f_new, m, dt, a_t, beta)
        f = f_new
        if step >= 300:
            acc += (k * x * x).mean().item()
            n_acc += 1
    assert abs(acc / n_acc / kT - 1.0) < 0.03, acc / n_acc


def test_gjf_strain_update_samples_the_harmonic_distribution():
    """The strain coordinate's GJF update in a harmonic well F = -K eps:
    <K eps^2> = kT for 20,000 independent cells."""
    from oracle import md as omd

    g = torch.Generator().manual_seed(12)
    n, big_k, kT, dt = 20000, 4.0, 0.5, 0.3
    q = torch.full((n,), 2.0, dtype=F64)
    ca = torch.full((n,), 1.5, dtype=F64)
    eps = torch.zeros(n, dtype=F64)
    eps_dot = torch.zeros(n, dtype=F64)
    fo = -big_k * eps
    acc, n_acc = 0.0, 0
    for step in range(800):
        beta = torch.sqrt(2.0 * ca * kT * dt) * torch.randn(n, generator=g, dtype=F64)
        eps = omd.gjf_strain_position(eps, eps_dot, fo, q, dt, ca, beta)
        fn = -big_k * eps
        eps_dot = omd.gjf_strain_velocity(eps_dot, fo, fn, q, dt, ca, beta)
        fo = fn
        if step >= 400:
            acc += (big_k * eps * eps).mean().item()
            n_acc += 1
    assert abs(acc / n_acc / kT - 1.0) < 0.03, acc / n_acc


def test_npt_step_without_forces_or_noise_scales_isotropically():
    """Zero forces, momenta and noise, no friction: the strain moves by the
    closed form b dt^2 F_eps / (2Q) with F_eps = 3 V (P - P_ext), the cell
    becomes exp(eps') h0 and the positions keep their fractional
    coordinates."""
    from oracle import md as omd

    g = torch.Generator().manual_seed(13)
    n_sys = 2
    idx = torch.tensor([0] * 5 + [1] * 7)
    ref = torch.stack([torch.eye(3, dtype=F64) * 9.0,
                       torch.tensor([[8.0, 0.5, 0.0], [0.0, 9.0, 0.3], [0.2, 0.0, 10.0]],
                                    dtype=F64)])
    frac = torch.rand(12, 3, generator=g, dtype=F64)
    x = torch.einsum("nij,nj->ni", ref[idx], frac)
    stress = torch.stack([torch.eye(3, dtype=F64) * -0.01, torch.eye(3, dtype=F64) * 0.02])
    zeros = torch.zeros(12, 3, dtype=F64)
    m = torch.full((12,), 2.0, dtype=F64)
    kT = torch.tensor([0.05, 0.08], dtype=F64)
    p_ext = torch.tensor([0.001, -0.002], dtype=F64)
    b_tau = torch.tensor([3.0, 4.0], dtype=F64)
    zero_s = torch.zeros(n_sys, dtype=F64)
    force_fn = lambda xx, cc: (zero_s, torch.zeros_like(xx), stress)  # noqa: E731
    dt = 0.2
    out = omd.npt_langevin_isotropic_step(
        x, zeros, zeros, m, idx, ref, zero_s, zero_s, stress, force_fn, dt, kT, p_ext,
        torch.zeros(n_sys, dtype=F64), torch.zeros(n_sys, dtype=F64), b_tau,
        torch.zeros(n_sys, dtype=F64), zeros)
    x1, p1, _, _, _, cell, eps, eps_dot = out
    n_at = torch.tensor([5.0, 7.0], dtype=F64)
    vol = torch.linalg.det(ref)
    f_eps = 3 * vol * (-torch.diagonal(stress, dim1=1, dim2=2).sum(1) / 3 + n_at * kT / vol - p_ext)
    q = (n_at + 1) * kT * b_tau ** 2
    torch.testing.assert_close(eps, dt * dt * f_eps / (2 * q), rtol=1e-14, atol=0)
    torch.testing.assert_close(cell, torch.exp(eps)[:, None, None] * ref, rtol=1e-15, atol=0)
    frac1 = torch.linalg.solve(cell[idx], x1.unsqueeze(-1)).squeeze(-1)
    torch.testing.assert_close(frac1, frac, rtol=0, atol=1e-14)
    assert p1.abs().max() == 0.0
    # the strain velocity averages the force at the old and the new volume
    vol1 = vol * torch.exp(3 * eps)
    f_eps1 = 3 * vol1 * (-torch.diagonal(stress, dim1=1, dim2=2).sum(1) / 3 + n_at * kT / vol1
                         - p_ext)
    torch.testing.assert_close(eps_dot, dt * (f_eps + f_eps1) / (2 * q), rtol=1e-13, atol=0)


def test_strain_force_of_a_perfect_lattice_closed_form():
    """F_eps = 3 V (P - P_ext) with the virial pressure of a compressed fcc
    Cu crystal from its shell sum, P_vir = -N / (6V) sum_s n_s r_s V'(r_s),
    plus the kinetic term N kT / V."""
    from oracle import md as omd

    a, rc = 3.50, 6.0
    pos = _fcc(a, (3, 3, 3))
    n = pos.shape[0]
    cell = (torch.eye(3, dtype=F64) * 3 * a).unsqueeze(0)
    fn = lambda d, zi, zj: om.morse_pair(d, zi, zj, **MORSE)  # noqa: E731
    out = om.pair_forward(pos, cell, torch.tensor([True] * 3), torch.ones(n, dtype=torch.long),
                          torch.zeros(n, dtype=torch.long), fn, rc)
    vol = (3 * a) ** 3
    shells = [(a * math.sqrt(k / 2), c) for k, c in SHELLS if a * math.sqrt(k / 2) < rc]
    p_vir = -n / (6 * vol) * sum(c * r * _morse_dv(r) for r, c in shells)
    kT, p_ext = 0.0862, 0.004
    f = omd.strain_force(out["stress"], torch.tensor([vol], dtype=F64),
                         torch.tensor([float(n)], dtype=F64), torch.tensor([kT], dtype=F64),
                         torch.tensor([p_ext], dtype=F64))
    expect = 3 * vol * (p_vir + n * kT / vol - p_ext)
    assert abs(f.item() - expect) < 1e-10 * abs(expect)
    assert p_vir > 0  # compressed: positive pressure


def test_particle_life_force_known_answers():
    """test_particle_life.py:13-35 and the closed form of the force."""
    z = torch.ones(6, dtype=torch.long)
    r = torch.tensor([0.1, 0.2, 0.3, 0.65, 1.0, 1.5], dtype=F64)
    f = om.particle_life_force(r, z, z, A=2.0, beta=0.3, sigma=1.0)
    assert (f[:2] < 0).all() and (f[4:] == 0).all()
    torch.testing.assert_close(f[:2], r[:2] / 0.3 - 1.0, rtol=0, atol=1e-15)
    assert abs(f[2].item()) < 1e-15 and abs(f[3].item() - 2.0) < 1e-15
    f1 = om.particle_life_force(r, z, z, A=1.0, beta=0.3, sigma=1.0)
    torch.testing.assert_close(om.particle_life_force(r, z, z, A=3.0, beta=0.3, sigma=1.0)[3:],
                               3.0 * f1[3:])


def test_pair_forces_forward_dimer_closed_form():
    """Two atoms 0.5 apart along (1, 2, 2)/3 in a non-periodic box: f = A (1 -
    0.3 / 0.7); a full list applies every entry's force and reaction (twice
    per pair), a half list once; the stress is -(f r / V) r_hat r_hat."""
    d = 0.5 * torch.tensor([1.0, 2.0, 2.0], dtype=F64) / 3.0
    x = torch.stack([torch.tensor([1.0, 1.0, 1.0], dtype=F64), torch.tensor([1.0, 1.0, 1.0],
                                                                            dtype=F64) + d])
    cell = 10.0 * torch.eye(3, dtype=F64).unsqueeze(0)
    pbc = torch.zeros(3, dtype=torch.bool)
    z = torch.ones(2, dtype=torch.long)
    idx = torch.zeros(2, dtype=torch.long)
    fn = lambda r, zi, zj: om.particle_life_force(r, zi, zj, A=1.5, beta=0.3, sigma=1.0)  # noqa: E731
    fmag = 1.5 * (1.0 - 0.3 / 0.7)
    rhat = d / 0.5
    for half, k in ((False, 2.0), (True, 1.0)):
        out = om.pair_forces_forward(x, cell, pbc, z, idx, fn, 1.0, half=half, per_atom=True)
        torch.testing.assert_close(out["forces"][0], -k * fmag * rhat, rtol=0, atol=1e-14)
        torch.testing.assert_close(out["forces"][1], k * fmag * rhat, rtol=0, atol=1e-14)
        sig = -(fmag * 0.5 / 1000.0) * torch.outer(rhat, rhat)
        torch.testing.assert_close(out["stress"][0], sig, rtol=0, atol=1e-16)
        torch.testing.assert_close(out["stresses"].sum(0), sig, rtol=0, atol=1e-16)
        assert out["energy"].tolist() == [0.0]


def test_pair_forces_forward_with_minus_dv_dr_equals_pair_forward():
    """force_fn = -dV/dr (Morse): on a half list the force model's forces,
    stress and per-atom stresses equal the potential model's; the full list
    doubles the forces (every pair applied twice) and keeps the stresses."""
    g = torch.Generator().manual_seed(41)
    pos = torch.cat([_fcc(3.61, (2, 2, 3)), _fcc(3.61, (3, 2, 2))])
    pos = pos + 0.08 * torch.randn(96, 3, generator=g, dtype=F64)
    cell = torch.stack([
        torch.tensor([[7.22, 0.3, 0.0], [0.0, 7.22, 0.2], [0.1, 0.0, 10.83]], dtype=F64),
        torch.tensor([[10.83, 0.0, -0.2], [0.4, 7.22, 0.0], [0.0, 0.0, 7.22]], dtype=F64)])
    z, idx = torch.ones(96, dtype=torch.long), torch.arange(2).repeat_interleave(48)
    pbc = torch.tensor([True] * 3)
    prm = MORSE

    def force_fn(r, zi, zj):
        rg = r.detach().requires_grad_()
        (dv,) = torch.autograd.grad(om.morse_pair(rg, zi, zj, **prm).sum(), rg)
        return -dv

    pot = om.pair_forward(pos, cell, pbc, z, idx,
                          lambda r, zi, zj: om.morse_pair(r, zi, zj, **prm), 6.0, per_atom=True)
    args = (pos, cell, pbc, z, idx, force_fn, 6.0)
    hf = om.pair_forces_forward(*args, half=True, per_atom=True)
    ff = om.pair_forces_forward(*args, half=False, per_atom=True)
    scale = pot["forces"].abs().max().item()
    assert scale > 0.1
    torch.testing.assert_close(hf["forces"], pot["forces"], rtol=0, atol=1e-12 * scale)
    torch.testing.assert_close(ff["forces"], 2.0 * pot["forces"], rtol=0, atol=2e-12 * scale)
    for out in (hf, ff):
        torch.testing.assert_close(out["stress"], pot["stress"], rtol=1e-12, atol=1e-15)
        torch.testing.assert_close(out["stresses"], pot["stresses"], rtol=1e-12, atol=1e-15)
