"""Pins of the oracle's thermostats and barostats (oracle/thermo.py), on the
CPU.  Each check is one a plausible mistake in the restatement (a dropped
factor, a wrong sign, a transposed map, a mis-typed weight) would fail:

* Suzuki-Yoshida weights: each set sums to 1, the composite is symmetric,
  and the weights are held at float32 as the reference's tensors hold them;
* Nose-Hoover chain (NVT): the extended energy of harmonic oscillators is
  conserved to second order in dt, and a cold start heats towards kT;
* CSVR: on an ideal gas the kinetic energy samples the canonical gamma
  distribution (mean dof kT / 2, variance dof kT^2 / 2);
* anisotropic NPT Langevin: the per-axis strain forces sum to the isotropic
  strain force; with balanced pressure, no noise and no friction the strains
  move ballistically and positions and cell rows scale per Cartesian axis;
* MTK Nose-Hoover NPT: the extended Hamiltonian of a soft-sphere crystal is
  conserved over 300 steps;
* C-rescale: at balanced pressure with no noise every mode leaves the cell
  unchanged and drifts positions by p dt / m; the triclinic deformation
  preserves volume (exp of a traceless matrix, Gram-Schmidt rotation), so
  det(cell') = det(cell) (sqrt(V') / sqrt(V))^2; the Gram-Schmidt rows keep
  the columns' lengths and angles.
"""

import math

import torch

from oracle import md as omd
from oracle import models as om
from oracle import thermo as oth

F64 = torch.float64


def test_suzuki_yoshida_weights():
    for k, w in oth.SY_WEIGHTS.items():
        assert w.dtype == torch.float32 and w.numel() == k
        assert abs(float(w.double().sum()) - 1.0) < 1e-6
        assert torch.equal(w, w.flip(0))


def _harmonic(k_spring):
    def force_fn(x, cell):
        return 0.5 * k_spring * (x * x).sum().reshape(1), -k_spring * x, torch.zeros(1, 3, 3)
    return force_fn


def test_nose_hoover_chain_conserves_the_extended_energy_of_oscillators():
    torch.manual_seed(0)
    n = 24
    k_spring = 2.0
    m = torch.linspace(1.0, 3.0, n, dtype=F64)
    x = torch.randn(n, 3, dtype=F64) * 0.3
    p = torch.randn(n, 3, dtype=F64) * 0.1  # cold: KE well below the target
    si = torch.zeros(n, dtype=torch.long)
    kT = torch.tensor([1.0], dtype=F64)
    ff = _harmonic(k_spring)
    dof = torch.tensor([3.0 * n - 3])
    ke = oth.kinetic_energy(p, m, si, 1)
    chain = oth.nhc_init(dof, ke, kT, torch.tensor([0.5], dtype=F64), 3)
    e, f, _ = ff(x, None)
    dt = 0.01
    h0 = oth.nvt_nose_hoover_invariant(e, p, m, si, chain, kT)
    kes = []
    for step in range(3000):
        x, p, f, e, chain = oth.nvt_nose_hoover_step(x, p, f, m, si, ff, None, dt, kT, chain, 3, 3)
        if step >= 1000:
            kes.append(float(oth.kinetic_energy(p, m, si, 1)))
    h1 = oth.nvt_nose_hoover_invariant(e, p, m, si, chain, kT)
    assert abs(float(h1 - h0)) < 1e-4 * float(dof * kT)
    # heated from cold towards dof kT / 2
    mean_ke = sum(kes) / len(kes)
    assert 0.8 < mean_ke / float(dof * kT / 2) < 1.2, mean_ke


def test_nose_hoover_chain_lengths_and_substeps_agree_at_small_dt():
    """the chain integration converges: one substep and 3 x 3 SY substeps
    give the same half step to O(dt^3)"""
    torch.manual_seed(1)
    n = 10
    m = torch.ones(n, dtype=F64)
    p = torch.randn(n, 3, dtype=F64)
    si = torch.zeros(n, dtype=torch.long)
    kT = torch.tensor([0.3], dtype=F64)
    ke = oth.kinetic_energy(p, m, si, 1)
    chain = oth.nhc_init(torch.tensor([27.0]), ke, kT, torch.tensor([0.2], dtype=F64), 4)
    chain = dict(chain, pxi=torch.tensor([[0.5, -0.2, 0.1, 0.05]], dtype=F64))
    for dt in (1e-2, 5e-3):
        p1, c1 = oth.nhc_half_step(p, chain, kT, si, dt, 1, 1)
        p9, c9 = oth.nhc_half_step(p, chain, kT, si, dt, 3, 3)
        err = float((p1 - p9).abs().max())
        assert err < 50 * dt ** 3, (dt, err)


def test_csvr_samples_the_canonical_kinetic_energy():
    g = torch.Generator().manual_seed(2)
    n = 4
    dof = torch.tensor([3.0 * n - 3], dtype=F64)
    m = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=F64)
    si = torch.zeros(n, dtype=torch.long)
    kT = torch.tensor([0.7], dtype=F64)
    p = torch.randn(n, 3, dtype=F64, generator=g)
    tau = torch.tensor([1.0], dtype=F64)
    kes = []
    for _ in range(20000):
        r1 = torch.randn(1, dtype=F64, generator=g)
        r2 = 2 * torch._standard_gamma((dof - 1) / 2, generator=g)
        p = oth.vrescale_update(p, m, si, dof, kT, tau, 1.0, r1, r2)
        kes.append(float(oth.kinetic_energy(p, m, si, 1)))
    ke = torch.tensor(kes[1000:], dtype=F64)
    target_mean = float(dof * kT / 2)
    target_var = float(dof * kT * kT / 2)
    assert abs(float(ke.mean()) / target_mean - 1) < 0.03, float(ke.mean()) / target_mean
    assert abs(float(ke.var()) / target_var - 1) < 0.12, float(ke.var()) / target_var


def test_anisotropic_strain_forces_sum_to_the_isotropic_force():
    torch.manual_seed(3)
    cell = torch.diag(torch.tensor([10.0, 11.0, 12.0], dtype=F64)).unsqueeze(0).repeat(2, 1, 1)
    stress = torch.randn(2, 3, 3, dtype=F64) * 1e-3
    stress = 0.5 * (stress + stress.mT)
    n = torch.tensor([40.0, 64.0], dtype=F64)
    kT = torch.tensor([0.02, 0.05], dtype=F64)
    pext = torch.tensor([1e-3, -2e-4], dtype=F64)
    f3 = oth.aniso_strain_force(cell, stress, n, kT, pext.unsqueeze(-1).expand(2, 3))
    f_iso = omd.strain_force(stress, torch.linalg.det(cell), n, kT, pext)
    torch.testing.assert_close(f3.sum(-1), f_iso, rtol=1e-13, atol=0)


def test_anisotropic_step_at_balanced_pressure_moves_strains_ballistically():
    n = 8
    si = torch.zeros(n, dtype=torch.long)
    m = torch.full((n,), 2.0, dtype=F64)
    x = torch.rand(n, 3, dtype=F64, generator=torch.Generator().manual_seed(4)) * 5
    h0 = torch.tensor([[[5.0, 0.5, 0.0], [0.0, 5.5, 0.3], [0.0, 0.0, 6.0]]], dtype=F64)
    eps = torch.tensor([[0.01, -0.02, 0.005]], dtype=F64)
    cell = torch.exp(eps).unsqueeze(-1) * h0
    eps_dot = torch.tensor([[0.3, -0.1, 0.2]], dtype=F64)
    kT = torch.tensor([0.1], dtype=F64)
    pext = torch.tensor([[1e-3, 2e-3, -1e-3]], dtype=F64)

    def stress_for(c):  # P_kk = pext_k exactly: -sigma_kk + n kT / V = pext_k
        vol = torch.linalg.det(c)
        return torch.diag_embed(n * kT / vol - pext)

    def ff(xx, c):
        return torch.zeros(1, dtype=F64), torch.zeros_like(xx), stress_for(c)

    zero = torch.zeros(1, dtype=F64)
    dt = 0.05
    out = oth.npt_langevin_anisotropic_step(
        x, torch.zeros(n, 3, dtype=F64), torch.zeros(n, 3, dtype=F64), m, si, h0, cell, eps,
        eps_dot, stress_for(cell), ff, dt, kT, pext, zero.expand(1).clone(), zero.clone(),
        torch.ones(1, dtype=F64), torch.zeros(1, 3, dtype=F64), torch.zeros(n, 3, dtype=F64))
    x1, p1, _, _, _, cell1, eps1, eps_dot1 = out
    torch.testing.assert_close(eps1, eps + dt * eps_dot, rtol=1e-14, atol=1e-16)
    torch.testing.assert_close(eps_dot1, eps_dot, rtol=1e-12, atol=1e-14)
    torch.testing.assert_close(cell1, torch.exp(eps1).unsqueeze(-1) * h0, rtol=1e-14, atol=0)
    # every Cartesian component of every position scales by exp(eps_k' - eps_k)
    torch.testing.assert_close(x1, x * torch.exp(dt * eps_dot), rtol=1e-14, atol=0)
    assert p1.abs().max() == 0


def _argon(reps=(2, 2, 2), a=5.26, jitter=0.02, seed=5):
    g = torch.Generator().manual_seed(seed)
    base = torch.tensor([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], dtype=F64)
    idx = torch.stack(torch.meshgrid(*[torch.arange(r) for r in reps], indexing="ij"), -1)
    frac = (idx.reshape(-1, 1, 3).to(F64) + base.unsqueeze(0)).reshape(-1, 3)
    x = frac * a + torch.randn(frac.shape, dtype=F64, generator=g) * jitter
    cell = torch.diag(torch.tensor(reps, dtype=F64) * a).unsqueeze(0)
    return x, cell


def _soft_force_fn(n):
    """soft spheres with alpha = 3: energy, force and curvature vanish at the
    cutoff, so no truncation jumps spoil the conservation check"""
    si = torch.zeros(n, dtype=torch.long)
    z = torch.full((n,), 18, dtype=torch.long)
    pbc = torch.tensor([True, True, True])

    def pair(d, zi, zj):
        return om.soft_sphere_pair(d, zi, zj, sigma=4.6, epsilon=5.0, alpha=3.0)

    def force_fn(x, cell):
        out = om.pair_forward(x, cell, pbc, z, si, pair, 4.6)
        return out["energy"], out["forces"], out["stress"]
    return force_fn


def test_mtk_nose_hoover_npt_conserves_its_extended_hamiltonian():
    x, cell = _argon()
    n = x.shape[0]
    si = torch.zeros(n, dtype=torch.long)
    m = torch.full((n,), 39.948, dtype=F64)
    ff = _soft_force_fn(n)
    kT = torch.tensor([40.0 * 8.617333262e-5], dtype=F64)
    g = torch.Generator().manual_seed(6)
    p = torch.randn(n, 3, dtype=F64, generator=g) * torch.sqrt(m * kT).unsqueeze(-1)
    p = p - p.mean(0)
    dt = 0.5 * 0.0982269  # 5 fs in internal units (1 ps = 98.2269)
    # near the crystal's virial pressure (8.1e-3 eV/A^3): the cell breathes
    # while the spheres stay in contact
    pext = torch.tensor([8e-3], dtype=F64)
    e, f, stress = ff(x, cell)
    dof = torch.tensor([3.0 * n - 3], dtype=F64)
    b_tau = torch.tensor([100 * dt], dtype=F64)
    therm = oth.nhc_init(dof, oth.kinetic_energy(p, m, si, 1), kT,
                         torch.tensor([20 * dt], dtype=F64), 3)
    baro = oth.nhc_init(torch.ones(1, dtype=F64), torch.zeros(1, dtype=F64), kT, b_tau, 3)
    eps = torch.zeros(1, dtype=F64)
    cm = torch.zeros(1, dtype=F64)
    w = 3 * (n + 1) * kT * b_tau ** 2
    h = []
    for step in range(300):
        x, p, f, e, stress, cell_new, eps, cm, w, therm, baro = oth.npt_nose_hoover_isotropic_step(
            x, p, f, m, si, cell, stress, ff, dt, kT, pext, eps, cm, therm, baro, 2, 3, b_tau)
        if step % 50 == 0 or step == 299:
            h.append(float(oth.npt_nose_hoover_isotropic_invariant(
                e, p, m, si, cell, eps, cm, w, therm, baro, kT, pext)))
    ke_scale = float(dof * kT / 2)
    assert max(h) - min(h) < 2e-3 * ke_scale, (max(h) - min(h)) / ke_scale
    assert abs(float(eps)) > 0


def _crescale_inputs(shear=True):
    torch.manual_seed(7)
    n = 12
    si = torch.cat([torch.zeros(6, dtype=torch.long), torch.ones(6, dtype=torch.long)])
    m = torch.linspace(1.0, 2.0, n, dtype=F64)
    x = torch.rand(n, 3, dtype=F64) * 4
    p = torch.randn(n, 3, dtype=F64) * 0.1
    cell = torch.diag(torch.tensor([6.0, 6.5, 7.0], dtype=F64)).repeat(2, 1, 1)
    if shear:
        cell[1, 0, 1] = 0.4
        cell[1, 1, 2] = -0.3
    return x, p, m, si, cell


def test_crescale_at_balanced_pressure_without_noise_leaves_the_cell():
    x, p, m, si, cell = _crescale_inputs()
    kT = torch.tensor([0.05, 0.08], dtype=F64)
    vol = torch.det(cell)
    beta_t = torch.tensor([0.01, 0.02], dtype=F64)
    tau_p = torch.tensor([1.0, 2.0], dtype=F64)
    dt = 0.01
    pext = torch.tensor([1e-3, -2e-3], dtype=F64)
    zero = torch.zeros(2, dtype=F64)
    for mode in ("isotropic", "anisotropic", "triclinic", "average"):
        # an internal pressure tensor p_int I with p_int = pext - kT / (2V):
        # the volume does not move and the deviatoric drive vanishes
        p_int = pext - kT / (2 * vol)
        if mode == "average":
            dof = 3 * torch.bincount(si).to(F64) - 3
            kin_p = dof / 3 * kT / vol
        else:
            kin = oth.per_system_sum(torch.einsum("bi,bj,b->bij", p, p, 1 / m), si, 2)
            kin_p = None
        stress = ((kin_p - p_int)[:, None, None] * torch.eye(3, dtype=F64) if mode == "average"
                  else kin / vol[:, None, None] - p_int[:, None, None] * torch.eye(3, dtype=F64))
        rz = torch.zeros((2, 3) if mode == "anisotropic" else (2, 3, 3), dtype=F64)
        x1, p1, c1 = oth.crescale_barostat(mode, x, p, m, si, cell, stress, kT, beta_t, tau_p,
                                           dt, pext, zero, rz, zero)
        torch.testing.assert_close(c1, cell, rtol=1e-13, atol=1e-13)
        torch.testing.assert_close(x1, x + p * dt / m.unsqueeze(-1), rtol=1e-12, atol=1e-13)
        torch.testing.assert_close(p1, p, rtol=1e-12, atol=1e-15)


def test_crescale_triclinic_deformation_preserves_volume():
    x, p, m, si, cell = _crescale_inputs()
    g = torch.Generator().manual_seed(8)
    kT = torch.tensor([0.05, 0.08], dtype=F64)
    beta_t = torch.tensor([0.05, 0.02], dtype=F64)
    tau_p = torch.tensor([0.5, 1.0], dtype=F64)
    dt = 0.02
    pext = torch.tensor([3e-3, -1e-3], dtype=F64)
    stress = torch.randn(2, 3, 3, dtype=F64, generator=g) * 1e-3
    stress = 0.5 * (stress + stress.mT)
    ra, rb = torch.randn(2, dtype=F64, generator=g), torch.randn(2, dtype=F64, generator=g)
    R = torch.randn(2, 3, 3, dtype=F64, generator=g)
    vol = torch.det(cell)
    for mode in ("triclinic", "average"):
        x1, p1, c1 = oth.crescale_barostat(mode, x, p, m, si, cell, stress, kT, beta_t, tau_p,
                                           dt, pext, ra, R, rb)
        if mode == "average":
            dof = 3 * torch.bincount(si).to(F64) - 3
            P = (dof / 3 * kT / vol)[:, None, None] * torch.eye(3, dtype=F64) - stress
        else:
            P = oth.instantaneous_pressure(p, m, si, stress, vol)
        sqv = torch.sqrt(vol)
        trp = torch.einsum("bii->b", P)
        half = -(beta_t * sqv / (2 * tau_p)) * (pext - trp / 3 - kT / (2 * vol)) * dt / 2
        pr = torch.sqrt(kT * beta_t * dt / (4 * tau_p))
        nsv = sqv + half + pr * ra + half + pr * rb
        torch.testing.assert_close(torch.det(c1), vol * (nsv / sqv) ** 2, rtol=1e-12, atol=0)
        assert (c1 - cell).abs().max() > 1e-6


def test_gram_schmidt_rows_keep_lengths_and_angles():
    g = torch.Generator().manual_seed(9)
    b = torch.eye(3, dtype=F64) + 0.3 * torch.randn(4, 3, 3, dtype=F64, generator=g)
    o = oth.rotate_gram_schmidt(b)
    assert torch.equal(o, torch.tril(o)) and (torch.diagonal(o, dim1=1, dim2=2) > 0).all()
    torch.testing.assert_close(o @ o.mT, b.mT @ b, rtol=1e-13, atol=1e-13)
    assert math.isclose(float(torch.det(o[0])), abs(float(torch.det(b[0]))), rel_tol=1e-12)
