"""Nose-Hoover chain, CSVR, anisotropic Langevin, MTK Nose-Hoover and
stochastic cell-rescaling integrators on the GPU against the oracle
(oracle/thermo.py) with the same draws.

Every step of the GPU path is compared with one oracle step started from a
CPU copy of the GPU state and driven by the same forces (the GPU model on the
oracle's positions and cell), so only the integrator arithmetic is compared:
positions, momenta, cells and the thermostat / barostat variables agree to
1e-11 of their scale (FP64 kernels in the reference's operation order; the
chain's momentum scale is applied once as the product of its substep scales,
the reductions run in a different order, and the C-rescale deformation uses
a Pade exponential where the reference calls torch.matrix_exp).

Beyond the per-step parity: the extended energies of both Nose-Hoover
integrators stay constant over 200 steps of a Cu batch, and integrate()
converts the new integrators' unit-carrying keywords.
"""

import pytest
import torch

import torch_sim_amd as tsa
from oracle import thermo as oth

pytestmark = pytest.mark.gpu
F64 = torch.float64
MORSE = dict(sigma=2.866, epsilon=0.3429, alpha=1.3588)
RC = 6.0
REPS = [(4, 4, 4), (3, 4, 5), (5, 4, 4)]
TOL = 1e-11


def _model(dev):
    return tsa.MorseModel(**MORSE, cutoff=RC, device=dev, compute_stress=True)


def _rel(a, b):
    a, b = torch.as_tensor(a).detach().double().cpu(), torch.as_tensor(b).detach().double().cpu()
    return (a - b).abs().max().item() / max(b.abs().max().item(), 1e-300)


def _state(dev, a=3.61, seed=81, shear=False):
    st = tsa.workloads.cu_batch(len(REPS), device=dev, a=a, jitter=0.03, seed=seed, reps=REPS)
    if shear:  # a triclinic batch: positions and cells sheared alike
        e = torch.tensor([[1.0, 0.04, -0.03], [0.0, 1.0, 0.05], [0.0, 0.0, 1.0]], dtype=F64,
                         device=dev)
        st.positions = (st.positions @ e.T).contiguous()
        st.cell = (e @ st.cell).contiguous()
    st.rng = seed
    return st


def _force_fn(model, ref):
    def force_fn(x, cell):
        probe = tsa.SimState(positions=x.to(ref.device), masses=ref.masses,
                             cell=cell.to(ref.device), pbc=ref.pbc,
                             atomic_numbers=ref.atomic_numbers, system_idx=ref.system_idx)
        o = model(probe)
        return o["energy"].cpu(), o["forces"].cpu(), o["stress"].cpu()
    return force_fn


def _cpu(st, *names):
    return [getattr(st, n).detach().cpu().clone() for n in names]


def _chain_cpu(c):
    return {"xi": c.positions.cpu().clone(), "pxi": c.momenta.cpu().clone(),
            "Q": c.masses.cpu().clone(), "tau": c.tau.cpu().clone(),
            "ke": c.kinetic_energy.cpu().clone(), "dof": c.degrees_of_freedom.cpu().clone()}


def _check(step, errs):
    print(f"step {step}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < TOL, (step, k, v)


def test_nvt_nose_hoover_steps_equal_the_oracle(dev):
    st = _state(dev)
    model = _model(dev)
    kt = 800.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    nh = tsa.nvt_nose_hoover_init(st, model, kT=kt, dt=dt)
    assert nh.chain.positions.shape == (st.n_systems, 3)
    n = torch.bincount(nh.system_idx).double()
    torch.testing.assert_close(nh.chain.masses[:, 1], torch.full_like(n, kt * (10 * dt) ** 2))
    torch.testing.assert_close(nh.chain.masses[:, 0], (3 * n - 3) * kt * (10 * dt) ** 2)
    kts = torch.full((st.n_systems,), kt, dtype=F64)
    ff = _force_fn(model, nh)
    for step in range(8):
        x, p, f, m, si, cell = _cpu(nh, "positions", "momenta", "forces", "masses", "system_idx",
                                    "cell")
        rx, rp, rf, re, rch = oth.nvt_nose_hoover_step(x, p, f, m, si, ff, cell, dt, kts,
                                                       _chain_cpu(nh.chain), 3, 3)
        nh = tsa.nvt_nose_hoover_step(nh, model, dt=dt, kT=kt)
        _check(step, {"x": _rel(nh.positions, rx), "p": _rel(nh.momenta, rp),
                      "xi": _rel(nh.chain.positions, rch["xi"]),
                      "pxi": _rel(nh.chain.momenta, rch["pxi"]),
                      "ke": _rel(nh.chain.kinetic_energy, rch["ke"])})
    assert (nh.chain.momenta.abs() > 0).all()


def test_nvt_nose_hoover_chain_variants_equal_the_oracle(dev):
    """chain_length 1 and 5, one substep (no Suzuki-Yoshida) and 7 SY steps"""
    model = _model(dev)
    kt = 500.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    kts = torch.full((len(REPS),), kt, dtype=F64)
    for cl, nc, nsy in ((1, 1, 1), (5, 2, 7), (2, 1, 5)):
        nh = tsa.nvt_nose_hoover_init(_state(dev, seed=82), model, kT=kt, dt=dt,
                                      tau=0.05 * tsa.MetalUnits.time,
                                      chain_length=cl, chain_steps=nc, sy_steps=nsy)
        ff = _force_fn(model, nh)
        for step in range(3):
            x, p, f, m, si, cell = _cpu(nh, "positions", "momenta", "forces", "masses",
                                        "system_idx", "cell")
            rx, rp, _, _, rch = oth.nvt_nose_hoover_step(x, p, f, m, si, ff, cell, dt, kts,
                                                         _chain_cpu(nh.chain), nc, nsy)
            nh = tsa.nvt_nose_hoover_step(nh, model, dt=dt, kT=kt)
            _check(step, {"x": _rel(nh.positions, rx), "p": _rel(nh.momenta, rp),
                          "pxi": _rel(nh.chain.momenta, rch["pxi"])})


def _smooth_model(dev):
    """soft spheres with alpha = 3 over the first fcc shell: energy, force
    and curvature vanish at the cutoff, so no truncation jumps enter the
    conservation checks"""
    return tsa.SoftSphereModel(sigma=3.2, epsilon=20.0, alpha=3.0, cutoff=3.2, device=dev,
                               compute_stress=True)


def test_nvt_nose_hoover_conserves_its_extended_energy(dev):
    st = _state(dev, seed=83)
    model = _smooth_model(dev)
    kt = 600.0 * tsa.MetalUnits.temperature
    dt = 0.001 * tsa.MetalUnits.time
    nh = tsa.nvt_nose_hoover_init(st, model, kT=kt, dt=dt, tau=0.02 * tsa.MetalUnits.time)
    h0 = tsa.nvt_nose_hoover_invariant(nh, kt)
    ke0 = tsa.calc_kinetic_energy(masses=nh.masses, momenta=nh.momenta, system_idx=nh.system_idx)
    for _ in range(200):
        nh = tsa.nvt_nose_hoover_step(nh, model, dt=dt, kT=kt)
    h1 = tsa.nvt_nose_hoover_invariant(nh, kt)
    drift = ((h1 - h0).abs() / ke0).max().item()
    print(f"extended-energy drift over 200 steps: {drift:.2e} of KE", flush=True)
    assert drift < 2e-3
    assert (nh.chain.positions[:, 0].abs() > 0).all()


def test_nvt_vrescale_steps_equal_the_oracle(dev):
    st = _state(dev, seed=84)
    model = _model(dev)
    kt = 700.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    vr = tsa.nvt_vrescale_init(st, model, kT=kt)
    s = vr.n_systems
    kts = torch.full((s,), kt, dtype=F64)
    ff = _force_fn(model, vr)
    g = torch.Generator().manual_seed(5)
    for step in range(6):
        r1 = torch.randn(s, dtype=F64, generator=g)
        r2 = 2 * torch._standard_gamma(torch.full((s,), 100.0, dtype=F64), generator=g)
        x, p, f, m, si, cell = _cpu(vr, "positions", "momenta", "forces", "masses", "system_idx",
                                    "cell")
        tau = 0.01 * tsa.MetalUnits.time
        rx, rp, _, _, _ = oth.nvt_vrescale_step(x, p, f, m, si, ff, cell, dt, kts,
                                                torch.full((s,), tau, dtype=F64), r1, r2)
        vr = tsa.nvt_vrescale_step(model, vr, dt=dt, kT=kt, tau=tau,
                                   noise=(r1.to(dev), r2.to(dev)))
        _check(step, {"x": _rel(vr.positions, rx), "p": _rel(vr.momenta, rp)})
    # the state's own draws run too, and with tau << dt the kinetic energy is
    # the target's gamma sample: its mean over systems and steps is near kT
    kin = []
    for _ in range(40):
        vr = tsa.nvt_vrescale_step(model, vr, dt=dt, kT=kt, tau=dt / 1000)
        kin.append(vr.calc_kT())
    mean = torch.stack(kin).mean().item()
    assert abs(mean / kt - 1) < 0.05, mean / kt


def test_npt_langevin_anisotropic_steps_equal_the_oracle(dev):
    st = _state(dev, a=3.52, seed=85)
    model = _model(dev)
    kt = 300.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    npt = tsa.npt_langevin_anisotropic_init(st, model, kT=kt, dt=dt)
    assert npt.cell_positions.shape == (st.n_systems, 3)
    s = npt.n_systems
    kts = torch.full((s,), kt, dtype=F64)
    pext = torch.tensor([[0.0, 1e-3, -5e-4], [2e-4, 2e-4, 2e-4], [0.0, 0.0, 1e-3]], dtype=F64)
    ff = _force_fn(model, npt)
    g = torch.Generator().manual_seed(6)
    for step in range(10):
        cn = torch.randn(s, 3, dtype=F64, generator=g)
        pn = torch.randn(npt.n_atoms, 3, dtype=F64, generator=g)
        c = {k: v for k, v in zip(
            ("x", "p", "f", "m", "si", "h0", "cell", "eps", "epsd", "sg", "al", "cal", "bt"),
            _cpu(npt, "positions", "momenta", "forces", "masses", "system_idx", "reference_cell",
                 "cell", "cell_positions", "cell_velocities", "stress", "alpha", "cell_alpha",
                 "b_tau"))}
        ref = oth.npt_langevin_anisotropic_step(
            c["x"], c["p"], c["f"], c["m"], c["si"], c["h0"], c["cell"], c["eps"], c["epsd"],
            c["sg"], ff, dt, kts, pext, c["al"], c["cal"], c["bt"], cn, pn)
        npt = tsa.npt_langevin_anisotropic_step(npt, model, dt=dt, kT=kt,
                                                external_pressure=pext.to(dev),
                                                noise=(cn.to(dev), pn.to(dev)))
        x, p, _, _, _, cell, eps, epsd = ref
        _check(step, {"x": _rel(npt.positions, x), "p": _rel(npt.momenta, p),
                      "cell": _rel(npt.cell, cell), "eps": _rel(npt.cell_positions, eps),
                      "eps_dot": _rel(npt.cell_velocities, epsd)})
        torch.testing.assert_close(npt.current_cell, npt.cell, rtol=1e-15, atol=0)
    # the axes move apart under different target pressures
    e = npt.cell_positions
    assert (e[0, 0] - e[0, 1]).abs() > 0 and (e.abs() > 0).all()


def test_npt_nose_hoover_isotropic_steps_equal_the_oracle(dev):
    st = _state(dev, a=3.53, seed=86)
    model = _model(dev)
    kt = 400.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    b_tau = 0.05 * tsa.MetalUnits.time
    npt = tsa.npt_nose_hoover_isotropic_init(st, model, kT=kt, dt=dt, b_tau=b_tau,
                                             t_tau=0.02 * tsa.MetalUnits.time)
    assert npt.cell_momentum.shape == (st.n_systems, 1)
    s = npt.n_systems
    kts = torch.full((s,), kt, dtype=F64)
    pext = torch.tensor([0.0, 5e-4, -2e-4], dtype=F64)
    ff = _force_fn(model, npt)
    for step in range(8):
        x, p, f, m, si, h0, sg, eps, cm = _cpu(npt, "positions", "momenta", "forces", "masses",
                                              "system_idx", "reference_cell", "stress",
                                              "cell_position", "cell_momentum")
        ref = oth.npt_nose_hoover_isotropic_step(
            x, p, f, m, si, h0, sg, ff, dt, kts, pext, eps, cm.reshape(-1),
            _chain_cpu(npt.thermostat), _chain_cpu(npt.barostat), 2, 3,
            torch.full((s,), b_tau, dtype=F64))
        npt = tsa.npt_nose_hoover_isotropic_step(npt, model, dt=dt, kT=kt,
                                                 external_pressure=pext.to(dev))
        rx, rp, _, _, _, rcell, reps, rcm, rw, rth, rba = ref
        _check(step, {"x": _rel(npt.positions, rx), "p": _rel(npt.momenta, rp),
                      "cell": _rel(npt.cell, rcell), "eps": _rel(npt.cell_position, reps),
                      "p_eps": _rel(npt.cell_momentum.reshape(-1), rcm),
                      "W": _rel(npt.cell_mass, rw),
                      "therm_pxi": _rel(npt.thermostat.momenta, rth["pxi"]),
                      "baro_pxi": _rel(npt.barostat.momenta, rba["pxi"]),
                      "baro_ke": _rel(npt.barostat.kinetic_energy, rba["ke"])})
        torch.testing.assert_close(npt.current_cell, npt.cell, rtol=1e-14, atol=0)
    assert (npt.cell_position.abs() > 0).all()


def test_npt_nose_hoover_conserves_its_extended_energy(dev):
    st = _state(dev, seed=87)
    model = _smooth_model(dev)
    kt = 300.0 * tsa.MetalUnits.temperature
    dt = 0.001 * tsa.MetalUnits.time
    # near the crystal's virial pressure: the cells breathe, the spheres
    # stay in contact
    pext = -torch.einsum("sii->s", model(st)["stress"]) / 3
    npt = tsa.npt_nose_hoover_isotropic_init(st, model, kT=kt, dt=dt,
                                             b_tau=0.1 * tsa.MetalUnits.time,
                                             t_tau=0.02 * tsa.MetalUnits.time)
    h0 = tsa.npt_nose_hoover_isotropic_invariant(npt, kt, pext)
    ke0 = tsa.calc_kinetic_energy(masses=npt.masses, momenta=npt.momenta,
                                  system_idx=npt.system_idx)
    for _ in range(200):
        npt = tsa.npt_nose_hoover_isotropic_step(npt, model, dt=dt, kT=kt, external_pressure=pext)
    h1 = tsa.npt_nose_hoover_isotropic_invariant(npt, kt, pext)
    drift = ((h1 - h0).abs() / ke0).max().item()
    print(f"MTK extended-energy drift over 200 steps: {drift:.2e} of KE", flush=True)
    assert drift < 2e-3
    assert (npt.cell_position.abs() > 0).all()


@pytest.mark.parametrize("mode,deviatoric", [("isotropic", False), ("anisotropic", False),
                                             ("triclinic", False), ("triclinic", True),
                                             ("average", False)])
def test_crescale_steps_equal_the_oracle(dev, mode, deviatoric):
    st = _state(dev, a=3.54, seed=88, shear=mode in ("triclinic", "average"))
    model = _model(dev)
    kt = 500.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    # a compliant barostat so that the cell visibly moves in a few steps
    beta_t = 1e-2
    npt = tsa.npt_crescale_init(st, model, kT=kt, dt=dt, tau_p=0.05 * tsa.MetalUnits.time,
                                isothermal_compressibility=beta_t)
    s = npt.n_systems
    kts = torch.full((s,), kt, dtype=F64)
    tau = 0.02 * tsa.MetalUnits.time
    if deviatoric:
        pe = torch.tensor([[2e-3, 3e-4, 0.0], [3e-4, -1e-3, 1e-4], [0.0, 1e-4, 5e-4]], dtype=F64)
        pext_gpu, pext_cpu = pe.to(dev), torch.zeros(s, dtype=F64)
        p_t = pe.unsqueeze(0).expand(s, 3, 3)
    else:
        pext_cpu = torch.tensor([1e-3, 0.0, -5e-4], dtype=F64)
        pext_gpu, p_t = pext_cpu.to(dev), None
    ff = _force_fn(model, npt)
    g = torch.Generator().manual_seed(7)
    step_fn = {"isotropic": tsa.npt_crescale_isotropic_step,
               "anisotropic": tsa.npt_crescale_anisotropic_step,
               "triclinic": tsa.npt_crescale_triclinic_step,
               "average": tsa.npt_crescale_triclinic_average_step}[mode]
    cell0 = npt.cell.clone()
    for step in range(6):
        z = {"r1a": torch.randn(s, dtype=F64, generator=g),
             "r2a": 2 * torch._standard_gamma(torch.full((s,), 200.0, dtype=F64), generator=g),
             "ra": torch.randn(s, dtype=F64, generator=g),
             "R": torch.randn((s, 3) if mode == "anisotropic" else (s, 3, 3), dtype=F64,
                              generator=g),
             "rb": torch.randn(s, dtype=F64, generator=g),
             "r1b": torch.randn(s, dtype=F64, generator=g),
             "r2b": 2 * torch._standard_gamma(torch.full((s,), 200.0, dtype=F64), generator=g)}
        x, p, f, m, si, cell, sg, c0i, v0 = _cpu(npt, "positions", "momenta", "forces", "masses",
                                                 "system_idx", "cell", "stress",
                                                 "initial_cell_inv", "initial_volume")
        ref = oth.npt_crescale_step(mode, x, p, f, m, si, cell, sg, ff, dt, kts, pext_cpu,
                                    torch.full((s,), beta_t, dtype=F64),
                                    npt.tau_p.cpu(), torch.full((s,), tau, dtype=F64), z,
                                    p_ext_tensor=p_t, cell0_inv=c0i, vol0=v0)
        npt = step_fn(npt, model, dt=dt, kT=kt, external_pressure=pext_gpu, tau=tau,
                      noise={k: v.to(dev) for k, v in z.items()})
        rx, rp, _, _, _, rcell = ref
        _check(step, {"x": _rel(npt.positions, rx), "p": _rel(npt.momenta, rp),
                      "cell": _rel(npt.cell, rcell)})
    assert _rel(npt.cell, cell0) > 1e-6  # the barostat moved the cell


def test_integrate_runs_the_new_integrators_with_unit_conversion(dev):
    model = _model(dev)
    runs = {
        tsa.Integrator.nvt_nose_hoover: dict(init_kwargs=dict(tau=0.05)),
        tsa.Integrator.nvt_vrescale: dict(tau=0.05),
        tsa.Integrator.npt_langevin_anisotropic: dict(
            external_pressure=torch.tensor([0.0, 100.0, -100.0], dtype=F64, device=dev)),
        tsa.Integrator.npt_nose_hoover_isotropic: dict(external_pressure=10.0,
                                                       init_kwargs=dict(b_tau=0.2, t_tau=0.05)),
        tsa.Integrator.npt_crescale_isotropic: dict(external_pressure=10.0, tau=0.05,
                                                    init_kwargs=dict(tau_p=0.5)),
        tsa.Integrator.npt_crescale_triclinic: dict(external_pressure=10.0, tau=0.05),
    }
    for integ, kw in runs.items():
        st = _state(dev, seed=89)
        out = tsa.integrate(st, model, integrator=integ, n_steps=5, temperature=500.0,
                            timestep=0.002, **kw)
        assert torch.isfinite(out.positions).all() and torch.isfinite(out.momenta).all(), integ
    # the loop by hand with the converted keywords equals integrate() bit for bit
    st = _state(dev, seed=90)
    out = tsa.integrate(st, model, integrator=tsa.Integrator.nvt_nose_hoover, n_steps=4,
                        temperature=500.0, timestep=0.002, init_kwargs=dict(tau=0.05))
    kt = 500.0 * tsa.MetalUnits.temperature
    dt = 0.002 * tsa.MetalUnits.time
    st = _state(dev, seed=90)
    nh = tsa.nvt_nose_hoover_init(st, model, kT=torch.tensor(kt, dtype=F64, device=dev), dt=dt,
                                  tau=0.05 * tsa.MetalUnits.time)
    for _ in range(4):
        nh = tsa.nvt_nose_hoover_step(nh, model, dt=dt, kT=kt)
    assert torch.equal(out.positions, nh.positions) and torch.equal(out.momenta, nh.momenta)
