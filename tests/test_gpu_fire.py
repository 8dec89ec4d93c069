"""FIRE steps and relaxations on the GPU (tsx_fire_pre / tsx_fire_post) against
the FP64 oracle (oracle/fire.py restating torch_sim/optimizers/fire.py and
cell_filters.py).

Tolerances: the model's pair arithmetic is FP32 on FP64 displacements
(force error ~1e-6 of the largest force), so positions, cells and velocities
are compared to 1e-6-level absolute tolerances over a dozen steps, and the
integer / branch state (n_pos, dt, alpha) must match exactly.
"""

import math

import pytest
import torch

import torch_sim_amd as tsa
from oracle import fire as of
from oracle import models as om
from oracle import nl as onl

pytestmark = pytest.mark.gpu
F64 = torch.float64
SIG = torch.tensor([[1.0, 1.2], [1.2, 1.4]], dtype=F64)


def _ss(dev, n_sys=3, n_min=32, n_max=64, seed=5, skin=0.0):
    st = tsa.workloads.soft_sphere_batch(n_sys, n_min=n_min, n_max=n_max, device=dev, seed=seed)
    model = tsa.SoftSphereMultiModel(
        atomic_numbers=torch.tensor([1, 2]), sigma_matrix=SIG, device=dev, compute_stress=True,
        skin=skin,
    )
    fn = om.MultiPairFn([1, 2], SIG, torch.ones(2, 2, dtype=F64), torch.full((2, 2), 2.0, dtype=F64))
    cpu = st.to("cpu")

    def model_fn(x, c):
        return om.pair_forward(x, c, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, 1.4,
                               compute_stress=True, nl=onl.linked_cell_nl)

    return st, model, model_fn


def _cmp(g, o, tag, *, cell=True, tol=2e-6):
    err = {}
    err["pos"] = (g.positions.cpu() - o["positions"]).abs().max().item()
    vel = o["velocities"]
    err["vel"] = (g.velocities.cpu() - torch.nan_to_num(vel, nan=0.0)).abs().max().item() \
        if not bool(vel.isnan().all()) else 0.0
    err["f"] = (g.forces.cpu() - o["forces"]).abs().max().item() / max(1.0, o["forces"].abs().max().item())
    err["e"] = ((g.energy.cpu() - o["energy"]).abs() / o["energy"].abs().clamp_min(1e-3)).max().item()
    if cell:
        err["cell"] = (g.cell.cpu() - o["cell"]).abs().max().item()
        err["cf"] = (g.cell_forces.cpu() - o["cell_forces"]).abs().max().item() / max(
            1.0, o["cell_forces"].abs().max().item())
        err["cp"] = (g.cell_positions.cpu() - o["cell_positions"]).abs().max().item()
    print(tag, {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert torch.equal(g.n_pos.cpu(), o["n_pos"]), (tag, g.n_pos, o["n_pos"])
    torch.testing.assert_close(g.dt.cpu(), o["dt"], rtol=1e-12, atol=0)
    torch.testing.assert_close(g.alpha.cpu(), o["alpha"], rtol=1e-12, atol=0)
    for k, v in err.items():
        lim = tol if k not in ("e",) else 1e-6
        assert v <= lim, (tag, k, v)


@pytest.mark.parametrize("filt,skin", [("frechet", 0.0), ("unit", 0.0), ("frechet", 0.25)])
def test_ase_fire_cell_filter_matches_oracle(dev, filt, skin):
    st, model, model_fn = _ss(dev, skin=skin)
    cpu = st.to("cpu")
    g = tsa.fire_init(st, model, cell_filter=filt)
    o = of.fire_init(cpu.positions.clone(), cpu.cell.clone(), cpu.masses, cpu.system_idx, model_fn,
                     frechet=filt == "frechet", unit=filt == "unit")
    _cmp(g, o, f"{filt} init")
    for k in range(12):
        tsa.fire_step(g, model)
        of.ase_fire_step(o, model_fn)
        _cmp(g, o, f"{filt} step {k}", tol=2e-6 * (k + 2))


def test_ase_fire_positions_only_matches_oracle(dev):
    st, model, model_fn = _ss(dev, seed=6)
    cpu = st.to("cpu")
    g = tsa.fire_init(st, model)
    o = of.fire_init(cpu.positions.clone(), cpu.cell.clone(), cpu.masses, cpu.system_idx, model_fn)
    for k in range(15):
        tsa.fire_step(g, model)
        of.ase_fire_step(o, model_fn)
        _cmp(g, o, f"pos-only step {k}", cell=False, tol=2e-6 * (k + 2))
    assert torch.equal(g.cell, st.cell)


def test_vv_fire_matches_oracle(dev):
    st, model, model_fn = _ss(dev, seed=8)
    cpu = st.to("cpu")
    g = tsa.fire_init(st, model, fire_flavor="vv_fire", dt_start=0.01)
    o = of.fire_init(cpu.positions.clone(), cpu.cell.clone(), cpu.masses, cpu.system_idx, model_fn,
                     dt_start=0.01)
    for k in range(15):
        tsa.fire_step(g, model, fire_flavor="vv_fire")
        of.vv_fire_step(o, model_fn)
        _cmp(g, o, f"vv step {k}", cell=False, tol=2e-6 * (k + 2))


def test_hydrostatic_frechet_matches_oracle(dev):
    st, model, model_fn = _ss(dev, seed=9)
    cpu = st.to("cpu")
    g = tsa.fire_init(st, model, cell_filter="frechet", hydrostatic_strain=True, scalar_pressure=0.5)
    o = of.fire_init(cpu.positions.clone(), cpu.cell.clone(), cpu.masses, cpu.system_idx, model_fn,
                     frechet=True, hydrostatic_strain=True, scalar_pressure=0.5)
    for k in range(8):
        tsa.fire_step(g, model)
        of.ase_fire_step(o, model_fn)
        _cmp(g, o, f"hydro step {k}", tol=2e-6 * (k + 2))


def test_fire_first_step_raw_forces(dev):
    """fire.py:313-315: with every velocity NaN the first ASE step is v = F dt."""
    st, model, _ = _ss(dev, seed=10)
    g = tsa.fire_init(st, model)
    f0 = g.forces.clone()
    dt0 = g.dt.clone()
    x0 = g.positions.clone()
    tsa.fire_step(g, model, max_step=1e9)
    idx = st.system_idx
    torch.testing.assert_close(g.positions - x0, f0 * dt0[idx, None] ** 2, rtol=1e-13, atol=1e-15)
    torch.testing.assert_close(g.dt, dt0)
    assert int(g.n_pos.sum()) == 0


def test_fire_max_step_clamp(dev):
    """No atom moves more than max_step in one ASE step (fire.py:404-415)."""
    st, model, _ = _ss(dev, seed=11)
    g = tsa.fire_init(st, model)
    for _ in range(3):
        x0 = g.positions.clone()
        tsa.fire_step(g, model, max_step=0.05)
        d = (g.positions - x0).norm(dim=1)
        assert float(d.max()) <= 0.05 * (1 + 1e-12)


def _ar_lattice_energy(a):
    # FP64 shell sum of LJ Ar fcc energy per atom at lattice constant a (rc = 8.5125)
    import numpy as np

    sig, eps, rc = 3.405, 0.0104, 8.5125
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    m = int(math.ceil(rc / a)) + 1
    g = np.stack(np.meshgrid(*[np.arange(-m, m + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    r = np.linalg.norm(pts, axis=1)
    r = r[(r > 1e-9) & (r < rc)]
    return float(np.sum(2 * eps * ((sig / r) ** 12 - (sig / r) ** 6)))


def test_frechet_relaxes_ar_lattice_constant(dev):
    """An Ar fcc cell relaxed with the Frechet filter reaches the lattice
    constant that minimises the closed-form shell-sum energy."""
    from scipy.optimize import minimize_scalar

    a_eq = minimize_scalar(_ar_lattice_energy, bounds=(5.0, 5.6), method="bounded",
                           options={"xatol": 1e-10}).x
    st = tsa.workloads.ar_fcc(device=dev, reps=(2, 2, 2), a=5.40)
    model = tsa.LennardJonesModel(sigma=3.405, epsilon=0.0104, cutoff=8.5125, device=dev,
                                  compute_stress=True)
    conv = tsa.generate_force_convergence_fn(1e-6, include_cell_forces=True)
    out = tsa.optimize(st, model, optimizer=tsa.Optimizer.fire, convergence_fn=conv,
                       max_steps=4000, init_kwargs=dict(cell_filter="frechet"))
    a = float(torch.det(out.cell[0]).item() ** (1.0 / 3.0)) / 2
    print("a_eq", a_eq, "relaxed", a)
    assert abs(a - a_eq) < 1e-5 * a_eq


def test_optimize_refills_and_keeps_order(dev):
    st, model, _ = _ss(dev, n_sys=6, n_min=32, n_max=96, seed=12)
    conv = tsa.generate_force_convergence_fn(1e-3, include_cell_forces=True)
    out = tsa.optimize(
        st, model, optimizer=tsa.Optimizer.fire, convergence_fn=conv, max_steps=5000,
        init_kwargs=dict(cell_filter="frechet"), autobatcher=tsa.InFlightAutoBatcher(max_atoms=200),
    )
    assert out.n_systems == st.n_systems
    assert torch.equal(out.system_idx, st.system_idx)
    assert torch.equal(out.atomic_numbers, st.atomic_numbers)
    res = model(out)
    f_atoms = torch.zeros(st.n_systems, dtype=F64, device=dev).scatter_reduce(
        0, out.system_idx, res["forces"].norm(dim=1), reduce="amax", include_self=False)
    assert bool((f_atoms < 1.5e-3).all()), f_atoms
    e0 = model(st)["energy"]
    assert bool((res["energy"] <= e0 + 1e-12).all())


def test_optimize_energy_convergence_default(dev):
    st, model, _ = _ss(dev, n_sys=2, n_min=32, n_max=48, seed=13)
    out = tsa.optimize(st, model, optimizer=tsa.Optimizer.fire, max_steps=400)
    assert out.n_systems == 2
    assert bool((model(out)["energy"] <= model(st)["energy"]).all())


def test_optimize_resident_equals_host_path_before_any_swap(dev):
    """With no system converging before max_steps, the device-resident path
    (batch gathered from the initialised pool, list rows relocated) and the
    host path (batch selected and initialised on its own, list built for
    it) step the same batch: bitwise the same result."""
    st, _, _ = _ss(dev, n_sys=5, n_min=32, n_max=96, seed=14)
    never = lambda state, last_energy: torch.zeros(state.n_systems, dtype=torch.bool,  # noqa: E731
                                                   device=state.device)
    outs = []
    for resident in (True, False):
        model = tsa.SoftSphereMultiModel(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=SIG,
                                         device=dev, compute_stress=True, skin=0.3)
        opt = (tsa.fire_init, tsa.fire_step) if resident else (
            lambda **kw: tsa.fire_init(**kw), lambda **kw: tsa.fire_step(**kw))
        stats = {}
        print("resident", resident, flush=True)
        with pytest.warns(UserWarning, match="maximum number of steps"):
            outs.append(tsa.optimize(st, model, optimizer=opt, convergence_fn=never, max_steps=40,
                                     init_kwargs=dict(cell_filter="frechet"), stats=stats))
        assert stats["batch_steps"] == 40 and stats["refills"] == 0
    for name in ("positions", "cell", "velocities", "forces", "energy", "dt", "alpha", "n_pos"):
        assert torch.equal(getattr(outs[0], name), getattr(outs[1], name)), name


def test_optimize_resident_refill_with_skin_converges_in_order(dev):
    """A tight atom budget forces many refills with the list rows relocated
    (skin 0.3): every system is returned in input order, relaxed below the
    tolerance by an independent model call, at or below its initial energy,
    and two runs give the same bits."""
    st, _, _ = _ss(dev, n_sys=10, n_min=32, n_max=128, seed=15)
    conv = tsa.generate_force_convergence_fn(1e-3, include_cell_forces=True)
    runs = []
    for _ in range(2):
        model = tsa.SoftSphereMultiModel(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=SIG,
                                         device=dev, compute_stress=True, skin=0.3)
        stats = {}
        out = tsa.optimize(st, model, optimizer=tsa.Optimizer.fire, convergence_fn=conv,
                           max_steps=5000, init_kwargs=dict(cell_filter="frechet"),
                           autobatcher=tsa.InFlightAutoBatcher(max_atoms=260), stats=stats)
        print(stats, flush=True)
        assert stats["refills"] >= 4 and stats["atom_steps"] > 0
        runs.append(out)
    out = runs[0]
    assert out.n_systems == st.n_systems
    assert torch.equal(out.system_idx, st.system_idx)
    assert torch.equal(out.atomic_numbers, st.atomic_numbers)
    fresh = tsa.SoftSphereMultiModel(atomic_numbers=torch.tensor([1, 2]), sigma_matrix=SIG,
                                     device=dev, compute_stress=True)
    res = fresh(out)
    f_atoms = torch.zeros(st.n_systems, dtype=F64, device=dev).scatter_reduce(
        0, out.system_idx, res["forces"].norm(dim=1), reduce="amax", include_self=False)
    assert bool((f_atoms < 1.5e-3).all()), f_atoms
    assert bool((res["energy"] <= fresh(st)["energy"] + 1e-12).all())
    for name in ("positions", "cell", "energy"):
        assert torch.equal(getattr(runs[0], name), getattr(runs[1], name)), name
