"""Backward kernels (tsx_pair_grad, tsx_nve_adjoint_update) against FP64
autograd of the oracle (create_graph=True through oracle.models.pair_forward,
the restatement of torch_sim/models/pair_potential.py:351-464) and against
central finite differences.

Tolerance: the pair derivatives are FP32 (relative error ~1e-6 per pair
term), sums FP64; gradients are compared to 2e-5 of their largest entry.
Species-pair parameter gradients are symmetrised on both sides
((G + G^T) / 2), the derivative split of a shared parameter.
"""

import pytest
import torch

import torch_sim_amd as tsa
from oracle import md as omd
from oracle import models as om
from oracle import nl as onl
from torch_sim_amd import autograd as tad

pytestmark = pytest.mark.gpu
F64 = torch.float64
KA_SIG = torch.tensor([[1.0, 0.8], [0.8, 0.88]], dtype=F64)
KA_EPS = torch.tensor([[1.0, 1.5], [1.5, 0.5]], dtype=F64)
RC = 2.5


def _setup(dev, n_sys=2, n_min=48, n_max=80, seed=3):
    st = tsa.workloads.binary_lj_batch(n_sys, n_min=n_min, n_max=n_max, device=dev, seed=seed)
    model = tsa.MultiLennardJonesModel(torch.tensor([1, 2]), KA_SIG, KA_EPS, device=dev, cutoff=RC)
    terms = tad.PairTerms(model)
    geo = tad.geometry(terms, st)
    return st, terms, geo


def _oracle(pos, cell, sig, eps, cpu, *, create_graph=True):
    fn = om.MultiPairFn([1, 2], sig, eps, kind="lennard_jones")
    return om.pair_forward(pos, cell, cpu.pbc, cpu.atomic_numbers, cpu.system_idx, fn, RC,
                           compute_stress=False, create_graph=create_graph, nl=onl.linked_cell_nl)


def _sym(g):
    return 0.5 * (g + g.T)


def _close(a, b, tag, rel=2e-5):
    a, b = a.detach().cpu(), b.detach().cpu()
    scale = max(b.abs().max().item(), 1e-12)
    err = (a - b).abs().max().item() / scale
    print(f"{tag}: rel err {err:.2e} (scale {scale:.3e})", flush=True)
    assert err <= rel, (tag, err)


def _leaves(st, dev):
    pos = st.positions.clone().requires_grad_(True)
    cell = st.cell.clone().requires_grad_(True)
    sig = KA_SIG.to(dev).clone().requires_grad_(True)
    eps = KA_EPS.to(dev).clone().requires_grad_(True)
    return pos, cell, sig, eps


def test_force_vjp_matches_oracle_autograd(dev):
    st, terms, geo = _setup(dev)
    cpu = st.to("cpu")
    g = torch.Generator().manual_seed(0)
    w = torch.randn(st.n_atoms, 3, generator=g, dtype=F64)
    pos, cell, sig, eps = _leaves(st, dev)
    params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
    f = tad.pair_forces(pos, cell, params, geo)
    (f * w.to(dev)).sum().backward()

    opos, ocell = cpu.positions.clone().requires_grad_(True), cpu.cell.clone().requires_grad_(True)
    osig, oeps = KA_SIG.clone().requires_grad_(True), KA_EPS.clone().requires_grad_(True)
    of = _oracle(opos, ocell, osig, oeps, cpu)["forces"]
    _close(f, of, "forces")
    gp, gc, gs, ge = torch.autograd.grad((of * w).sum(), [opos, ocell, osig, oeps])
    _close(pos.grad, gp, "d<w,F>/dx (Hessian-vector product)")
    _close(cell.grad, gc, "d<w,F>/dcell")
    _close(sig.grad, _sym(gs), "d<w,F>/dsigma")
    _close(eps.grad, _sym(ge), "d<w,F>/deps")


def test_energy_gradient_matches_oracle_autograd(dev):
    st, terms, geo = _setup(dev, seed=4)
    cpu = st.to("cpu")
    ebar = torch.tensor([0.7, -1.3], dtype=F64)
    pos, cell, sig, eps = _leaves(st, dev)
    params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
    e = tad.pair_energy(pos, cell, params, geo)
    (e * ebar.to(dev)).sum().backward()

    opos, ocell = cpu.positions.clone().requires_grad_(True), cpu.cell.clone().requires_grad_(True)
    osig, oeps = KA_SIG.clone().requires_grad_(True), KA_EPS.clone().requires_grad_(True)
    oe = _oracle(opos, ocell, osig, oeps, cpu)["energy"]
    _close(e, oe, "energy", rel=1e-6)
    gp, gc, gs, ge = torch.autograd.grad((oe * ebar).sum(), [opos, ocell, osig, oeps])
    _close(pos.grad, gp, "dE/dx")
    _close(cell.grad, gc, "dE/dcell")
    _close(sig.grad, _sym(gs), "dE/dsigma")
    _close(eps.grad, _sym(ge), "dE/deps")


def test_force_matching_gradient_finite_differences(dev):
    """d/dtheta of the force-matching loss against FP64 central differences
    of the oracle (the shared off-diagonal parameter moves both entries)."""
    st, terms, geo = _setup(dev, seed=5)
    cpu = st.to("cpu")
    target = _oracle(cpu.positions, cpu.cell, KA_SIG * 1.03, KA_EPS * 0.95, cpu,
                     create_graph=False)["forces"]
    pos, cell, sig, eps = _leaves(st, dev)
    params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
    f = tad.pair_forces(pos, cell, params, geo)
    ((f - target.to(dev)) ** 2).mean().backward()

    def loss(s, e):
        of = _oracle(cpu.positions, cpu.cell, s, e, cpu, create_graph=False)["forces"]
        return ((of - target) ** 2).mean().item()

    h = 1e-6
    for (a, b) in [(0, 0), (0, 1), (1, 1)]:
        for which, grad in (("sigma", sig.grad), ("eps", eps.grad)):
            dp = torch.zeros(2, 2, dtype=F64)
            dp[a, b] = dp[b, a] = h
            if which == "sigma":
                fd = (loss(KA_SIG + dp, KA_EPS) - loss(KA_SIG - dp, KA_EPS)) / (2 * h)
            else:
                fd = (loss(KA_SIG, KA_EPS + dp) - loss(KA_SIG, KA_EPS - dp)) / (2 * h)
            got = grad[a, b].item() * (1 if a == b else 2)
            print(which, a, b, "fd", fd, "kernel", got, flush=True)
            assert abs(got - fd) <= 2e-5 * max(abs(fd), 1e-3), (which, a, b, got, fd)


def test_gradients_are_deterministic(dev):
    st, terms, geo = _setup(dev, seed=6)
    outs = []
    for _ in range(2):
        pos, cell, sig, eps = _leaves(st, dev)
        params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
        f = tad.pair_forces(pos, cell, params, geo)
        (f.square()).sum().backward()
        outs.append((pos.grad.clone(), cell.grad.clone(), sig.grad.clone(), eps.grad.clone()))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def _nve_oracle(cpu, x0, p0, sig, eps, dt, n_steps, c):
    def force_fn(x):
        o = _oracle(x, cpu.cell, sig, eps, cpu)
        return o["energy"], o["forces"]

    x, p = x0, p0
    _, f = force_fn(x)
    for _ in range(n_steps):
        x, p, f, _ = omd.nve_step(x, p, f, cpu.masses, force_fn, dt)
    e = _oracle(x, cpu.cell, sig, eps, cpu)["energy"]
    return e.sum() + (c * p).sum(), x, p


@pytest.mark.parametrize("ckpt", [1, 5])
def test_nve_adjoint_matches_oracle_autograd(dev, ckpt):
    st, terms, geo = _setup(dev, n_sys=2, n_min=32, n_max=48, seed=7)
    cpu = st.to("cpu")
    dt, n_steps = 0.002, 12
    g = torch.Generator().manual_seed(1)
    p0 = torch.randn(st.n_atoms, 3, generator=g, dtype=F64) * 0.7
    c = torch.randn(st.n_atoms, 3, generator=g, dtype=F64)

    x0 = st.positions.clone().requires_grad_(True)
    p0g = p0.to(dev).requires_grad_(True)
    sig = KA_SIG.to(dev).clone().requires_grad_(True)
    eps = KA_EPS.to(dev).clone().requires_grad_(True)
    params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
    xn, pn = tad.nve_trajectory(x0, p0g, params, geo, st.masses, st.cell, dt=dt, n_steps=n_steps,
                                checkpoint_every=ckpt)
    loss = tad.pair_energy(xn, st.cell, params, geo).sum() + (c.to(dev) * pn).sum()
    loss.backward()

    ox0 = cpu.positions.clone().requires_grad_(True)
    op0 = p0.clone().requires_grad_(True)
    osig, oeps = KA_SIG.clone().requires_grad_(True), KA_EPS.clone().requires_grad_(True)
    oloss, oxn, opn = _nve_oracle(cpu, ox0, op0, osig, oeps, dt, n_steps, c)
    _close(xn, oxn, "x_N", rel=1e-6)
    _close(pn, opn, "p_N", rel=1e-5)
    gx, gp, gs, ge = torch.autograd.grad(oloss, [ox0, op0, osig, oeps])
    _close(x0.grad, gx, "dL/dx0", rel=1e-4)
    _close(p0g.grad, gp, "dL/dp0", rel=1e-4)
    _close(sig.grad, _sym(gs), "dL/dsigma", rel=1e-4)
    _close(eps.grad, _sym(ge), "dL/deps", rel=1e-4)


def test_nve_adjoint_checkpoint_invariance(dev):
    st, terms, geo = _setup(dev, n_sys=2, n_min=32, n_max=48, seed=8)
    g = torch.Generator().manual_seed(2)
    p0 = (torch.randn(st.n_atoms, 3, generator=g, dtype=F64) * 0.7).to(dev)
    grads = []
    for ckpt in (1, 4, 10):
        sig = KA_SIG.to(dev).clone().requires_grad_(True)
        eps = KA_EPS.to(dev).clone().requires_grad_(True)
        params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
        xn, pn = tad.nve_trajectory(st.positions, p0, params, geo, st.masses, st.cell, dt=0.002,
                                    n_steps=10, checkpoint_every=ckpt)
        tad.pair_energy(xn, st.cell, params, geo).sum().backward()
        grads.append((sig.grad.clone(), eps.grad.clone()))
    for other in grads[1:]:
        assert torch.equal(grads[0][0], other[0]) and torch.equal(grads[0][1], other[1])


def test_skin_list_gradients_equal_rebuilt_list(dev):
    """Gradients through an NVE trajectory with a Verlet skin list equal those
    with a list rebuilt at every evaluation (same pair sets at the cutoff)."""
    st = tsa.workloads.binary_lj_batch(2, n_min=32, n_max=48, device=dev, seed=9)
    model = tsa.MultiLennardJonesModel(torch.tensor([1, 2]), KA_SIG, KA_EPS, device=dev, cutoff=RC)
    g = torch.Generator().manual_seed(3)
    p0 = (torch.randn(st.n_atoms, 3, generator=g, dtype=F64) * 0.7).to(dev)
    out = []
    for skin in (0.0, 0.3):
        terms = tad.PairTerms(model, skin=skin)
        geo = tad.geometry(terms, st)
        sig = KA_SIG.to(dev).clone().requires_grad_(True)
        eps = KA_EPS.to(dev).clone().requires_grad_(True)
        params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
        xn, pn = tad.nve_trajectory(st.positions, p0, params, geo, st.masses, st.cell, dt=0.002,
                                    n_steps=30, checkpoint_every=7)
        tad.pair_energy(xn, st.cell, params, geo).sum().backward()
        out.append((xn.detach(), sig.grad.clone(), eps.grad.clone()))
    for a, b in zip(*out):
        torch.testing.assert_close(a, b, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("skin", [0.0, 0.3])
def test_nve_graph_segments_equal_eager(dev, skin):
    """Segments replayed from captured HIP graphs (captured on the first
    trajectory of a geometry, replayed on the next ones, eager work between
    them) give bitwise the trajectory and gradients of the same kernels
    launched one by one on a fresh geometry; uneven last segment included
    (30 steps, checkpoints every 7)."""
    st = tsa.workloads.binary_lj_batch(2, n_min=32, n_max=48, device=dev, seed=12)
    model = tsa.MultiLennardJonesModel(torch.tensor([1, 2]), KA_SIG, KA_EPS, device=dev, cutoff=RC)
    g = torch.Generator().manual_seed(5)
    p0 = (torch.randn(st.n_atoms, 3, generator=g, dtype=F64) * 0.7).to(dev)
    c = torch.randn(st.n_atoms, 3, generator=g, dtype=F64).to(dev)

    def run(geo, use_graph):
        sig = KA_SIG.to(dev).clone().requires_grad_(True)
        eps = KA_EPS.to(dev).clone().requires_grad_(True)
        params = tad.pair_params(sig, eps, torch.zeros(2, 2, dtype=F64, device=dev))
        xn, pn = tad.nve_trajectory(st.positions, p0, params, geo, st.masses, st.cell, dt=0.002,
                                    n_steps=30, checkpoint_every=7, use_graph=use_graph)
        loss = tad.pair_energy(xn, st.cell, params, geo).sum() + (c * pn).sum()
        loss.backward()
        torch.cuda.synchronize()
        return xn.detach().clone(), pn.detach().clone(), sig.grad.clone(), eps.grad.clone()

    ref = run(tad.geometry(tad.PairTerms(model, skin=skin), st), False)
    geo = tad.geometry(tad.PairTerms(model, skin=skin), st)
    for rep in range(3):
        print("graph replay", rep, "skin", skin, flush=True)
        out = run(geo, True)
        for a, b in zip(ref, out):
            assert torch.equal(a, b), rep
