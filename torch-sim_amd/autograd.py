"""Differentiable pair terms and the NVE adjoint, on the HIP backward kernels.

The reference obtains these gradients from torch.autograd through
``PairPotentialModel.forward`` (torch_sim/models/pair_potential.py:351-464;
force-matching needs create_graph=True, which the reference's forward does
not set, pair_potential.py:419-425) and through an unrolled NVE loop
(torch_sim/integrators/nve.py:74-130).  Here each piece is a
``torch.autograd.Function`` whose backward calls ``tsx_pair_grad``:

* :func:`pair_forces` -- forces [N,3]; backward = force VJP (a Hessian-vector
  product for positions, the cell term, and the (sigma, epsilon, alpha)
  term): the "force of force" second derivatives of force matching.
* :func:`pair_energy` -- per-system energy [S]; backward = -F, the cell
  gradient and dE/dparams.
* :func:`nve_trajectory` -- n velocity-Verlet steps; backward = the discrete
  adjoint, recomputing each segment of ``checkpoint_every`` steps from its
  stored checkpoint (memory O(checkpoints + segment) instead of O(steps)).

``params`` is the [n_species, n_species, 3] tensor of (sigma, epsilon,
alpha); build it with :func:`pair_params` from tensors that require grad.
Gradients for a species pair come back symmetrised: entries (a, b) and
(b, a) each hold half of the derivative with respect to their shared value.
"""

from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream
from .neighbors import NeighborList, SystemIndex, VerletList


class PairTerms:
    """The fixed part of a differentiable pair model (kind, species, cutoff)
    taken from a :class:`~torch_sim_amd.models.PairPotentialModel`."""

    def __init__(self, model, *, tpa: int | None = None, skin: float = 0.0):
        self.model = model
        self.kind = model.kind
        self.n_species = model.n_species
        self.cutoff = model.cutoff
        self.tpa = tpa or model.tpa
        # skin > 0: a Verlet list kept between calls (trajectories rebuild only
        # the systems that left the skin); every pair is still cut at cutoff
        self.vl = VerletList(self.cutoff, skin, tpa=self.tpa) if skin > 0 else None
        self.nl = NeighborList(self.cutoff, tpa=self.tpa) if self.vl is None else self.vl.nl

    def default_params(self) -> torch.Tensor:
        n = self.n_species
        return self.model._params.view(n, n, 3).clone()


class Geometry:
    """Per-call structure: system layout, species rows and a neighbor list
    built at the given positions (no gradient flows through it)."""

    def __init__(self, terms: PairTerms, positions, cell, pbc, system_idx, atomic_numbers,
                 n_systems: int | None = None):
        # the list is (re)built at the positions of every forward / backward call
        _lib.require_gpu(positions, "positions")
        self.terms = terms
        self.n_systems = int(cell.shape[0]) if n_systems is None else n_systems
        self.sysidx = SystemIndex(system_idx, self.n_systems)
        self.sysidx.counts  # validates the layout
        self.species = terms.model.species_of(atomic_numbers)
        self.pbc = pbc
        self.n_atoms = int(positions.shape[0])
        self.rebuild(positions, cell)

    def rebuild(self, positions, cell) -> None:
        pos = positions.detach().to(torch.float64).contiguous()
        cl = cell.detach().to(torch.float64).contiguous()
        if self.terms.vl is not None:
            self.terms.vl.update(pos, cl, self.pbc, self.sysidx)
        else:
            self.terms.nl.build(pos, cl, self.pbc, self.sysidx)

    # raw kernel calls -----------------------------------------------------
    def forces(self, pos, cell, params) -> torch.Tensor:
        n = self.n_atoms
        f = torch.empty((n, 3), dtype=torch.float64, device=pos.device)
        e = torch.empty(n, dtype=torch.float64, device=pos.device)
        nl = self.terms.nl
        if n:
            call("tsx_pair_forces", self.terms.kind, ptr(params), self.terms.n_species,
                 ptr(self.species), ptr(pos), 3, ptr(cell), ptr(self.sysidx.sys_start),
                 ptr(self.sysidx.atom_sys), ptr(nl.nbr), ptr(nl.count), ptr(nl.order), n,
                 nl.jbits, nl.tpa, nl.ld, self.terms.cutoff, 0, ptr(f), ptr(e), None, stream())
        return f

    def energy(self, pos, cell, params) -> torch.Tensor:
        n, s = self.n_atoms, self.n_systems
        f = torch.empty((n, 3), dtype=torch.float64, device=pos.device)
        e = torch.empty(n, dtype=torch.float64, device=pos.device)
        nl = self.terms.nl
        if n:
            call("tsx_pair_forces", self.terms.kind, ptr(params), self.terms.n_species,
                 ptr(self.species), ptr(pos), 3, ptr(cell), ptr(self.sysidx.sys_start),
                 ptr(self.sysidx.atom_sys), ptr(nl.nbr), ptr(nl.count), ptr(nl.order), n,
                 nl.jbits, nl.tpa, nl.ld, self.terms.cutoff, 0, ptr(f), ptr(e), None, stream())
        energy = torch.empty(s, dtype=torch.float64, device=pos.device)
        call("tsx_system_reduce", ptr(e), None, ptr(cell), ptr(self.sysidx.sys_start),
             self.sysidx.max_atoms, s, ptr(energy), None, None,
             ptr(self.sysidx.reduce_ws), stream())
        return energy

    def grad(self, pos, cell, params, *, w=None, ebar=None, want_pos=True, want_cell=False,
             want_params=False, dparams_acc=None):
        """tsx_pair_grad: returns (dpos, dcell, dparams); dparams accumulates
        into ``dparams_acc`` when given."""
        n, s = self.n_atoms, self.n_systems
        nl = self.terms.nl
        dev = pos.device
        dpos = torch.empty((n, 3), dtype=torch.float64, device=dev) if want_pos else None
        dcell = torch.empty((s, 3, 3), dtype=torch.float64, device=dev) if want_cell else None
        ns = self.terms.n_species
        acc = dparams_acc is not None
        dpar = None
        if want_params:
            dpar = dparams_acc if acc else torch.empty((ns, ns, 3), dtype=torch.float64, device=dev)
        ws_bytes = int(_lib.lib().tsx_grad_ws_bytes(n, nl.tpa, ns))
        ws = torch.empty(max(1, ws_bytes // 8), dtype=torch.float64, device=dev) \
            if (want_cell or want_params) else None
        if n:
            call("tsx_pair_grad", self.terms.kind, ptr(params), ns, ptr(self.species), ptr(pos), 3,
                 ptr(cell), ptr(self.sysidx.sys_start), ptr(self.sysidx.atom_sys), s,
                 ptr(nl.nbr), ptr(nl.count), ptr(nl.order), n, nl.jbits, nl.tpa, nl.ld,
                 self.terms.cutoff, ptr(w), ptr(ebar), ptr(dpos), ptr(dcell), ptr(dpar),
                 int(acc), ptr(ws), stream())
        elif want_params and not acc:
            dpar.zero_()
        return dpos, dcell, dpar


def _f64(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class _PairForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, cell, params, geo: Geometry):
        pos, cl, par = _f64(positions), _f64(cell), _f64(params)
        ctx.geo = geo
        ctx.save_for_backward(pos, cl, par)
        geo.rebuild(pos, cl)
        return geo.forces(pos, cl, par).to(positions.dtype)

    @staticmethod
    def backward(ctx, w):
        pos, cl, par = ctx.saved_tensors
        need = ctx.needs_input_grad
        ctx.geo.rebuild(pos, cl)  # the list may have been rebuilt elsewhere since forward
        dpos, dcell, dpar = ctx.geo.grad(pos, cl, par, w=_f64(w), want_pos=need[0],
                                         want_cell=need[1], want_params=need[2])
        return dpos, dcell, dpar, None


class _PairEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, cell, params, geo: Geometry):
        pos, cl, par = _f64(positions), _f64(cell), _f64(params)
        ctx.geo = geo
        ctx.save_for_backward(pos, cl, par)
        geo.rebuild(pos, cl)
        return geo.energy(pos, cl, par).to(positions.dtype)

    @staticmethod
    def backward(ctx, ebar):
        pos, cl, par = ctx.saved_tensors
        need = ctx.needs_input_grad
        ctx.geo.rebuild(pos, cl)
        dpos, dcell, dpar = ctx.geo.grad(pos, cl, par, ebar=_f64(ebar), want_pos=need[0],
                                         want_cell=need[1], want_params=need[2])
        return dpos, dcell, dpar, None


def pair_params(sigma, epsilon, alpha=None) -> torch.Tensor:
    """[ns, ns, 3] parameter tensor from symmetric (sigma, epsilon, alpha)
    matrices (scalars for one species); differentiable in its inputs."""
    sigma = torch.as_tensor(sigma, dtype=torch.float64)
    epsilon = torch.as_tensor(epsilon, dtype=torch.float64, device=sigma.device)
    if sigma.ndim == 0:
        sigma = sigma.view(1, 1)
    n = sigma.shape[0]
    epsilon = epsilon.expand(n, n) if epsilon.ndim == 0 else epsilon.view(n, n)
    if alpha is None:
        alpha = torch.full((n, n), 2.0, dtype=torch.float64, device=sigma.device)
    alpha = torch.as_tensor(alpha, dtype=torch.float64, device=sigma.device)
    alpha = alpha.expand(n, n) if alpha.ndim == 0 else alpha.view(n, n)
    return torch.stack([sigma, epsilon, alpha], dim=-1)


def geometry(terms: PairTerms, state) -> Geometry:
    """Geometry (layout, species, neighbor list) of a SimState."""
    return Geometry(terms, state.positions, state.cell, state.pbc, state.system_idx,
                    state.atomic_numbers, state.n_systems)


def pair_forces(positions, cell, params, geo: Geometry) -> torch.Tensor:
    """Differentiable forces [N,3] (positions, cell [S,3,3] column, params)."""
    return _PairForces.apply(positions, cell, params.to(positions.device), geo)


def pair_energy(positions, cell, params, geo: Geometry) -> torch.Tensor:
    """Differentiable per-system energy [S]."""
    return _PairEnergy.apply(positions, cell, params.to(positions.device), geo)


class _NVETrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, p0, params, geo: Geometry, masses, cell, dt: float, n_steps: int,
                checkpoint_every: int):
        x, p, par = _f64(x0).clone(), _f64(p0).clone(), _f64(params)
        m, cl = _f64(masses), _f64(cell)
        n = x.shape[0]
        ckpts = []
        geo.rebuild(x, cl)
        f = geo.forces(x, cl, par)
        for k in range(n_steps):
            if k % checkpoint_every == 0:
                ckpts.append((k, x.clone(), p.clone()))
            call("tsx_md_vv_half", ptr(x), ptr(p), ptr(f), ptr(m), dt, n, stream())
            geo.rebuild(x, cl)
            f = geo.forces(x, cl, par)
            call("tsx_md_kick", ptr(p), ptr(f), 0.5 * dt, n, stream())
        ctx.geo, ctx.dt, ctx.n_steps = geo, dt, n_steps
        ctx.ckpts = ckpts
        ctx.save_for_backward(par, m, cl)
        return x, p

    @staticmethod
    def backward(ctx, xbar_n, pbar_n):
        par, m, cl = ctx.saved_tensors
        geo, dt, n_steps = ctx.geo, ctx.dt, ctx.n_steps
        n = m.shape[0]
        dev = m.device
        xbar = _f64(xbar_n).clone() if xbar_n is not None else torch.zeros((n, 3), dtype=torch.float64, device=dev)
        pbar = _f64(pbar_n).clone() if pbar_n is not None else torch.zeros((n, 3), dtype=torch.float64, device=dev)
        dpar = torch.zeros_like(par)
        want_par = ctx.needs_input_grad[2]
        w = 0.5 * dt * pbar  # cotangent of F(x_N): only the last half kick
        for seg in reversed(range(len(ctx.ckpts))):
            k0, xs, ps = ctx.ckpts[seg]
            k1 = ctx.ckpts[seg + 1][0] if seg + 1 < len(ctx.ckpts) else n_steps
            # recompute x_{k0+1 .. k1} from the checkpoint
            xs_list = []
            x, p = xs.clone(), ps.clone()
            geo.rebuild(x, cl)
            f = geo.forces(x, cl, par)
            for _ in range(k0, k1):
                call("tsx_md_vv_half", ptr(x), ptr(p), ptr(f), ptr(m), dt, n, stream())
                xs_list.append(x.clone())
                geo.rebuild(x, cl)
                f = geo.forces(x, cl, par)
                call("tsx_md_kick", ptr(p), ptr(f), 0.5 * dt, n, stream())
            for step in range(k1, k0, -1):
                xn = xs_list[step - k0 - 1]
                geo.rebuild(xn, cl)
                vx, _, _ = geo.grad(xn, cl, par, w=w, want_pos=True, want_params=want_par,
                                    dparams_acc=dpar if want_par else None)
                # step - 1 == 0: the next VJP is at x_0 (first half kick only)
                call("tsx_nve_adjoint_update", ptr(xbar), ptr(pbar), ptr(vx), ptr(m), dt,
                     dt if step - 1 > 0 else 0.5 * dt, ptr(w), n, stream())
        x0 = ctx.ckpts[0][1]
        geo.rebuild(x0, cl)
        vx, _, _ = geo.grad(x0, cl, par, w=w, want_pos=True, want_params=want_par,
                            dparams_acc=dpar if want_par else None)
        xbar = xbar + vx
        return xbar, pbar, (dpar if want_par else None), None, None, None, None, None, None


def nve_trajectory(x0, p0, params, geo: Geometry, masses, cell, *, dt: float, n_steps: int,
                   checkpoint_every: int = 20):
    """(x_n, p_n) after ``n_steps`` velocity-Verlet steps, differentiable in
    x0, p0 and params through the checkpointed discrete adjoint."""
    if n_steps < 1 or checkpoint_every < 1:
        raise ValueError("n_steps and checkpoint_every must be >= 1")
    return _NVETrajectory.apply(x0, p0, params.to(x0.device), geo, masses, cell, float(dt),
                                int(n_steps), int(checkpoint_every))


__all__ = [
    "Geometry",
    "PairTerms",
    "geometry",
    "nve_trajectory",
    "pair_energy",
    "pair_forces",
    "pair_params",
]
