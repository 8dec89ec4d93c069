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
        if not getattr(model, "fused_kernels", False):
            raise TypeError("the backward kernels take the closed-form pair potentials "
                            "(LJ, Morse, soft sphere and their species-pair forms)")
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

    def rebuild(self, positions, cell, *, full: bool = False) -> None:
        """Make the list valid at ``positions`` (``full``: rebuild every
        system, so the list no longer depends on earlier calls)."""
        pos = positions.detach().to(torch.float64).contiguous()
        cl = cell.detach().to(torch.float64).contiguous()
        if self.terms.vl is not None:
            if full:
                self.terms.vl.rebuild_all(pos, cl, self.pbc, self.sysidx)
            else:
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


class _TrajRunner:
    """Static buffers and (optionally) captured HIP graphs for an NVE
    trajectory and its adjoint on one :class:`Geometry`.

    A segment of velocity-Verlet steps (forward), the recomputation of a
    segment from its checkpoint (storing every position), and the reverse
    sweep of a segment (force VJP + adjoint update per step) each run without
    a host synchronisation: the list is kept valid by the capturable skin
    check and gated rebuild, and its capacity status is read once per
    segment.  On an overflow the runner grows the list, rebuilds it at the
    segment's start and redoes the segment eagerly.  Every kernel sequence is
    the same with and without graphs, so results are bitwise equal."""

    def __init__(self, geo: Geometry, masses, cell, dt: float, use_graph: bool):
        self.geo, self.dt, self.use_graph = geo, float(dt), bool(use_graph)
        n, s, ns = geo.n_atoms, geo.n_systems, geo.terms.n_species
        dev = masses.device
        z3 = lambda: torch.zeros((n, 3), dtype=torch.float64, device=dev)  # noqa: E731
        self.x, self.p, self.f = z3(), z3(), z3()
        self.xbar, self.pbar, self.w, self.vx = z3(), z3(), z3(), z3()
        self.e = torch.zeros(max(1, n), dtype=torch.float64, device=dev)
        self.m = _f64(masses).clone()
        self.cl = _f64(cell).clone()
        self.par = torch.zeros((ns, ns, 3), dtype=torch.float64, device=dev)
        self.dpar = torch.zeros_like(self.par)
        self.pbc_u8 = geo.pbc.to(torch.uint8).contiguous()
        self.pbc_per = geo.pbc.ndim == 2
        nl = geo.terms.nl
        self.gws = torch.zeros(max(1, int(_lib.lib().tsx_grad_ws_bytes(n, nl.tpa, ns)) // 8 + 1),
                               dtype=torch.float64, device=dev)
        self.xs = None
        self.graphs: dict = {}
        self.gen = None

    # kernel sequences (no allocation, no host read) -------------------------
    def _update(self, pos):
        geo, t = self.geo, self.geo.terms
        # the cell is constant along a trajectory: the per-system grids of the
        # host-checked build at its start stay valid (setup=False)
        if t.vl is not None:
            t.vl.enqueue_update(pos, self.cl, self.pbc_u8, self.pbc_per, geo.sysidx, setup=False)
        else:
            t.nl.enqueue(pos, 3, self.cl, self.pbc_u8, self.pbc_per, geo.sysidx, setup=False)

    def _forces(self, pos):
        geo, nl = self.geo, self.geo.terms.nl
        call("tsx_pair_forces", geo.terms.kind, ptr(self.par), geo.terms.n_species,
             ptr(geo.species), ptr(pos), 3, ptr(self.cl), ptr(geo.sysidx.sys_start),
             ptr(geo.sysidx.atom_sys), ptr(nl.nbr), ptr(nl.count), ptr(nl.order), geo.n_atoms,
             nl.jbits, nl.tpa, nl.ld, geo.terms.cutoff, 0, ptr(self.f), ptr(self.e), None,
             stream())

    def _vjp(self, pos, want_par: bool):
        geo, nl = self.geo, self.geo.terms.nl
        call("tsx_pair_grad", geo.terms.kind, ptr(self.par), geo.terms.n_species,
             ptr(geo.species), ptr(pos), 3, ptr(self.cl), ptr(geo.sysidx.sys_start),
             ptr(geo.sysidx.atom_sys), geo.n_systems, ptr(nl.nbr), ptr(nl.count), ptr(nl.order),
             geo.n_atoms, nl.jbits, nl.tpa, nl.ld, geo.terms.cutoff, ptr(self.w), None,
             ptr(self.vx), None, ptr(self.dpar) if want_par else None, 1, ptr(self.gws),
             stream())

    def _vv(self):
        n, dt = self.geo.n_atoms, self.dt
        call("tsx_md_vv_half", ptr(self.x), ptr(self.p), ptr(self.f), ptr(self.m), dt, n, stream())
        self._update(self.x)
        self._forces(self.x)
        call("tsx_md_kick", ptr(self.p), ptr(self.f), 0.5 * dt, n, stream())

    def _seq_forward(self, k: int):
        for _ in range(k):
            self._vv()

    def _seq_recompute(self, k: int):
        self._update(self.x)
        self._forces(self.x)
        for j in range(k):
            n, dt = self.geo.n_atoms, self.dt
            call("tsx_md_vv_half", ptr(self.x), ptr(self.p), ptr(self.f), ptr(self.m), dt, n,
                 stream())
            self.xs[j].copy_(self.x)
            self._update(self.x)
            self._forces(self.x)
            call("tsx_md_kick", ptr(self.p), ptr(self.f), 0.5 * dt, n, stream())

    def _seq_reverse(self, k: int, first: bool, want_par: bool):
        n, dt = self.geo.n_atoms, self.dt
        for j in range(k, 0, -1):
            xn = self.xs[j - 1]
            self._update(xn)
            self._vjp(xn, want_par)
            # the global step 1 is last: the next VJP is at x_0 (half kick only)
            wscale = 0.5 * dt if (first and j == 1) else dt
            call("tsx_nve_adjoint_update", ptr(self.xbar), ptr(self.pbar), ptr(self.vx),
                 ptr(self.m), dt, wscale, ptr(self.w), n, stream())

    # execution ---------------------------------------------------------------
    def _generation(self):
        t = self.geo.terms
        return t.vl.generation if t.vl is not None else t.nl.generation

    def _status(self) -> int:
        return int(self.geo.terms.nl.status.item())

    def _run(self, key, fn, *args, graph_ok: bool = True):
        gen = self._generation()
        if gen != self.gen:  # a buffer a graph points to was replaced
            self.graphs.clear()
            self.gen = gen
        self.geo.terms.nl.status.zero_()
        if not (self.use_graph and graph_ok):
            fn(*args)
            return
        g = self.graphs.get(key)
        if g is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                fn(*args)
            self.graphs[key] = g
        g.replay()

    def _recover(self, st: int, pos):
        t = self.geo.terms
        if t.vl is not None:
            t.vl.recover(st, pos, self.cl, self.geo.pbc, self.geo.sysidx)
        else:
            t.nl.build(pos, self.cl, self.geo.pbc, self.geo.sysidx)

    def segment(self, kind: str, k: int, *, first: bool = False, want_par: bool = False):
        """Run one segment; redo it eagerly after growing the list when the
        status word reports an overflow."""
        snap = (self.x.clone(), self.p.clone(), self.f.clone(), self.xbar.clone(),
                self.pbar.clone(), self.w.clone(), self.dpar.clone())
        for attempt in range(8):
            if kind == "fwd":
                self._run(("fwd", k), self._seq_forward, k, graph_ok=attempt == 0)
            elif kind == "rec":
                self._run(("rec", k), self._seq_recompute, k, graph_ok=attempt == 0)
            else:
                self._run(("rev", k, first, want_par), self._seq_reverse, k, first, want_par,
                          graph_ok=attempt == 0)
            st = self._status()
            if not st:
                return
            for dst, src in zip((self.x, self.p, self.f, self.xbar, self.pbar, self.w, self.dpar),
                                snap):
                dst.copy_(src)
            start = self.xs[k - 1] if kind == "rev" else self.x
            self._recover(st, start)
        raise _lib.TsxError("neighbor list capacity could not be resolved in a trajectory")


class _NVETrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, p0, params, geo: Geometry, masses, cell, dt: float, n_steps: int,
                checkpoint_every: int, use_graph: bool):
        # one runner (static buffers, captured segments) per geometry and
        # (dt, use_graph); its inputs are reloaded at every forward / backward
        key = (float(dt), bool(use_graph))
        cache = geo.__dict__.setdefault("_traj_runners", {})
        r = cache.get(key)
        if r is None:
            r = cache[key] = _TrajRunner(geo, masses, cell, dt, use_graph)
        par, m, cl = _f64(params).clone(), _f64(masses).clone(), _f64(cell).clone()
        r.par.copy_(par)
        r.m.copy_(m)
        r.cl.copy_(cl)
        r.x.copy_(_f64(x0))
        r.p.copy_(_f64(p0))
        geo.rebuild(r.x, r.cl, full=True)  # every system at x_0: history-independent
        r._forces(r.x)
        ckpts = []
        for k0 in range(0, n_steps, checkpoint_every):
            ckpts.append((k0, r.x.clone(), r.p.clone()))
            r.segment("fwd", min(checkpoint_every, n_steps - k0))
        ctx.runner, ctx.n_steps, ctx.ckpts = r, n_steps, ckpts
        ctx.inputs = (par, m, cl)
        return r.x.clone(), r.p.clone()

    @staticmethod
    def backward(ctx, xbar_n, pbar_n):
        r, n_steps, ckpts = ctx.runner, ctx.n_steps, ctx.ckpts
        want_par = ctx.needs_input_grad[2]
        for dst, src in zip((r.par, r.m, r.cl), ctx.inputs):
            dst.copy_(src)
        seg_max = max(min(ckpts[i + 1][0] if i + 1 < len(ckpts) else n_steps, n_steps) - ckpts[i][0]
                      for i in range(len(ckpts)))
        if r.xs is None or r.xs.shape[0] < seg_max:
            r.xs = torch.zeros((seg_max, r.geo.n_atoms, 3), dtype=torch.float64,
                               device=r.x.device)
            r.graphs.clear()  # the recompute and reverse graphs point to xs
        r.xbar.copy_(_f64(xbar_n)) if xbar_n is not None else r.xbar.zero_()
        r.pbar.copy_(_f64(pbar_n)) if pbar_n is not None else r.pbar.zero_()
        r.dpar.zero_()
        torch.mul(r.pbar, 0.5 * r.dt, out=r.w)  # cotangent of F(x_N): the last half kick only
        for seg in reversed(range(len(ckpts))):
            k0, xs, ps = ckpts[seg]
            k1 = ckpts[seg + 1][0] if seg + 1 < len(ckpts) else n_steps
            r.x.copy_(xs)
            r.p.copy_(ps)
            r.segment("rec", k1 - k0)
            r.segment("rev", k1 - k0, first=k0 == 0, want_par=want_par)
        # the VJP at x_0 (first half kick)
        r.x.copy_(ckpts[0][1])
        r.geo.rebuild(r.x, r.cl)
        r._vjp(r.x, want_par)
        xbar = r.xbar + r.vx
        return (xbar, r.pbar.clone(), (r.dpar.clone() if want_par else None), None, None, None,
                None, None, None, None)


def nve_trajectory(x0, p0, params, geo: Geometry, masses, cell, *, dt: float, n_steps: int,
                   checkpoint_every: int = 20, use_graph: bool = True):
    """(x_n, p_n) after ``n_steps`` velocity-Verlet steps, differentiable in
    x0, p0 and params through the checkpointed discrete adjoint.  Each
    segment of ``checkpoint_every`` steps runs as one captured HIP graph
    (``use_graph=False``: the same kernels launched one by one)."""
    if n_steps < 1 or checkpoint_every < 1:
        raise ValueError("n_steps and checkpoint_every must be >= 1")
    return _NVETrajectory.apply(x0, p0, params.to(x0.device), geo, masses, cell, float(dt),
                                int(n_steps), int(checkpoint_every), bool(use_graph))


__all__ = [
    "Geometry",
    "PairTerms",
    "geometry",
    "nve_trajectory",
    "pair_energy",
    "pair_forces",
    "pair_params",
]
