"""Fused, graph-captured batched MD (NVE / Langevin NVT) on one GPU.

The reference advances every system with ~10 small torch launches, a fresh
neighbor list and several host synchronisations per step
(torch_sim/integrators/nvt.py:144-240 with pair_potential.py:351-464).  Here
one step is four launches, all device-resident and captured into one
hipGraph for two steps:

  1. gated cell-list rebuild (tsx_nl_build with ``gate``): runs only when the
     previous step found an atom farther than skin/2 from its position at the
     last build; otherwise every kernel exits at once,
  2. gated copy of the positions into the displacement reference,
  3. tsx_md_fused_step: pair forces on the current positions, the closing
     B(dt/2) of the step and the opening B A O A (or B A) of the next step
     into the other position buffer, plus the displacement test,
  4. tsx_md_reduce: fixed-order per-system energies, Philox step counter.

Positions live in a packed [N,4] float64 layout (x, y, z, species bits); the
list is built at r_list = cutoff + skin and every pair is re-tested against
the cutoff in FP64 inside the force kernel, so the pair set used for forces
equals the reference's (pair_potential.py:194).

Langevin noise is drawn from a counter-based Philox4x32-10 stream keyed by
(seed, step, atom index), or supplied as a tensor per step for comparisons.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from ._lib import call, ptr, stream
from .models import PairPotentialModel
from .neighbors import STATUS_BIN_OVERFLOW, STATUS_NL_OVERFLOW, NeighborList, SystemIndex

NVE, LANGEVIN = 0, 1
WORDS_STENCIL = 1  # TSX_WORDS_STENCIL


class BatchedMD:
    """Graph-captured batched MD for one of the closed-form pair models.

    Args:
        model: a :class:`PairPotentialModel` (LJ / Morse / soft sphere).
        state: an MD state (positions, momenta, masses, cell, pbc, ...).
        dt: time step in internal units.
        kT: temperature in energy units, scalar or per system (Langevin).
        gamma: friction in internal inverse time units (Langevin); defaults
            to 1 / (100 dt) as nvt.py:227.
        langevin: BAOAB Langevin (True) or velocity Verlet NVE (False).
        skin: Verlet skin added to the model cutoff for the stored list.
        tpa: lanes per atom in the pair kernel (1, 2, 4 or 8).
        bins_per_rlist: cell-list bins per list radius (1 to 4).
        seed: Philox key for the Langevin noise.
        use_graph: capture two steps into a hipGraph and replay it.
        chunk_steps: steps between host checks of the device status word.
    """

    def __init__(
        self,
        model: PairPotentialModel,
        state,
        *,
        dt: float,
        kT=0.0,
        gamma=None,
        langevin: bool = True,
        skin: float = 0.5,
        tpa: int = 4,
        bins_per_rlist: int = 2,
        seed: int = 0,
        use_graph: bool = True,
        chunk_steps: int = 200,
    ):
        if not (isinstance(model, PairPotentialModel) and model.fused_kernels):
            raise TypeError("BatchedMD runs the closed-form pair potentials of torch_sim_amd "
                            "(LJ, Morse, soft sphere and their species-pair forms)")
        _lib.require_gpu(state.positions, "state.positions")
        if state.dtype != torch.float64:
            raise TypeError("BatchedMD integrates in float64")
        self.model = model
        self.dt = float(dt)
        self.langevin = bool(langevin)
        self.skin = float(skin)
        self.seed = int(seed)
        self.use_graph = bool(use_graph)
        self.chunk_steps = int(chunk_steps)
        dev = state.device
        n, s = state.n_atoms, state.n_systems
        self.n_atoms, self.n_systems = n, s
        self.device = dev
        self.state_template = state
        self.sysidx = SystemIndex(state.system_idx, s)
        self.sysidx.reduce_ws  # allocated (zeroed) before any graph capture
        self.cell = state.cell.detach().to(torch.float64).contiguous()
        # every cell orthorhombic (off-diagonal entries exactly zero): the
        # fused kernel's stencil shift then takes one product per component
        diag = torch.diag_embed(torch.diagonal(self.cell, dim1=-2, dim2=-1))
        self.orthorhombic = bool((self.cell == diag).all())
        pbc = state.pbc
        self.pbc_per_system = pbc.ndim == 2
        self.pbc = pbc.to(torch.uint8).contiguous()
        self._pbc_bool = pbc
        # packed positions, two buffers
        self.pos = [torch.zeros((n, 4), dtype=torch.float64, device=dev) for _ in range(2)]
        self.pos[0][:, :3] = state.positions.detach()
        species = model.species_of(state.atomic_numbers)
        if species is not None:
            for b in self.pos:
                b.view(torch.int64)[:, 3] = species.to(torch.int64)
        self.pos[1].copy_(self.pos[0])
        self.xref = self.pos[0].clone()
        mom = getattr(state, "momenta", None)
        self.p = (mom.detach().clone() if mom is not None else torch.zeros((n, 3), dtype=torch.float64, device=dev)).contiguous()
        self.m = state.masses.detach().to(torch.float64).contiguous()
        self.f = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        self.e_atom = torch.zeros(n, dtype=torch.float64, device=dev)
        self.energy = torch.zeros(s, dtype=torch.float64, device=dev)
        kT_t = torch.as_tensor(kT, dtype=torch.float64, device=dev)
        self.kT = (kT_t.expand(s) if kT_t.ndim == 0 else kT_t).contiguous().clone()
        g = (1.0 / (100.0 * self.dt)) if gamma is None else gamma
        g_t = torch.as_tensor(g, dtype=torch.float64, device=dev)
        self.gamma = (g_t.expand(s) if g_t.ndim == 0 else g_t).contiguous().clone()
        self.step_ctr = torch.ones(1, dtype=torch.int64, device=dev)
        # per-system rebuild gates, one row per step parity
        self.flags = torch.zeros((2, s), dtype=torch.int32, device=dev)
        self.rebuilds = torch.zeros(1, dtype=torch.int32, device=dev)
        self.nl = NeighborList(model.cutoff + self.skin, tpa=tpa, bins_per_rlist=bins_per_rlist,
                               word_mode=WORDS_STENCIL)
        # bin-relative frame of the stencil-word list: anchors, r - A of each
        # position buffer (float32, one 16-byte record per neighbour gather),
        # per-system bin steps
        self.anchor = torch.zeros((n, 4), dtype=torch.float64, device=dev)
        self.rel = [torch.zeros((n, 4), dtype=torch.float32, device=dev) for _ in range(2)]
        self.ustep = torch.zeros((s, 12), dtype=torch.float64, device=dev)
        self._frames = [
            _lib.NLFrame(self.anchor.data_ptr(), self.rel[q].data_ptr(), self.ustep.data_ptr())
            for q in range(2)
        ]
        # With systems of at most 2048 atoms the per-system reduction of a step
        # (energies, clearing the gate row its rebuild read, the step counter)
        # runs inside the next step's rebuild launch; the last step of a run
        # and the other layouts keep the reduction launch of their own.
        self._merge_reduce = self.nl.one_block_build and 0 < self.sysidx.max_atoms <= 2048
        # a system this large reuses a neighbour's position only after many
        # megabytes of list words: stream the words past L2
        self.stream_words = self.sysidx.max_atoms > 16384
        self._frames_red = [
            _lib.NLFrame(self.anchor.data_ptr(), self.rel[q].data_ptr(), self.ustep.data_ptr(),
                         self.e_atom.data_ptr(), self.energy.data_ptr(),
                         self._gate(1 - q).data_ptr(), self.step_ctr.data_ptr())
            for q in range(2)
        ]
        self.cur = 0
        self.steps_done = 0
        self._graph = None
        self._initialized = False
        self._is_open = False
        self._noise = None

    # ---------------------------------------------------------- kernels
    def _gate(self, q: int) -> torch.Tensor:
        return self.flags[q]

    def _fused(self, q: int, *, end_kick: bool, do_next: bool, noise=None) -> None:
        nl = self.nl
        m = self.model
        call(
            "tsx_md_fused_step",
            m.kind, ptr(m._params), m.n_species, ptr(self.pos[q]), ptr(self.pos[1 - q]),
            ptr(self.xref), ptr(self.cell), ptr(self.sysidx.sys_start), ptr(self.sysidx.atom_sys),
            ptr(nl.nbr), ptr(nl.count), ptr(nl.order), self.n_atoms, nl.jbits, nl.tpa, nl.ld,
            m.cutoff,
            ptr(self.p), ptr(self.m), ptr(self.f), ptr(self.e_atom), None, self.dt,
            int(self.langevin), ptr(self.kT), ptr(self.gamma), ptr(noise), self.seed,
            ptr(self.step_ctr), int(end_kick), int(do_next), 0.5 * self.skin,
            ptr(self._gate(1 - q)), ptr(self.rel[q]), ptr(self.rel[1 - q]), ptr(self.anchor),
            ptr(self.ustep), int(self.orthorhombic), int(self.stream_words), stream(),
        )

    def _reduce(self, q: int) -> None:
        call(
            "tsx_md_reduce", ptr(self.e_atom), ptr(self.sysidx.sys_start),
            self.sysidx.max_atoms, self.n_systems, ptr(self.energy), ptr(self.step_ctr), 1,
            ptr(self._gate(q)), ptr(self.sysidx.reduce_ws), stream(),
        )

    def _rebuild(self, q: int) -> None:
        """Gated list rebuild of the systems flagged in gate row q (cells are
        constant during MD, so the geometry from the first build is reused)."""
        self.nl.enqueue(
            self.pos[q], 4, self.cell, self.pbc, self.pbc_per_system, self.sysidx,
            setup=False, gate=self._gate(q), xref=self.xref, counter=self.rebuilds,
            frame=self._frames_red[q] if self._merge_reduce else self._frames[q],
        )

    def _step(self, q: int, do_next: bool, noise=None) -> None:
        self._rebuild(q)
        self._fused(q, end_kick=True, do_next=do_next, noise=noise)
        if not (self._merge_reduce and do_next):  # else the next rebuild reduces
            self._reduce(q)

    # ------------------------------------------------------------ control
    def initialize(self) -> None:
        """Host-checked list build on the initial positions."""
        if self.n_atoms:
            self.nl.build(
                self.pos[0], self.cell, self._pbc_bool, self.sysidx, pos_stride=4, xref=self.xref,
                frame=self._frames[0],
            )
        self.flags.zero_()
        self.cur = 0
        self._initialized = True
        self._is_open = False

    def _open(self) -> None:
        """Forces and energies at the current positions (the state of the
        step boundary) and the opening half of the next step."""
        q = self.cur
        self._fused(q, end_kick=False, do_next=True, noise=self._next_noise())
        if not self._merge_reduce:  # else the next rebuild reduces
            self._reduce(q)
        self.cur = 1 - q

    def _next_noise(self):
        if self._noise is None:
            return None
        return next(self._noise)

    def _capture(self) -> None:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step(1, True)
            self._step(0, True)
        self._graph = g

    def _snapshot(self):
        return (
            self.cur,
            self.pos[self.cur].clone(),
            self.p.clone(),
            self.step_ctr.clone(),
            self.rebuilds.clone(),
        )

    def _restore(self, snap) -> None:
        cur, pos, p, ctr, nreb = snap
        self.cur = cur
        self.pos[cur].copy_(pos)
        self.p.copy_(p)
        self.step_ctr.copy_(ctr)
        self.rebuilds.copy_(nreb)

    def _run_chunk(self, n_full: int) -> None:
        """Run ``n_full`` steps that each open the next step (do_next)."""
        k = 0
        while k < n_full:
            if self._noise is not None:
                self._step(self.cur, True, noise=self._next_noise())
                self.cur = 1 - self.cur
                k += 1
            elif self.use_graph and self.cur == 1 and n_full - k >= 2:
                if self._graph is None:
                    self._capture()
                self._graph.replay()
                k += 2
            else:
                self._step(self.cur, True)
                self.cur = 1 - self.cur
                k += 1

    def run(self, n_steps: int, noise=None) -> None:
        """Advance all systems by ``n_steps`` reference steps.

        The last step is left open: its positions are final, its closing
        half kick waits for the next run (which does it in the same launch as
        the next step's forces) or for :meth:`close` / :meth:`state`.  A run
        of n steps therefore costs n force passes (n + 1 from a closed
        state), and consecutive runs equal one run of their total length bit
        for bit.

        ``noise``: optional iterable of [N,3] float64 tensors, one per step
        (Langevin trajectory comparisons); runs without graph replay.
        """
        if n_steps <= 0:
            return
        self._noise = iter(noise) if noise is not None else None
        if not self._initialized:
            self.initialize()
        if self.n_atoms == 0:
            return
        remaining = n_steps
        if not self._is_open:
            self._open()
            self._is_open = True
            remaining -= 1
        done = 0
        while done < remaining:
            todo = min(self.chunk_steps, remaining - done)
            snap = self._snapshot()
            self.nl.status.zero_()
            self._run_chunk(todo)
            st = int(self.nl.status.item())
            if st & (STATUS_NL_OVERFLOW | STATUS_BIN_OVERFLOW):
                if self._noise is not None:
                    raise _lib.TsxError("neighbor-list overflow during a supplied-noise run")
                self._restore(snap)
                kmax = self.nl.kmax * 2 if st & STATUS_NL_OVERFLOW else self.nl.kmax
                bins = self.nl.bin_cap * 2 if st & STATUS_BIN_OVERFLOW else self.nl.bin_cap
                self.nl.ensure(self.n_atoms, self.n_systems, kmax, bins, self.nl.jbits, self.device)
                self._graph = None
                # full rebuild (with geometry) at the restored positions
                self.nl.build(self.pos[self.cur], self.cell, self._pbc_bool, self.sysidx,
                              pos_stride=4, xref=self.xref, frame=self._frames[self.cur])
                self.flags.zero_()
                continue
            if st:
                raise _lib.TsxError(f"engine status {st:#x} (shift overflow or singular cell)")
            done += todo
        self.steps_done += n_steps
        self._noise = None

    def close(self) -> None:
        """Finish the open step: forces at its final positions, the closing
        half kick and the per-system energies (positions unchanged)."""
        if self._is_open:
            self._step(self.cur, False)
            self.cur = 1 - self.cur
            self._is_open = False

    # ------------------------------------------------------------- output
    @property
    def positions(self) -> torch.Tensor:
        """Positions at the end of the last step (final also while the step
        is open: only its momenta wait for the closing kick)."""
        return self.pos[self.cur][:, :3]

    def state(self):
        """MD state with the engine's current positions, momenta and forces
        (closes an open step first)."""
        from .integrators import MDState

        self.close()
        t = self.state_template
        return MDState(
            positions=self.positions.clone(),
            masses=t.masses,
            cell=t.cell,
            pbc=t.pbc,
            atomic_numbers=t.atomic_numbers,
            system_idx=t.system_idx,
            momenta=self.p.clone(),
            energy=self.energy.clone(),
            forces=self.f.clone(),
        )

    @property
    def n_rebuilds(self) -> int:
        """Number of per-system list rebuilds run by the gated builds."""
        return int(self.rebuilds.item())

    @property
    def mean_neighbors(self) -> float:
        return float(self.nl.count[: self.n_atoms].float().mean().item())

    def algorithmic_bytes_per_atom_step(self) -> float:
        """Bytes the fused step must move per atom (HBM roofline numerator).

        reads: own packed position 32, anchor 32, displacement reference 32,
        the float32 bin-relative position 16 (every atom's record is gathered
        at least once), momentum 24, mass 8, row order 4, system index 4, row
        length 4, one 4-byte list word per stored neighbour; writes: position
        32, its float32 bin-relative copy 16, momentum 24, force 24, per-atom
        energy 8; reduce: per-atom energy 8 (read).
        """
        k = self.mean_neighbors
        reads = 32 + 32 + 32 + 16 + 24 + 8 + 4 + 4 + 4 + 4.0 * k
        writes = 32 + 16 + 24 + 24 + 8
        return reads + writes + 8


def estimate_skin(cutoff: float) -> float:
    """Default Verlet skin: 8 % of the cutoff, at least 0.3 length units."""
    return max(0.3, 0.08 * cutoff) if math.isfinite(cutoff) else 0.3
