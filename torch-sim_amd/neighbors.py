"""Neighbor lists on the GPU through the tsx C ABI.

``torchsim_nl`` keeps the reference signature and output contract
(torch_sim/neighbors/__init__.py:46-98, torch_nl.py:156-215):

    torchsim_nl(positions [N,3], cell [S,3,3] row-vector, pbc [S,3] or [3],
                cutoff, system_idx [N], self_interaction=False)
      -> (mapping [2,P] int64, system_mapping [P] int64, shifts_idx [P,3])

with the full ordered pair set {(i, j, S): |r_j - r_i + S.L| < cutoff},
(i, i, 0) excluded unless ``self_interaction``, shifts valid for the
unwrapped input positions.  The order of the pairs is the build's row order
(the reference order is unspecified, transforms.py:999, 1291).

:class:`NeighborList` owns the device buffers of one list (ELL rows, counts,
the bin-sorted atom order, workspace and the device status word) and is
reused by the models and the batched engine.
"""

from __future__ import annotations

import ctypes
import itertools
import math
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream

STATUS_NL_OVERFLOW = 1
STATUS_BIN_OVERFLOW = 2
STATUS_SHIFT_OVERFLOW = 4
STATUS_SINGULAR_CELL = 8
STATUS_BAD_SIZE = 32


class TensorKey:
    """Identity of a tensor's contents for cache lookups.

    It holds the tensor object through a weak reference plus its in-place
    version counter and shape.  A freed tensor never matches, even when the
    caching allocator hands its memory to a new tensor of the same size (a
    key made of the data pointer would match that new tensor and return stale
    results); an in-place write bumps the version and misses too.
    """

    __slots__ = ("_ref", "_version", "_shape")

    def __init__(self, t: torch.Tensor):
        self._ref = weakref.ref(t)
        self._version = t._version
        self._shape = tuple(t.shape)

    def matches(self, t: torch.Tensor) -> bool:
        return self._ref() is t and t._version == self._version and tuple(t.shape) == self._shape


_SERIAL = itertools.count(1)


class SystemIndex:
    """Device bookkeeping of the system layout: int32 atom->system map and the
    [S+1] atom offsets of each system (no host synchronisation).

    ``serial`` is unique per instance: lists built for one layout record it
    and rebuild in full when they are handed another."""

    def __init__(self, system_idx: torch.Tensor, n_systems: int):
        dev = system_idx.device
        self.n_systems = int(n_systems)
        self.n_atoms = int(system_idx.shape[0])
        self.atom_sys = system_idx.to(torch.int32).contiguous()
        bounds = torch.arange(self.n_systems + 1, device=dev, dtype=system_idx.dtype)
        self.sys_start = torch.searchsorted(system_idx, bounds).to(torch.int32).contiguous()
        self._source = TensorKey(system_idx)
        self.serial = next(_SERIAL)
        self._counts = None
        self._red_ws = None

    @property
    def max_atoms(self) -> int:
        """Atoms of the largest system (host, cached)."""
        c = self.counts
        return int(c.max()) if c.size else 0

    @property
    def reduce_ws(self) -> torch.Tensor:
        """Zeroed workspace of the per-system reductions (tsx_system_reduce /
        tsx_md_reduce); reductions that share it run on one stream."""
        if self._red_ws is None:
            nbytes = int(_lib.lib().tsx_reduce_ws_bytes(self.max_atoms, self.n_systems))
            self._red_ws = torch.zeros((nbytes + 7) // 8, dtype=torch.float64,
                                       device=self.atom_sys.device)
        return self._red_ws

    def matches(self, system_idx: torch.Tensor, n_systems: int) -> bool:
        """True when ``system_idx`` is the very tensor this index was made
        from, unmodified since."""
        return self._source.matches(system_idx) and int(n_systems) == self.n_systems

    @property
    def counts(self) -> np.ndarray:
        """Atoms per system on the host (one synchronisation, cached).

        Also validates the layout the kernels index with: system_idx sorted,
        within [0, n_systems), so that the offsets cover every atom."""
        if self._counts is None:
            start = self.sys_start.cpu().numpy()
            if self.n_atoms and (start[0] != 0 or start[-1] != self.n_atoms):
                raise ValueError(
                    "system_idx must be sorted with values in [0, n_systems) "
                    f"(n_systems={self.n_systems})"
                )
            self._counts = np.diff(start)
        return self._counts


def state_system_index(state) -> SystemIndex:
    """The SystemIndex of a state, kept on the state and rebuilt when its
    ``system_idx`` tensor changes (the layout is validated once)."""
    si = getattr(state, "_tsx_sysidx", None)
    if si is None or not si.matches(state.system_idx, state.n_systems):
        si = SystemIndex(state.system_idx, state.n_systems)
        si.counts  # validates the layout once
        object.__setattr__(state, "_tsx_sysidx", si)
    return si


def jbits_for(max_atoms_per_system: int) -> int:
    """Bits for a system-local atom index."""
    return max(1, int(math.ceil(math.log2(max(2, max_atoms_per_system)))))


def plan_capacity(
    cell_col: np.ndarray,
    pbc: np.ndarray,
    counts: np.ndarray,
    r_list: float,
    bins_per_rlist: int,
) -> tuple[int, int]:
    """Host estimate of (bin count, per-atom neighbor capacity).

    Mirrors the device grid rule nb_k = max(1, floor(q w_k / r_list)) on
    periodic axes; the neighbor estimate is 1.3 rho (4/3) pi r^3 + 24 for
    periodic systems and n_s for systems without periodicity.
    """
    n_sys = cell_col.shape[0]
    pbc = np.broadcast_to(pbc.reshape(-1, 3), (n_sys, 3))
    total_bins = 0
    kmax = 1
    for s in range(n_sys):
        lrow = cell_col[s].T
        det = np.linalg.det(lrow)
        active = bool(pbc[s].any()) and np.isfinite(det) and abs(det) > 1e-12
        nb = [1, 1, 1]
        if active:
            inv = np.linalg.inv(lrow)
            w = 1.0 / np.linalg.norm(inv, axis=0)
            for k in range(3):
                if pbc[s, k]:
                    nb[k] = int(min(1024, max(1, math.floor(bins_per_rlist * w[k] / r_list))))
            rho = counts[s] / abs(det)
            k_est = 1.3 * rho * 4.0 / 3.0 * math.pi * r_list**3 + 24
            if not pbc[s].all():
                k_est = max(k_est, float(counts[s]))
        else:
            k_est = float(counts[s])
        total_bins += nb[0] * nb[1] * nb[2]
        kmax = max(kmax, int(math.ceil(k_est)))
    return total_bins, kmax


class NeighborList:
    """Device neighbor list (ELL rows in bin order) with overflow handling."""

    def __init__(
        self,
        r_list: float,
        *,
        half: bool = False,
        self_interaction: bool = False,
        tpa: int = 4,
        bins_per_rlist: int = 1,
        fill_mode: int = 0,
        word_mode: int = 0,
        one_block_build: bool = True,
    ):
        self.r_list = float(r_list)
        self.fill_mode = int(fill_mode)  # TSX_FILL_AUTO / _BIN / _ATOM / _WAVE / _WAVE_SPARSE / _WAVE_DENSE
        self.word_mode = int(word_mode)  # TSX_WORDS_SHIFT / _STENCIL
        # batches of systems of <= 2048 atoms: bin / scan / sort by one
        # workgroup per system (False: the grid-wide kernels)
        self.one_block_build = bool(one_block_build)
        self.half = bool(half)
        self.self_interaction = bool(self_interaction)
        self.tpa = int(tpa)
        self.bins_per_rlist = int(bins_per_rlist)
        self.kmax = 0
        self.bin_cap = 0
        self.jbits = 1
        self.n_atoms = -1
        self.n_systems = -1
        self.nbr = self.count = self.order = self.ws = self.status = None
        self.generation = 0  # bumped whenever the buffers are reallocated

    # ------------------------------------------------------------ sizing
    def ensure(self, n_atoms: int, n_systems: int, kmax: int, bin_cap: int, jbits: int,
               device: torch.device) -> None:
        """(Re)allocate buffers for at least the given capacities."""
        kmax = max(1, kmax)
        kmax = ((kmax + self.tpa - 1) // self.tpa) * self.tpa
        realloc = (
            self.nbr is None
            or n_atoms != self.n_atoms
            or n_systems != self.n_systems
            or kmax > self.kmax
            or bin_cap > self.bin_cap
        )
        self.jbits = jbits
        if not realloc:
            return
        self.generation += 1
        self.kmax = max(kmax, self.kmax if n_atoms == self.n_atoms else 0)
        self.bin_cap = max(bin_cap, self.bin_cap if n_systems == self.n_systems else 0)
        self.n_atoms, self.n_systems = n_atoms, n_systems
        rows = max(1, n_atoms)
        self.nbr = torch.empty(
            (self.kmax // self.tpa) * rows * self.tpa, dtype=torch.int32, device=device
        )
        self.count = torch.zeros(rows, dtype=torch.int32, device=device)
        self.order = torch.zeros(rows, dtype=torch.int32, device=device)
        wsb = int(_lib.lib().tsx_nl_ws_bytes(n_atoms, n_systems, self.bin_cap))
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)

    @property
    def ld(self) -> int:
        return max(1, self.n_atoms)

    # ------------------------------------------------------------ building
    def enqueue(
        self,
        pos: torch.Tensor,
        pos_stride: int,
        cell_col: torch.Tensor,
        pbc_u8: torch.Tensor,
        pbc_per_system: bool,
        sysidx: SystemIndex,
        *,
        setup: bool = True,
        gate: torch.Tensor | None = None,
        xref: torch.Tensor | None = None,
        counter: torch.Tensor | None = None,
        cell_build: torch.Tensor | None = None,
        frame: _lib.NLFrame | None = None,
    ) -> None:
        """Enqueue a build on the current stream (capturable, no host sync).

        ``gate`` [S] int32 restricts the rebuild to flagged systems; ``setup``
        recomputes the per-system geometry (needed whenever a cell changed);
        ``frame`` receives the bin-relative frame of a stencil-word list."""
        call(
            "tsx_nl_build",
            ptr(pos), pos_stride, ptr(cell_col), ptr(pbc_u8), int(pbc_per_system),
            ptr(sysidx.sys_start), ptr(sysidx.atom_sys), sysidx.n_atoms, sysidx.n_systems,
            self.r_list, self.bins_per_rlist, int(self.self_interaction), int(self.half),
            self.jbits, self.tpa, self.kmax, self.ld, self.bin_cap, int(setup), ptr(self.ws),
            ptr(self.nbr), ptr(self.count), ptr(self.order), ptr(xref), ptr(self.status),
            ptr(gate), ptr(counter), ptr(cell_build), self.fill_mode, self.word_mode,
            None if frame is None else ctypes.addressof(frame),
            sysidx.max_atoms if self.one_block_build else 0, stream(),
        )

    def build(
        self,
        pos: torch.Tensor,
        cell_col: torch.Tensor,
        pbc: torch.Tensor,
        sysidx: SystemIndex,
        pos_stride: int = 3,
        max_retries: int = 12,
        xref: torch.Tensor | None = None,
        cell_build: torch.Tensor | None = None,
        frame: _lib.NLFrame | None = None,
    ) -> None:
        """Build with host-side capacity planning and overflow retries."""
        _lib.require_gpu(pos, "positions")
        pbc_per_system = pbc.ndim == 2
        pbc_u8 = pbc.to(torch.uint8).contiguous()
        counts = sysidx.counts
        jbits = jbits_for(int(counts.max()) if counts.size else 1)
        if self.nbr is None or sysidx.n_atoms != self.n_atoms or sysidx.n_systems != self.n_systems:
            bins, kmax = plan_capacity(
                cell_col.detach().cpu().numpy(),
                pbc.detach().cpu().numpy(),
                counts,
                self.r_list,
                self.bins_per_rlist,
            )
            self.ensure(sysidx.n_atoms, sysidx.n_systems, kmax, bins + 64, jbits, pos.device)
        self.jbits = jbits
        for _ in range(max_retries):
            self.status.zero_()
            self.enqueue(pos, pos_stride, cell_col, pbc_u8, pbc_per_system, sysidx, xref=xref,
                         cell_build=cell_build, frame=frame)
            st = int(self.status.item())
            if st & STATUS_SINGULAR_CELL:
                raise _lib.TsxError("periodic system with a singular cell")
            if st & STATUS_BAD_SIZE:
                raise _lib.TsxError("a system has more atoms than max_system_atoms")
            if st & STATUS_SHIFT_OVERFLOW:
                raise _lib.TsxError(
                    "image shift outside the packed range: positions are too far "
                    "outside their cells; wrap them first"
                )
            if not st & (STATUS_NL_OVERFLOW | STATUS_BIN_OVERFLOW):
                return
            kmax = self.kmax * 2 if st & STATUS_NL_OVERFLOW else self.kmax
            bins = self.bin_cap * 4 if st & STATUS_BIN_OVERFLOW else self.bin_cap
            self.ensure(sysidx.n_atoms, sysidx.n_systems, kmax, bins, jbits, pos.device)
        raise _lib.TsxError("neighbor list capacity could not be resolved")

    # ----------------------------------------------------------- exporting
    def pairs(self, sysidx: SystemIndex, dtype: torch.dtype):
        """Reference triple (mapping, system_mapping, shifts) from the list."""
        dev = self.count.device
        n = sysidx.n_atoms
        if n == 0:
            return (
                torch.zeros((2, 0), dtype=torch.long, device=dev),
                torch.zeros(0, dtype=torch.long, device=dev),
                torch.zeros((0, 3), dtype=dtype, device=dev),
            )
        row_ptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
        ws = torch.empty(int(_lib.lib().tsx_scan_ws_bytes(n)), dtype=torch.uint8, device=dev)
        call("tsx_nl_count_pairs", ptr(self.count), n, ptr(row_ptr), ptr(ws), stream())
        n_pairs = int(row_ptr[n].item())
        mapping = torch.empty((2, n_pairs), dtype=torch.int64, device=dev)
        sysmap = torch.empty(n_pairs, dtype=torch.int64, device=dev)
        shifts = torch.empty((n_pairs, 3), dtype=torch.float64, device=dev)
        if n_pairs:
            call(
                "tsx_nl_emit_pairs",
                ptr(self.nbr), ptr(self.count), ptr(self.order), ptr(row_ptr),
                ptr(sysidx.sys_start), ptr(sysidx.atom_sys), n, self.jbits, self.tpa, self.ld,
                ptr(mapping), n_pairs, ptr(sysmap), ptr(shifts), stream(),
            )
        return mapping, sysmap, shifts.to(dtype)


class VerletList:
    """A neighbor list with skin kept across calls.

    ``update`` rebuilds only the systems whose atoms, pulled back to the cell
    of their last build, moved out of the skin or whose cell strained past
    what the skin absorbs (``tsx_nl_check`` then a gated ``tsx_nl_build``);
    every pair is still tested against the cutoff in the force kernels, so
    results equal a list rebuilt at every call (the reference behaviour,
    neighbors/__init__.py:46-98).  One host read of the build status per
    update (capacity overflow falls back to a full build with more room).
    """

    def __init__(self, cutoff: float, skin: float, *, tpa: int = 4, bins_per_rlist: int = 2,
                 self_interaction: bool = False):
        if skin <= 0:
            raise ValueError("skin must be positive")
        self.cutoff = float(cutoff)
        self.skin = float(skin)
        self.nl = NeighborList(self.cutoff + self.skin, self_interaction=self_interaction,
                               tpa=tpa, bins_per_rlist=bins_per_rlist)
        self._key = None
        self.xref = self.cell_build = self.flags = self.ws = None
        self.full_builds = 0
        self.updates = 0
        self._gen = 0

    def _full(self, pos, cell_col, pbc, sysidx, pos_stride):
        dev = pos.device
        self.xref = torch.zeros((max(1, sysidx.n_atoms), 4), dtype=torch.float64, device=dev)
        self.cell_build = torch.zeros((sysidx.n_systems, 3, 3), dtype=torch.float64, device=dev)
        self.flags = torch.zeros(sysidx.n_systems, dtype=torch.int32, device=dev)
        nbytes = int(_lib.lib().tsx_nl_check_ws_bytes(sysidx.n_systems))
        self.ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=dev)
        self.nl.build(pos, cell_col, pbc, sysidx, pos_stride, xref=self.xref,
                      cell_build=self.cell_build)
        self._key = sysidx.serial
        self.full_builds += 1
        self._gen += 1

    @property
    def generation(self) -> tuple:
        """Changes whenever a buffer a captured graph points to is replaced."""
        return (self._gen, self.nl.generation)

    def ready(self, sysidx: SystemIndex) -> bool:
        return self.xref is not None and self._key == sysidx.serial

    def rebuild_all(self, pos: torch.Tensor, cell_col: torch.Tensor, pbc: torch.Tensor,
                    sysidx: SystemIndex, pos_stride: int = 3) -> NeighborList:
        """Host-checked build of every system at ``pos`` (the list state then
        depends on ``pos`` only), keeping the buffers when they fit."""
        if not self.ready(sysidx):
            self._full(pos, cell_col, pbc, sysidx, pos_stride)
        else:
            self.nl.build(pos, cell_col, pbc, sysidx, pos_stride, xref=self.xref,
                          cell_build=self.cell_build)
        return self.nl

    def enqueue_update(self, pos: torch.Tensor, cell_col: torch.Tensor, pbc_u8: torch.Tensor,
                       pbc_per_system: bool, sysidx: SystemIndex, pos_stride: int = 3,
                       setup: bool = True) -> None:
        """The skin check and gated rebuild of :meth:`update` without the host
        read (capturable; the list must be ready; ``setup=False`` keeps the
        per-system grids of the last setup build, for cells that do not change).  Capacity problems
        accumulate in ``nl.status``: the caller checks it after a run of
        updates and, on overflow, calls :meth:`recover` and redoes the run."""
        if sysidx.n_atoms == 0:
            return
        call("tsx_nl_check", ptr(pos), pos_stride, ptr(cell_col), ptr(self.cell_build),
             ptr(self.xref), ptr(sysidx.atom_sys), sysidx.n_atoms, sysidx.n_systems,
             self.nl.r_list, self.cutoff, ptr(self.flags), ptr(self.ws), stream())
        self.nl.enqueue(pos, pos_stride, cell_col, pbc_u8, pbc_per_system, sysidx, setup=setup,
                        gate=self.flags, xref=self.xref, cell_build=self.cell_build)

    def recover(self, status: int, pos, cell_col, pbc, sysidx, pos_stride: int = 3) -> None:
        """After an overflow seen in ``status``: more room, then a full build
        at ``pos`` (raises for a singular cell or an oversized system)."""
        if status & STATUS_SINGULAR_CELL:
            raise _lib.TsxError("periodic system with a singular cell")
        if status & STATUS_BAD_SIZE:
            raise _lib.TsxError("a system has more atoms than max_system_atoms")
        kmax = self.nl.kmax * 2 if status & STATUS_NL_OVERFLOW else self.nl.kmax
        bins = self.nl.bin_cap * 4 if status & STATUS_BIN_OVERFLOW else self.nl.bin_cap
        self.nl.ensure(sysidx.n_atoms, sysidx.n_systems, kmax, bins, self.nl.jbits, pos.device)
        self._full(pos, cell_col, pbc, sysidx, pos_stride)

    def update(self, pos: torch.Tensor, cell_col: torch.Tensor, pbc: torch.Tensor,
               sysidx: SystemIndex, pos_stride: int = 3,
               defer_status: bool = False) -> NeighborList:
        """Make the list valid for ``pos`` / ``cell_col``; returns it.

        With ``defer_status`` the skin check and the gated rebuild are only
        queued: the caller queues its work on the returned list, then calls
        :meth:`settle`, which reads the build status once (one host
        synchronisation per call instead of one between the build and the
        force kernels) and says whether that work must be redone on a list
        rebuilt with more room.
        """
        self.updates += 1
        self._pending = None
        if self._key != sysidx.serial or self.xref is None:
            self._full(pos, cell_col, pbc, sysidx, pos_stride)
            return self.nl
        if sysidx.n_atoms == 0:
            return self.nl
        call("tsx_nl_check", ptr(pos), pos_stride, ptr(cell_col), ptr(self.cell_build),
             ptr(self.xref), ptr(sysidx.atom_sys), sysidx.n_atoms, sysidx.n_systems,
             self.nl.r_list, self.cutoff, ptr(self.flags), ptr(self.ws), stream())
        self.nl.status.zero_()
        self.nl.enqueue(pos, pos_stride, cell_col, pbc.to(torch.uint8).contiguous(), pbc.ndim == 2,
                        sysidx, setup=True, gate=self.flags, xref=self.xref,
                        cell_build=self.cell_build)
        if defer_status:
            self._pending = (pos, cell_col, pbc, sysidx, pos_stride)
            return self.nl
        self._settle(pos, cell_col, pbc, sysidx, pos_stride)
        return self.nl

    def settle(self) -> bool:
        """Read the status of a deferred :meth:`update`.  Returns True when
        the list had to be rebuilt (every system, with more room), so that
        results computed from it since the update must be computed again."""
        pending, self._pending = getattr(self, "_pending", None), None
        return False if pending is None else self._settle(*pending)

    def _settle(self, pos, cell_col, pbc, sysidx, pos_stride) -> bool:
        st = int(self.nl.status.item())
        if st & STATUS_SINGULAR_CELL:
            raise _lib.TsxError("periodic system with a singular cell")
        if st & (STATUS_NL_OVERFLOW | STATUS_BIN_OVERFLOW | STATUS_SHIFT_OVERFLOW):
            # more room (or wrapped images) needed: full build of every system
            if st & STATUS_NL_OVERFLOW:
                self.nl.ensure(sysidx.n_atoms, sysidx.n_systems, self.nl.kmax * 2,
                               self.nl.bin_cap, self.nl.jbits, pos.device)
            if st & STATUS_BIN_OVERFLOW:
                self.nl.ensure(sysidx.n_atoms, sysidx.n_systems, self.nl.kmax,
                               self.nl.bin_cap * 4, self.nl.jbits, pos.device)
            self._full(pos, cell_col, pbc, sysidx, pos_stride)
            return True
        return False


def segment_index(src_starts: torch.Tensor, counts: torch.Tensor, total: int) -> torch.Tensor:
    """The ranges [src_starts[k], src_starts[k] + counts[k]) concatenated
    (device int64; ``total`` = sum of counts, given by the caller so that
    nothing synchronises)."""
    dst_starts = torch.cumsum(counts, 0) - counts
    return torch.arange(total, device=counts.device) + torch.repeat_interleave(
        src_starts - dst_starts, counts, output_size=total)


def assemble_verlet(dst: VerletList, parts, sysidx: SystemIndex) -> None:
    """Make ``dst`` the Verlet list of the layout ``sysidx`` from rows of
    other Verlet lists, without a build.

    ``parts`` is a sequence of ``(src, src_sysidx, sys_ids, n_atoms)``: the
    systems ``sys_ids`` (device int64, in order) of a list ``src`` valid for
    the layout ``src_sysidx``, holding ``n_atoms`` atoms together; the new
    layout is the concatenation of the parts' systems.  A system's rows are
    the contiguous block of its atom range and its list words index atoms
    within the system, so the rows move with block gathers; only the row
    order entries shift to the new atom offsets.  The displacement
    references (positions and cells at the last build) move with their
    systems, so every system keeps its rebuild schedule.  All sources need
    the same word format (jbits, lanes per atom, list radius)."""
    srcs = [p[0] for p in parts]
    nl0 = srcs[0].nl
    for v in srcs:
        if (v.nl.jbits, v.nl.tpa, v.nl.r_list) != (nl0.jbits, nl0.tpa, nl0.r_list):
            raise ValueError("assembled Verlet lists need one word format")
    n, s = sysidx.n_atoms, sysidx.n_systems
    dev = sysidx.atom_sys.device
    kmax = max(v.nl.kmax for v in srcs)
    bin_cap = max(v.nl.bin_cap for v in srcs)
    nl = dst.nl
    nl.n_atoms = -1  # reallocate for the new layout
    nl.ensure(n, s, kmax, bin_cap, nl0.jbits, dev)
    tpa = nl.tpa
    out = nl.nbr.view(nl.kmax // tpa, nl.ld, tpa)
    dst.xref = torch.zeros((max(1, n), 4), dtype=torch.float64, device=dev)
    dst.cell_build = torch.empty((s, 3, 3), dtype=torch.float64, device=dev)
    dst.flags = torch.zeros(s, dtype=torch.int32, device=dev)
    nbytes = int(_lib.lib().tsx_nl_check_ws_bytes(s))
    dst.ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=dev)
    a0 = s0 = 0
    for src, ssi, ids, n_part in parts:
        k = ids.numel()
        if k == 0:
            continue
        cnt = (ssi.sys_start[ids + 1] - ssi.sys_start[ids]).to(torch.int64)
        old_start = ssi.sys_start[ids].to(torch.int64)
        rows = segment_index(old_start, cnt, n_part)
        new_start = sysidx.sys_start[s0:s0 + k].to(torch.int64)
        shift = torch.repeat_interleave((new_start - old_start).to(torch.int32), cnt,
                                        output_size=n_part)
        sn = src.nl
        kg = sn.kmax // tpa
        src_nbr = sn.nbr.view(kg, sn.ld, tpa)
        out[:kg, a0:a0 + n_part, :] = src_nbr.index_select(1, rows)
        nl.count[a0:a0 + n_part] = sn.count.index_select(0, rows)
        nl.order[a0:a0 + n_part] = sn.order.index_select(0, rows) + shift
        dst.xref[a0:a0 + n_part] = src.xref.index_select(0, rows)
        dst.cell_build[s0:s0 + k] = src.cell_build.index_select(0, ids)
        a0 += n_part
        s0 += k
    nl.status.zero_()
    dst._key = sysidx.serial
    dst._gen += 1


def _normalize(cell: torch.Tensor, pbc: torch.Tensor, n_systems: int):
    """Cell -> [S,3,3], pbc -> [S,3] (torch_sim/neighbors/utils.py:7-34)."""
    if cell.ndim == 2:
        cell = cell.unsqueeze(0).expand(n_systems, -1, -1) if cell.shape[0] == 3 else cell.reshape(n_systems, 3, 3)
    if pbc.ndim == 1:
        pbc = pbc.unsqueeze(0).expand(n_systems, -1) if pbc.shape[0] == 3 else pbc.reshape(n_systems, 3)
    return cell, pbc


def torchsim_nl(
    positions: torch.Tensor,
    cell: torch.Tensor,
    pbc: torch.Tensor,
    cutoff: float,
    system_idx: torch.Tensor,
    self_interaction: bool = False,  # noqa: FBT001, FBT002
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Full neighbor list with the reference contract (neighbors/__init__.py:46-98).

    ``cell`` uses the row-vector convention of the reference NL functions.
    """
    _lib.require_gpu(positions, "positions")
    dtype = positions.dtype
    if positions.shape[0] == 0:
        dev = positions.device
        return (
            torch.zeros((2, 0), dtype=torch.long, device=dev),
            torch.zeros(0, dtype=torch.long, device=dev),
            torch.zeros((0, 3), dtype=dtype, device=dev),
        )
    n_systems = int(system_idx.max().item()) + 1
    cell, pbc = _normalize(cell, pbc, n_systems)
    cell_col = cell.to(torch.float64).transpose(1, 2).contiguous()
    pos = positions.to(torch.float64).contiguous()
    sysidx = SystemIndex(system_idx, n_systems)
    nl = NeighborList(cutoff, self_interaction=self_interaction, tpa=1)
    nl.build(pos, cell_col, pbc.contiguous(), sysidx)
    return nl.pairs(sysidx, dtype)


# The reference exposes the same contract under several names; on this
# backend all of them are the HIP cell-list search.
torch_nl_linked_cell = torchsim_nl
torch_nl_n2 = torchsim_nl
default_batched_nl = torchsim_nl


def wrap_positions(
    positions: torch.Tensor, cell_row: torch.Tensor, system_idx: torch.Tensor, pbc: torch.Tensor
) -> tuple[torch.Tensor, torch.Tensor]:
    """Wrapped positions and integer shifts (transforms.py:160-210) on the GPU."""
    _lib.require_gpu(positions, "positions")
    n_systems = cell_row.shape[0]
    if pbc.ndim == 1:
        pbc = pbc.unsqueeze(0).expand(n_systems, -1)
    dev = positions.device
    cell_col = cell_row.to(torch.float64).transpose(1, 2).contiguous()
    pos = positions.to(torch.float64).contiguous()
    atom_sys = system_idx.to(torch.int32).contiguous()
    ws = torch.empty(int(_lib.lib().tsx_wrap_ws_bytes(n_systems)), dtype=torch.uint8, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    wrapped = torch.empty_like(pos)
    shifts = torch.empty_like(pos)
    call(
        "tsx_wrap_positions", ptr(pos), ptr(cell_col), ptr(pbc.to(torch.uint8).contiguous()), 1,
        ptr(atom_sys), pos.shape[0], n_systems, ptr(ws), ptr(status), ptr(wrapped), ptr(shifts),
        stream(),
    )
    return wrapped.to(positions.dtype), shifts.to(positions.dtype)
