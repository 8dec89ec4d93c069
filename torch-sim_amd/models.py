"""Pair-potential models behind the reference ModelInterface surface.

``forward(state) -> {"energy" [S], "forces" [N,3], "stress" [S,3,3],
"energies" [N], "stresses" [N,3,3]}`` as torch_sim/models/interface.py:169-201
and pair_potential.py:351-464.  ``PairPotentialModel(pair_fn, *, cutoff,
...)`` takes the reference's constructor (pair_potential.py:300-345); the
convenience models build their pair function with ``functools.partial``
exactly as LennardJonesModel (lennard_jones.py:92-137), MorseModel
(morse.py:96-144), SoftSphereModel (soft_sphere.py:192-246) and
SoftSphereMultiModel (soft_sphere.py:265-377) do.

For the closed forms the pair function is not called during a forward
pass: the model recognises ``lennard_jones_pair``, ``morse_pair``,
``soft_sphere_pair`` bound with ``functools.partial`` and the species-pair
``MultiSoftSpherePairFn`` / ``MultiLennardJonesPairFn``, and hands their
parameters to the HIP pair kernel (FP32 pair arithmetic, FP64
accumulation).  Any other callable (a user potential such as the
Born-Mayer-Huggins-Tosi-Fumi example of pair_potential.py:35-61) runs on the
generic path: the pairs come from the same HIP cell list, in the
reference's (mapping, system_mapping, shifts) form, and the pair function,
its autograd derivative and the scatters are torch operations on the GPU,
in the operation order of pair_potential.py:351-464.

``PairForcesModel`` (pair_potential.py:467-612) takes a force function
instead of an energy; ``particle_life_pair_force`` (particle_life.py:24-56)
is recognised and runs in the HIP kernel, any other force function on the
generic path.  ``SumModel`` / ``SerialSumModel`` (interface.py:204-351) add
the outputs of several models.

Each call rebuilds the neighbor list (as the reference does,
pair_potential.py:168-174) with the HIP cell list, or keeps a Verlet list
with ``skin``.  The kernel walks full lists (each pair from both ends, no
atomics); the result equals the reference's half-list sum, so
``reduce_to_half_list`` changes nothing but is accepted.

``retain_graph=True`` (pair_potential.py:419-425, 461-462): ``energy`` and
``forces`` come back attached to autograd through the HIP backward kernels
(positions, cell and the pair function's parameter tensors); the forces'
graph includes the second-derivative term that the reference's
``create_graph=False`` leaves out.  Per-atom outputs and stress stay
detached.

Reference quirk kept: per-atom ``energies`` give every atom the full energy
of each of its pairs (pair_potential.py:411-417), so they sum to twice the
system energy.
"""

from __future__ import annotations

import functools
import inspect

import torch

from . import _lib
from ._lib import call, ptr, stream
from .neighbors import (
    NeighborList,
    SystemIndex,
    TensorKey,
    VerletList,
    default_batched_nl,
    torch_nl_linked_cell,
    torch_nl_n2,
    torchsim_nl,
)

LJ, MORSE, SOFT_SPHERE, PARTICLE_LIFE = 0, 1, 2, 3
GENERIC = -1  # a user callable: torch operations over the HIP list's pairs

# every name the reference gives its batched neighbor list; on this backend
# they are all the HIP cell-list search
_NL_FNS = (torchsim_nl, torch_nl_linked_cell, torch_nl_n2, default_batched_nl)


# ------------------------------------------------------------ pair functions
def lennard_jones_pair(dr, zi, zj, sigma=1.0, epsilon=1.0):  # noqa: ARG001
    """V = 4 eps [(s/r)^12 - (s/r)^6], 0 at r <= 0 (lennard_jones.py:23-47)."""
    idr6 = (sigma / dr).pow(6)
    energy = 4.0 * epsilon * (idr6 * idr6 - idr6)
    return torch.where(dr > 0, energy, torch.zeros_like(energy))


def morse_pair(dr, zi, zj, sigma=1.0, epsilon=5.0, alpha=5.0):  # noqa: ARG001
    """V = eps (1 - exp(-a (r - s)))^2 - eps, 0 at r <= 0 (morse.py:23-47)."""
    energy = epsilon * (1.0 - torch.exp(-alpha * (dr - sigma))).pow(2) - epsilon
    return torch.where(dr > 0, energy, torch.zeros_like(energy))


def soft_sphere_pair(dr, zi, zj, sigma=1.0, epsilon=1.0, alpha=2.0):  # noqa: ARG001
    """V = eps/a (1 - r/s)^a for r < s, else 0 (soft_sphere.py:35-59)."""
    energy = epsilon / alpha * (1.0 - dr / sigma).pow(alpha)
    return torch.where(dr < sigma, energy, torch.zeros_like(energy))


def particle_life_pair_force(dr, zi, zj, A=1.0, beta=0.3, sigma=1.0):  # noqa: ARG001, N803
    """Particle-life force magnitude (particle_life.py:24-56): r/beta - 1 for
    r < beta, A (1 - |2r - 1 - beta| / (1 - beta)) for beta <= r < sigma,
    else 0.  A force, not an energy: for PairForcesModel."""
    zero = torch.zeros_like(dr)
    inner = torch.where(dr < beta, dr / beta - 1.0, zero)
    ramp = A * (1.0 - torch.abs(2.0 * dr - 1.0 - beta) / (1.0 - beta))
    return inner + torch.where((dr >= beta) & (dr < sigma), ramp, zero)


class _SpeciesPairFn(torch.nn.Module):
    """Per-species-pair parameter matrices indexed through the atomic numbers
    (row k belongs to ``atomic_numbers[k]``, soft_sphere.py:132-140)."""

    _names = ("sigma_matrix", "epsilon_matrix", "alpha_matrix")

    def __init__(self, atomic_numbers, sigma_matrix, epsilon_matrix, alpha_matrix=None,
                 alpha_default=2.0):
        super().__init__()
        atomic_numbers = torch.as_tensor(atomic_numbers)
        n = len(atomic_numbers)
        if sigma_matrix.shape != (n, n):
            raise ValueError(f"sigma_matrix must have shape ({n}, {n})")
        if epsilon_matrix.shape != (n, n):
            raise ValueError(f"epsilon_matrix must have shape ({n}, {n})")
        if alpha_matrix is not None and alpha_matrix.shape != (n, n):
            raise ValueError(f"alpha_matrix must have shape ({n}, {n})")
        self.register_buffer("atomic_numbers", atomic_numbers)
        self.sigma_matrix = sigma_matrix
        self.epsilon_matrix = epsilon_matrix
        self.alpha_matrix = (alpha_matrix if alpha_matrix is not None
                             else torch.full((n, n), alpha_default, dtype=sigma_matrix.dtype,
                                             device=sigma_matrix.device))
        z_to_idx = torch.full((int(atomic_numbers.max().item()) + 1,), -1, dtype=torch.long)
        for idx, z in enumerate(atomic_numbers.tolist()):
            z_to_idx[int(z)] = idx
        self.register_buffer("z_to_idx", z_to_idx)

    def _lookup(self, zi, zj):
        a, b = self.z_to_idx[zi], self.z_to_idx[zj]
        return self.sigma_matrix[a, b], self.epsilon_matrix[a, b], self.alpha_matrix[a, b]


class MultiSoftSpherePairFn(_SpeciesPairFn):
    """Species-pair soft sphere (soft_sphere.py:90-167)."""

    def forward(self, dr, zi, zj):
        sigma, epsilon, alpha = self._lookup(zi, zj)
        energy = epsilon / alpha * (1.0 - dr / sigma).pow(alpha)
        return torch.where(dr < sigma, energy, torch.zeros_like(energy))


class MultiLennardJonesPairFn(_SpeciesPairFn):
    """Species-pair LJ 12-6 (sigma, epsilon matrices).  The reference has no
    such class; it is the LJ counterpart of MultiSoftSpherePairFn for binary
    LJ mixtures (alpha is unused)."""

    def __init__(self, atomic_numbers, sigma_matrix, epsilon_matrix):
        super().__init__(atomic_numbers, sigma_matrix, epsilon_matrix,
                         torch.zeros_like(sigma_matrix))

    def forward(self, dr, zi, zj):
        sigma, epsilon, _ = self._lookup(zi, zj)
        idr6 = (sigma / dr).pow(6)
        energy = 4.0 * epsilon * (idr6 * idr6 - idr6)
        return torch.where(dr > 0, energy, torch.zeros_like(energy))


_CLOSED_FORMS = {lennard_jones_pair: LJ, morse_pair: MORSE, soft_sphere_pair: SOFT_SPHERE}


def _partial_parts(fn):
    """(function, keywords) of a keyword-only functools.partial or a bare
    function; (None, {}) when dr / zi / zj are bound positionally."""
    if isinstance(fn, functools.partial):
        if fn.args:
            return None, {}
        return fn.func, dict(fn.keywords)
    return fn, {}


def _bound_values(func, kw, names):
    """The keyword values of ``names`` (defaults from the signature); a
    keyword the function does not take is a TypeError, as the call would be."""
    extra = set(kw) - set(names)
    if extra:
        raise TypeError(f"unknown parameters {sorted(extra)} for {func.__name__}")
    sig = inspect.signature(func).parameters
    return [kw.get(k, sig[k].default) for k in names]


def resolve_pair_fn(pair_fn, *, generic: bool = False):
    """(kind, sigma, epsilon, alpha, species_z) of a recognised pair function.

    Scalars come back as given (floats or 0-d tensors, which may require
    grad); species-pair functions give [n, n] matrices and the atomic numbers
    of their rows.  Another callable is a TypeError, or with ``generic`` the
    tuple (GENERIC, None, None, None, None)."""
    if isinstance(pair_fn, (MultiSoftSpherePairFn, MultiLennardJonesPairFn)):
        kind = SOFT_SPHERE if isinstance(pair_fn, MultiSoftSpherePairFn) else LJ
        return (kind, pair_fn.sigma_matrix, pair_fn.epsilon_matrix, pair_fn.alpha_matrix,
                pair_fn.atomic_numbers)
    func, kw = _partial_parts(pair_fn)
    kind = _CLOSED_FORMS.get(func) if func is not None else None
    if kind is None:
        if generic and callable(pair_fn):
            return GENERIC, None, None, None, None
        raise TypeError(
            "the HIP pair kernels take functools.partial(lennard_jones_pair | morse_pair | "
            "soft_sphere_pair, sigma=..., epsilon=..., alpha=...) and MultiSoftSpherePairFn / "
            f"MultiLennardJonesPairFn; got {pair_fn!r}"
        )
    names = ("sigma", "epsilon") + (("alpha",) if kind != LJ else ())
    vals = _bound_values(func, kw, names)
    return kind, vals[0], vals[1], vals[2] if kind != LJ else 2.0, None


def resolve_force_fn(force_fn):
    """(PARTICLE_LIFE, sigma, A, beta) for a bound particle_life_pair_force,
    else (GENERIC, None, None, None)."""
    func, kw = _partial_parts(force_fn)
    if func is particle_life_pair_force:
        a, beta, sigma = _bound_values(func, kw, ("A", "beta", "sigma"))
        return PARTICLE_LIFE, sigma, a, beta
    if not callable(force_fn):
        raise TypeError(f"force_fn must be callable, got {force_fn!r}")
    return GENERIC, None, None, None


class ModelInterface(torch.nn.Module):
    """Abstract model surface (torch_sim/models/interface.py:71-201)."""

    _device: torch.device
    _dtype: torch.dtype
    _compute_stress: bool
    _compute_forces: bool

    @property
    def device(self) -> torch.device:
        return self._device

    @device.setter
    def device(self, device: torch.device) -> None:
        raise NotImplementedError(
            f"{type(self).__name__} fixes its device at construction; build a new model "
            "for another device"
        )

    @property
    def dtype(self) -> torch.dtype:
        return self._dtype

    @dtype.setter
    def dtype(self, dtype: torch.dtype) -> None:
        raise NotImplementedError(
            f"{type(self).__name__} fixes its dtype at construction; build a new model "
            "for another dtype"
        )

    @property
    def compute_stress(self) -> bool:
        return self._compute_stress

    @compute_stress.setter
    def compute_stress(self, compute_stress: bool) -> None:
        raise NotImplementedError(
            f"{type(self).__name__} fixes compute_stress at construction"
        )

    @property
    def compute_forces(self) -> bool:
        return self._compute_forces

    @compute_forces.setter
    def compute_forces(self, compute_forces: bool) -> None:
        raise NotImplementedError(
            f"{type(self).__name__} fixes compute_forces at construction"
        )

    @property
    def memory_scales_with(self) -> str:
        return getattr(self, "_memory_scales_with", "n_atoms_x_density")

    def forward(self, state, **kwargs) -> dict[str, torch.Tensor]:  # pragma: no cover
        raise NotImplementedError


def _default_device(device) -> torch.device:
    if device is None:
        return torch.device("cuda")
    return torch.device(device)


def _check_dtype(state, dtype) -> None:
    if state.dtype != dtype:
        raise TypeError(
            f"SimState dtype {state.dtype} does not match model dtype {dtype}. "
            f"Either set the model dtype to {state.dtype} or convert the SimState "
            f"to {dtype} using sim_state.to(dtype={dtype})."
        )


def _check_nl_fn(neighbor_list_fn) -> None:
    if neighbor_list_fn is not None and neighbor_list_fn not in _NL_FNS:
        raise NotImplementedError(
            "torch_sim_amd models build their neighbor list with the HIP cell list "
            "(torchsim_nl and its aliases torch_nl_linked_cell, torch_nl_n2, "
            f"default_batched_nl); got {neighbor_list_fn!r}"
        )


class _HipPairModel(ModelInterface):
    """What the pair models share: the parameter table of a closed-form kind
    for the HIP pair kernel, the list(s), and the generic path's pairs."""

    def _setup_pairs(self, kind, params, species_z, *, cutoff, device, dtype, half, tpa,
                     skin) -> None:
        self._device = _default_device(device)
        self._dtype = dtype
        self.kind = int(kind)
        self.cutoff = float(cutoff)
        self.reduce_to_half_list = bool(half)
        self.tpa = tpa
        self._sysidx: SystemIndex | None = None
        self._species_key = None
        self._species = None
        self._z_to_idx = None
        self._generic_nl = None
        self.skin = float(skin)
        self._vl = None
        if self.kind == GENERIC:
            self.n_species = 1
            self._param_src = None
            return
        self._param_src = params
        self._param_versions = None
        sig = self._matrix(params[0])
        self.n_species = n = sig.shape[0]
        if n > 8:
            raise ValueError("at most 8 species are supported")
        self._refresh_params()
        if n > 1:
            if species_z is None:
                raise ValueError("species-pair parameters need the atomic numbers of the rows")
            z = torch.as_tensor(species_z, dtype=torch.long).flatten().cpu()
            table = torch.full((int(z.max()) + 1,), -1, dtype=torch.int32)
            for k, zz in enumerate(z.tolist()):
                table[int(zz)] = k
            self._z_to_idx = table.to(self._device)
        self._nl = NeighborList(self.cutoff, tpa=tpa)
        self._vl = VerletList(self.cutoff, self.skin, tpa=tpa) if self.skin > 0 else None

    @property
    def fused_kernels(self) -> bool:
        """True when the fused MD / FIRE engines can run this model (an
        energy model of a closed-form kind)."""
        return False

    # ------------------------------------------------------------- params
    @staticmethod
    def _matrix(v) -> torch.Tensor:
        t = torch.as_tensor(v).detach().to(torch.float64).cpu()
        return t.view(1, 1) if t.ndim == 0 else t

    def _refresh_params(self) -> None:
        """The [n^2, 3] device table of (sigma, epsilon, alpha), rebuilt when a
        parameter tensor changed in place since the last build."""
        if self._param_src is None:
            return
        versions = tuple(v._version if isinstance(v, torch.Tensor) else None
                         for v in self._param_src)
        if versions == self._param_versions:
            return
        sig, eps, alp = (self._matrix(v) for v in self._param_src)
        n = sig.shape[0]
        eps = eps.expand(n, n) if eps.numel() == 1 else eps
        alp = alp.expand(n, n) if alp.numel() == 1 else alp
        for name, m in (("sigma", sig), ("epsilon", eps), ("alpha", alp)):
            if m.shape != (n, n):
                raise ValueError(f"{name} matrix must have shape ({n}, {n})")
            if not torch.allclose(m, m.T):
                raise ValueError(f"{name}_matrix is not symmetric")
        self._params = torch.stack([sig, eps, alp], dim=-1).reshape(n * n, 3).contiguous().to(
            self._device)
        self._param_versions = versions

    # ------------------------------------------------------------- helpers
    def _system_index(self, state) -> SystemIndex:
        si = self._sysidx
        if si is None or not si.matches(state.system_idx, state.n_systems):
            si = SystemIndex(state.system_idx, state.n_systems)
            self._sysidx = si
        return si

    def species_of(self, atomic_numbers: torch.Tensor) -> torch.Tensor | None:
        """Species row of every atom (int32); every atomic number must be one
        of the model's species (checked on the host: the pair kernels index
        their parameter table with it)."""
        if self.n_species == 1:
            return None
        if self._species_key is not None and self._species_key.matches(atomic_numbers):
            return self._species
        if int(atomic_numbers.max()) >= self._z_to_idx.numel() or int(atomic_numbers.min()) < 0:
            raise ValueError("atomic number outside the model's species table")
        sp = self._z_to_idx[atomic_numbers].contiguous()
        if int(sp.min()) < 0:
            raise ValueError("state contains an atomic number that is not a model species")
        self._species_key, self._species = TensorKey(atomic_numbers), sp
        return sp

    def _hip_pass(self, state, *, forces_needed: bool, stress_needed: bool,
                  per_atom_stresses: bool):
        """The HIP pair kernel and the per-system reduction: (energy [S],
        e_atom [N] (half of each pair's energy per end), forces, stress,
        stresses), FP64."""
        _lib.require_gpu(state.positions, "state.positions")
        self._refresh_params()
        pos = state.positions.detach().to(torch.float64).contiguous()
        cell = state.cell.detach().to(torch.float64).contiguous()
        n, s = state.n_atoms, state.n_systems
        si = self._system_index(state)
        pbc = state.pbc
        if self._vl is not None:
            # the build status is read once, after the force kernels are queued
            nl = self._vl.update(pos, cell, pbc, si, defer_status=True)
        else:
            nl = self._nl
            nl.build(pos, cell, pbc, si)
        dev = pos.device
        forces = torch.empty((n, 3), dtype=torch.float64, device=dev)
        e_atom = torch.empty(n, dtype=torch.float64, device=dev)
        w_atom = torch.empty((n, 6), dtype=torch.float64, device=dev) if stress_needed else None
        species = self.species_of(state.atomic_numbers)

        def pair_forces(nl):
            call(
                "tsx_pair_forces",
                self.kind, ptr(self._params), self.n_species, ptr(species), ptr(pos), 3,
                ptr(cell), ptr(si.sys_start), ptr(si.atom_sys), ptr(nl.nbr), ptr(nl.count),
                ptr(nl.order), n, nl.jbits, nl.tpa, nl.ld, self.cutoff, 0, ptr(forces),
                ptr(e_atom), ptr(w_atom), stream(),
            )

        if n:
            pair_forces(nl)
            if self._vl is not None and self._vl.settle():
                pair_forces(self._vl.nl)  # the list was rebuilt with more room
        energy = torch.empty(s, dtype=torch.float64, device=dev)
        stress = torch.empty((s, 3, 3), dtype=torch.float64, device=dev) if stress_needed else None
        stresses = (
            torch.empty((n, 3, 3), dtype=torch.float64, device=dev)
            if stress_needed and per_atom_stresses
            else None
        )
        call(
            "tsx_system_reduce", ptr(e_atom), ptr(w_atom), ptr(cell), ptr(si.sys_start),
            si.max_atoms, s, ptr(energy), ptr(stress), ptr(stresses), ptr(si.reduce_ws),
            stream(),
        )
        return energy, e_atom, forces, stress, stresses

    def _generic_pairs(self, state):
        """The generic path's pairs (pair_potential.py:130-197): the HIP cell
        list at the cutoff (half with reduce_to_half_list) in the reference's
        triple form; dr = r_j - r_i + S . L_row keeps the autograd graph of
        the positions and the cell."""
        _lib.require_gpu(state.positions, "state.positions")
        if self._generic_nl is None:
            self._generic_nl = NeighborList(self.cutoff, half=self.reduce_to_half_list, tpa=1)
        si = self._system_index(state)
        pos = state.positions
        self._generic_nl.build(pos.detach().to(torch.float64).contiguous(),
                               state.cell.detach().to(torch.float64).contiguous(), state.pbc, si)
        mapping, sysmap, shifts = self._generic_nl.pairs(si, pos.dtype)
        row_cell = state.cell.mT
        cart = (shifts.unsqueeze(-1) * row_cell[sysmap]).sum(dim=1)
        dr = pos[mapping[1]] - pos[mapping[0]] + cart
        return mapping, sysmap, dr, dr.norm(dim=1), row_cell

    def _generic_stress(self, state, mapping, sysmap, dr, fvec, row_cell, per_atom: bool):
        """Virial stress (pair_potential.py:200-274): -(1 / V) sum dr (x) f,
        weight 1 on a half list and 1/2 on a full one; per atom, a quarter
        (full) or half (half list) of each pair to each end."""
        half = self.reduce_to_half_list
        vol = torch.linalg.det(row_cell).abs()
        per_pair = dr.unsqueeze(-1) * fvec.unsqueeze(-2)
        s = state.n_systems
        stress = torch.zeros((s, 3, 3), dtype=dr.dtype, device=dr.device)
        stress = stress.index_add(0, sysmap, -per_pair) / vol[:, None, None]
        out = {"stress": stress * (1.0 if half else 0.5)}
        if per_atom:
            w = 0.5 if half else 0.25
            at = torch.zeros((state.n_atoms, 3, 3), dtype=dr.dtype, device=dr.device)
            at = at.index_add(0, mapping[0], -w * per_pair).index_add(0, mapping[1],
                                                                      -w * per_pair)
            out["stresses"] = at / vol[state.system_idx, None, None]
        return out


class PairPotentialModel(_HipPairModel):
    """Batched pair potential (pair_potential.py:277-464).

    A recognised closed-form ``pair_fn`` (module docstring) runs in the HIP
    pair kernel; its parameters may be floats, tensors or, for species-pair
    functions, symmetric [n, n] matrices.  Any other callable runs on the
    generic path (HIP list, torch pair arithmetic and autograd on the GPU).
    Extra keywords of this backend: ``tpa`` (lanes per atom in the pair
    kernel) and ``skin`` (keep a Verlet list between calls; results equal a
    rebuild at every call); both apply to the closed forms.  ``device``
    defaults to the current GPU (there is no CPU path).
    """

    def __init__(
        self,
        pair_fn,
        *,
        cutoff: float,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        reduce_to_half_list: bool = False,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        kind, sigma, epsilon, alpha, species_z = resolve_pair_fn(pair_fn, generic=True)
        _check_nl_fn(neighbor_list_fn)
        super().__init__()
        self._setup_pairs(kind, (sigma, epsilon, alpha), species_z, cutoff=cutoff,
                          device=device, dtype=dtype, half=reduce_to_half_list, tpa=tpa,
                          skin=skin)
        self._compute_forces = compute_forces
        self._compute_stress = compute_stress
        self.per_atom_energies = per_atom_energies
        self.per_atom_stresses = per_atom_stresses
        self.pair_fn = pair_fn
        self.neighbor_list_fn = torchsim_nl
        self.retain_graph = retain_graph
        self._diff_terms = None

    @property
    def fused_kernels(self) -> bool:
        return self.kind in (LJ, MORSE, SOFT_SPHERE)

    def graph_params(self) -> torch.Tensor:
        """[n, n, 3] (sigma, epsilon, alpha) built from the pair function's own
        values, so tensors among them that require grad stay in the graph."""
        from .autograd import pair_params

        sig, eps, alp = self._param_src
        return pair_params(sig, eps, alp).to(self._device)

    def _attached(self, state, out) -> None:
        """Replace energy / forces by outputs attached to autograd."""
        from . import autograd as tad

        if self._diff_terms is None:
            self._diff_terms = tad.PairTerms(self)
        geo = tad.geometry(self._diff_terms, state)
        params = self.graph_params()
        dt = self._dtype
        out["energy"] = tad.pair_energy(state.positions, state.cell, params, geo).to(dt)
        if self._compute_forces:
            out["forces"] = tad.pair_forces(state.positions, state.cell, params, geo).to(dt)

    def _forward_generic(self, state) -> dict[str, torch.Tensor]:
        """pair_potential.py:351-464 over the HIP list's pairs."""
        half = self.reduce_to_half_list
        mapping, sysmap, dr, dist, row_cell = self._generic_pairs(state)
        z = state.atomic_numbers
        need_grad = self._compute_forces or self._compute_stress
        d = dist.requires_grad_() if need_grad else dist
        e = self.pair_fn(d, z[mapping[0]], z[mapping[1]])
        e = torch.where(dist < self.cutoff, e, torch.zeros_like(e))
        ew = 1.0 if half else 0.5
        dev, dt = dist.device, dist.dtype
        out = {"energy": torch.zeros(state.n_systems, dtype=dt, device=dev).index_add(
            0, sysmap, ew * e)}
        if self.per_atom_energies:
            out["energies"] = torch.zeros(state.n_atoms, dtype=dt, device=dev).index_add(
                0, mapping[0], ew * e).index_add(0, mapping[1], ew * e)
        if need_grad:
            (dvdr,) = torch.autograd.grad(e.sum(), d, retain_graph=self.retain_graph)
            safe = torch.where(dist > 0, dist, torch.ones_like(dist))
            fvec = (-dvdr / safe).unsqueeze(-1) * dr  # on j due to i
            if self._compute_forces:
                f = torch.zeros_like(state.positions).index_add(0, mapping[0], -fvec)
                if half:
                    f = f.index_add(0, mapping[1], fvec)
                out["forces"] = f
            if self._compute_stress:
                out.update(self._generic_stress(state, mapping, sysmap, dr, fvec, row_cell,
                                                self.per_atom_stresses))
        if not self.retain_graph:
            out = {k: v.detach() for k, v in out.items()}
        return out

    def forward(self, state, **_kwargs) -> dict[str, torch.Tensor]:
        _check_dtype(state, self._dtype)
        if self.kind == GENERIC:
            return self._forward_generic(state)
        energy, e_atom, forces, stress, stresses = self._hip_pass(
            state, forces_needed=self._compute_forces, stress_needed=self._compute_stress,
            per_atom_stresses=self.per_atom_stresses)
        dt = self._dtype
        out = {"energy": energy.to(dt)}
        if self.per_atom_energies:
            out["energies"] = (2.0 * e_atom).to(dt)
        if self._compute_forces:
            out["forces"] = forces.to(dt)
        if self._compute_stress:
            out["stress"] = stress.to(dt)
            if stresses is not None:
                out["stresses"] = stresses.to(dt)
        if self.retain_graph and state.n_atoms:
            self._attached(state, out)
        return out


class PairForcesModel(_HipPairModel):
    """Pair model defined by a force magnitude (pair_potential.py:467-612).

    ``force_fn(dr, zi, zj) -> f`` gives the force on j due to i as
    f (r_j - r_i + S.L) / r; energies are zero.  A bound
    ``particle_life_pair_force`` runs in the HIP kernel, any other callable on
    the generic path.  Reference behaviour kept: every list entry applies its
    force and the reaction (pair_potential.py:586-588), so on the default
    full list, where each pair appears twice, a symmetric force function
    gives twice the forces of ``reduce_to_half_list=True``; the stresses are
    weighted as for the potential models and agree between the two.
    """

    def __init__(
        self,
        force_fn,
        *,
        cutoff: float,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        compute_stress: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        reduce_to_half_list: bool = False,
        tpa: int = 4,
    ) -> None:
        kind, sigma, amp, beta = resolve_force_fn(force_fn)
        _check_nl_fn(neighbor_list_fn)
        super().__init__()
        self._setup_pairs(kind, (sigma, amp, beta), None, cutoff=cutoff, device=device,
                          dtype=dtype, half=reduce_to_half_list, tpa=tpa, skin=0.0)
        self._compute_forces = True
        self._compute_stress = compute_stress
        self.per_atom_stresses = per_atom_stresses
        self.force_fn = force_fn
        self.neighbor_list_fn = torchsim_nl

    def _forward_generic(self, state) -> dict[str, torch.Tensor]:
        """pair_potential.py:539-612 over the HIP list's pairs."""
        mapping, sysmap, dr, dist, row_cell = self._generic_pairs(state)
        z = state.atomic_numbers
        f = self.force_fn(dist, z[mapping[0]], z[mapping[1]])
        f = torch.where(dist < self.cutoff, f, torch.zeros_like(f))
        safe = torch.where(dist > 0, dist, torch.ones_like(dist))
        fvec = (f / safe).unsqueeze(-1) * dr
        forces = torch.zeros_like(state.positions).index_add(0, mapping[0], -fvec).index_add(
            0, mapping[1], fvec)
        out = {"energy": torch.zeros(state.n_systems, dtype=dist.dtype, device=dist.device),
               "forces": forces}
        if self._compute_stress:
            out.update(self._generic_stress(state, mapping, sysmap, dr, fvec, row_cell,
                                            self.per_atom_stresses))
        return out

    def forward(self, state, **_kwargs) -> dict[str, torch.Tensor]:
        _check_dtype(state, self._dtype)
        if self.kind == GENERIC:
            return self._forward_generic(state)
        energy, _, forces, stress, stresses = self._hip_pass(
            state, forces_needed=True, stress_needed=self._compute_stress,
            per_atom_stresses=self.per_atom_stresses)
        dt = self._dtype
        # the kernel's row sums give each pair once; a full list applies
        # every pair twice
        if not self.reduce_to_half_list:
            forces = 2.0 * forces
        out = {"energy": energy.to(dt), "forces": forces.to(dt)}
        if self._compute_stress:
            out["stress"] = stress.to(dt)
            if stresses is not None:
                out["stresses"] = stresses.to(dt)
        return out


class ParticleLifeModel(PairForcesModel):
    """PairForcesModel with particle_life_pair_force (particle_life.py:59-122):
    ``epsilon`` is the amplitude A; the cutoff defaults to 2.5 sigma."""

    def __init__(
        self,
        sigma: float = 1.0,
        epsilon: float = 1.0,
        beta: float = 0.3,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        compute_stress: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        cutoff: float | None = None,
        tpa: int = 4,
        **_kwargs,
    ) -> None:
        self.sigma_param = sigma
        self.epsilon = epsilon
        self.beta = beta
        super().__init__(
            force_fn=functools.partial(particle_life_pair_force, A=epsilon, beta=beta,
                                       sigma=sigma),
            cutoff=cutoff if cutoff is not None else 2.5 * sigma,
            device=device,
            dtype=dtype,
            compute_stress=compute_stress,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            reduce_to_half_list=False,
            tpa=tpa,
        )


_CANONICAL_KEYS = frozenset({"energy", "forces", "stress"})


def _accumulate_output(combined: dict, output: dict) -> None:
    """Energy, forces and stress add; any other key is replaced by the later
    model's value (interface.py:56-68)."""
    for key, value in output.items():
        if key in combined and key in _CANONICAL_KEYS:
            combined[key] = combined[key] + value
        else:
            combined[key] = value


class SumModel(ModelInterface):
    """Sum of two or more models on one device and dtype (interface.py:204-322)."""

    def __init__(self, *models) -> None:
        super().__init__()
        if len(models) < 2:
            raise ValueError("SumModel requires at least two child models")
        first = models[0]
        for k, m in enumerate(models[1:], start=1):
            if m.device != first.device:
                raise ValueError(
                    f"Device mismatch: model 0 has {first.device}, model {k} has {m.device}")
            if m.dtype != first.dtype:
                raise ValueError(
                    f"Dtype mismatch: model 0 has {first.dtype}, model {k} has {m.dtype}")
        self.models = torch.nn.ModuleList(models)
        self._device = first.device
        self._dtype = first.dtype
        self._compute_stress = all(m.compute_stress for m in models)
        self._compute_forces = all(m.compute_forces for m in models)

    def _children(self) -> list:
        return list(self.models.children())

    def _propagate(self, name: str, value: bool) -> None:
        for m in self._children():
            try:
                setattr(m, name, value)
            except NotImplementedError:
                if value:
                    raise

    @ModelInterface.compute_stress.setter
    def compute_stress(self, value: bool) -> None:
        """Set on every child that allows it (a child that cannot is an error
        only when switching on)."""
        self._propagate("compute_stress", value)
        self._compute_stress = value

    @ModelInterface.compute_forces.setter
    def compute_forces(self, value: bool) -> None:
        self._propagate("compute_forces", value)
        self._compute_forces = value

    @property
    def retain_graph(self) -> bool:
        return all(getattr(m, "retain_graph", False) for m in self._children())

    @retain_graph.setter
    def retain_graph(self, value: bool) -> None:
        for m in self._children():
            if hasattr(m, "retain_graph"):
                m.retain_graph = value

    @property
    def memory_scales_with(self) -> str:
        order = {"n_atoms": 0, "n_atoms_x_density": 1}
        return max((m.memory_scales_with for m in self._children()),
                   key=lambda v: order.get(v, 2))

    def forward(self, state, **kwargs) -> dict[str, torch.Tensor]:
        combined: dict[str, torch.Tensor] = {}
        for m in self._children():
            _accumulate_output(combined, m(state, **kwargs))
        return combined


class SerialSumModel(SumModel):
    """SumModel whose children see the earlier children's extra outputs as
    state extras (interface.py:325-351)."""

    def forward(self, state, **kwargs) -> dict[str, torch.Tensor]:
        combined: dict[str, torch.Tensor] = {}
        serial = state.clone()
        for m in self._children():
            out = m(serial, **kwargs)
            _accumulate_output(combined, out)
            serial.store_model_extras(out)
        return combined


class LennardJonesModel(PairPotentialModel):
    """LJ 12-6 model, keywords of lennard_jones.py:92-106."""

    def __init__(
        self,
        sigma: float = 1.0,
        epsilon: float = 1.0,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        cutoff: float | None = None,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        self.sigma = sigma
        self.epsilon = epsilon
        super().__init__(
            pair_fn=functools.partial(lennard_jones_pair, sigma=sigma, epsilon=epsilon),
            cutoff=cutoff if cutoff is not None else 2.5 * float(sigma),
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            reduce_to_half_list=True,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


class MorseModel(PairPotentialModel):
    """Morse model, keywords of morse.py:96-111."""

    def __init__(
        self,
        sigma: float = 1.0,
        epsilon: float = 5.0,
        alpha: float = 5.0,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        cutoff: float | None = None,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        self.sigma = sigma
        self.epsilon = epsilon
        self.alpha = alpha
        super().__init__(
            pair_fn=functools.partial(morse_pair, sigma=sigma, epsilon=epsilon, alpha=alpha),
            cutoff=cutoff if cutoff is not None else 2.5 * float(sigma),
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            reduce_to_half_list=True,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


class SoftSphereModel(PairPotentialModel):
    """Soft-sphere model, keywords of soft_sphere.py:192-208."""

    def __init__(
        self,
        sigma: float = 1.0,
        epsilon: float = 1.0,
        alpha: float = 2.0,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        use_neighbor_list: bool = True,  # noqa: ARG002
        cutoff: float | None = None,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        self.sigma = sigma
        self.epsilon = epsilon
        self.alpha = alpha
        super().__init__(
            pair_fn=functools.partial(soft_sphere_pair, sigma=sigma, epsilon=epsilon,
                                      alpha=alpha),
            cutoff=cutoff if cutoff is not None else float(sigma),
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            reduce_to_half_list=True,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


def _species_matrix(m, n, default) -> torch.Tensor:
    if m is None:
        return torch.full((n, n), default, dtype=torch.float64)
    return m


class SoftSphereMultiModel(PairPotentialModel):
    """Species-pair soft-sphere model, keywords of soft_sphere.py:265-283."""

    def __init__(
        self,
        atomic_numbers: torch.Tensor,
        sigma_matrix: torch.Tensor | None = None,
        epsilon_matrix: torch.Tensor | None = None,
        alpha_matrix: torch.Tensor | None = None,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        pbc: torch.Tensor | bool = True,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        use_neighbor_list: bool = True,
        neighbor_list_fn=torchsim_nl,
        cutoff: float | None = None,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        self.pbc = torch.tensor([pbc] * 3) if isinstance(pbc, bool) else pbc
        self.use_neighbor_list = use_neighbor_list
        unique_z = torch.unique(torch.as_tensor(atomic_numbers).cpu()).sort().values.long()
        n = len(unique_z)
        sig = _species_matrix(sigma_matrix, n, 1.0)
        eps = _species_matrix(epsilon_matrix, n, 1.0)
        alp = _species_matrix(alpha_matrix, n, 2.0)
        for name, m in (("sigma_matrix", sig), ("epsilon_matrix", eps), ("alpha_matrix", alp)):
            if m.shape != (n, n):
                raise ValueError(f"{name} must have shape ({n}, {n})")
            if not torch.allclose(m.detach().cpu().double(), m.detach().cpu().double().T):
                raise ValueError(f"{name} is not symmetric")
        self.sigma_matrix, self.epsilon_matrix, self.alpha_matrix = sig, eps, alp
        super().__init__(
            pair_fn=MultiSoftSpherePairFn(unique_z, sig, eps, alp),
            cutoff=cutoff or float(sig.detach().max()),
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            reduce_to_half_list=True,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


class MultiLennardJonesModel(PairPotentialModel):
    """Species-pair LJ model (per-pair sigma / epsilon matrices).

    The reference has no multi-species LJ class; its counterpart is a
    PairPotentialModel with a species-lookup pair function in the style of
    MultiSoftSpherePairFn (soft_sphere.py:90-167), which is what this builds.
    """

    def __init__(
        self,
        atomic_numbers: torch.Tensor,
        sigma_matrix: torch.Tensor,
        epsilon_matrix: torch.Tensor,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        *,
        cutoff: float,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        unique_z = torch.unique(torch.as_tensor(atomic_numbers).cpu()).sort().values.long()
        self.sigma_matrix = sigma_matrix
        self.epsilon_matrix = epsilon_matrix
        super().__init__(
            pair_fn=MultiLennardJonesPairFn(unique_z, sigma_matrix, epsilon_matrix),
            cutoff=cutoff,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            reduce_to_half_list=True,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )
