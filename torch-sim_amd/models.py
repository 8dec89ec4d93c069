"""Pair-potential models behind the reference ModelInterface surface.

``forward(state) -> {"energy" [S], "forces" [N,3], "stress" [S,3,3],
"energies" [N], "stresses" [N,3,3]}`` as torch_sim/models/interface.py:169-201
and pair_potential.py:351-464.  ``PairPotentialModel(pair_fn, *, cutoff,
...)`` takes the reference's constructor (pair_potential.py:300-345); the
convenience models build their pair function with ``functools.partial``
exactly as LennardJonesModel (lennard_jones.py:92-137), MorseModel
(morse.py:96-144), SoftSphereModel (soft_sphere.py:192-246) and
SoftSphereMultiModel (soft_sphere.py:265-377) do.

The pair function is not called during a forward pass: the model recognises
the closed forms (``lennard_jones_pair``, ``morse_pair``,
``soft_sphere_pair`` bound with ``functools.partial``, and the species-pair
``MultiSoftSpherePairFn`` / ``MultiLennardJonesPairFn``) and hands their
parameters to the HIP pair kernel (FP32 pair arithmetic, FP64
accumulation).  Any other callable is refused with a TypeError.

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

LJ, MORSE, SOFT_SPHERE = 0, 1, 2

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


def resolve_pair_fn(pair_fn):
    """(kind, sigma, epsilon, alpha, species_z) of a recognised pair function.

    Scalars come back as given (floats or 0-d tensors, which may require
    grad); species-pair functions give [n, n] matrices and the atomic numbers
    of their rows."""
    if isinstance(pair_fn, (MultiSoftSpherePairFn, MultiLennardJonesPairFn)):
        kind = SOFT_SPHERE if isinstance(pair_fn, MultiSoftSpherePairFn) else LJ
        return (kind, pair_fn.sigma_matrix, pair_fn.epsilon_matrix, pair_fn.alpha_matrix,
                pair_fn.atomic_numbers)
    func, kw = pair_fn, {}
    if isinstance(pair_fn, functools.partial):
        func, kw = pair_fn.func, dict(pair_fn.keywords)
        if pair_fn.args:
            func = None  # positional binding of dr / zi / zj: not a closed form here
    kind = _CLOSED_FORMS.get(func) if func is not None else None
    if kind is None:
        raise TypeError(
            "torch_sim_amd evaluates pair potentials in HIP kernels and recognises "
            "functools.partial(lennard_jones_pair | morse_pair | soft_sphere_pair, sigma=..., "
            "epsilon=..., alpha=...) and MultiSoftSpherePairFn / MultiLennardJonesPairFn; "
            f"got {pair_fn!r}"
        )
    sig = inspect.signature(func).parameters
    allowed = {"sigma", "epsilon"} | ({"alpha"} if kind != LJ else set())
    extra = set(kw) - allowed
    if extra:
        raise TypeError(f"unknown parameters {sorted(extra)} for {func.__name__}")
    vals = {k: kw.get(k, sig[k].default) for k in allowed}
    return kind, vals["sigma"], vals["epsilon"], vals.get("alpha", 2.0), None


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


class PairPotentialModel(ModelInterface):
    """Batched pair potential (pair_potential.py:277-464) on the HIP kernels.

    ``pair_fn`` must be one of the recognised closed forms (module
    docstring); its parameters may be floats, tensors or, for species-pair
    functions, symmetric [n, n] matrices.  Extra keywords of this backend:
    ``tpa`` (lanes per atom in the pair kernel) and ``skin`` (keep a Verlet
    list between calls; results equal a rebuild at every call).  ``device``
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
        kind, sigma, epsilon, alpha, species_z = resolve_pair_fn(pair_fn)
        _check_nl_fn(neighbor_list_fn)
        super().__init__()
        self._device = _default_device(device)
        self._dtype = dtype
        self._compute_forces = compute_forces
        self._compute_stress = compute_stress
        self.per_atom_energies = per_atom_energies
        self.per_atom_stresses = per_atom_stresses
        self.pair_fn = pair_fn
        self.neighbor_list_fn = torchsim_nl
        self.reduce_to_half_list = reduce_to_half_list
        self.retain_graph = retain_graph
        self.kind = int(kind)
        self.cutoff = float(cutoff)
        self._param_src = (sigma, epsilon, alpha)
        self._param_versions = None
        sig = self._matrix(sigma)
        self.n_species = n = sig.shape[0]
        if n > 8:
            raise ValueError("at most 8 species are supported")
        self._refresh_params()
        self._z_to_idx = None
        if n > 1:
            if species_z is None:
                raise ValueError("species-pair parameters need the atomic numbers of the rows")
            z = torch.as_tensor(species_z, dtype=torch.long).flatten().cpu()
            table = torch.full((int(z.max()) + 1,), -1, dtype=torch.int32)
            for k, zz in enumerate(z.tolist()):
                table[int(zz)] = k
            self._z_to_idx = table.to(self._device)
        self._nl = NeighborList(self.cutoff, tpa=tpa)
        self.skin = float(skin)
        self._vl = VerletList(self.cutoff, self.skin, tpa=tpa) if self.skin > 0 else None
        self._sysidx: SystemIndex | None = None
        self._species_key = None
        self._species = None
        self._diff_terms = None
        self.tpa = tpa

    # ------------------------------------------------------------- params
    @staticmethod
    def _matrix(v) -> torch.Tensor:
        t = torch.as_tensor(v).detach().to(torch.float64).cpu()
        return t.view(1, 1) if t.ndim == 0 else t

    def _refresh_params(self) -> None:
        """The [n^2, 3] device table of (sigma, epsilon, alpha), rebuilt when a
        parameter tensor changed in place since the last build."""
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

    def graph_params(self) -> torch.Tensor:
        """[n, n, 3] (sigma, epsilon, alpha) built from the pair function's own
        values, so tensors among them that require grad stay in the graph."""
        from .autograd import pair_params

        sig, eps, alp = self._param_src
        return pair_params(sig, eps, alp).to(self._device)

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

    def forward(self, state, **_kwargs) -> dict[str, torch.Tensor]:
        _check_dtype(state, self._dtype)
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
        need_w = self._compute_stress
        w_atom = torch.empty((n, 6), dtype=torch.float64, device=dev) if need_w else None
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
        stress = torch.empty((s, 3, 3), dtype=torch.float64, device=dev) if need_w else None
        stresses = (
            torch.empty((n, 3, 3), dtype=torch.float64, device=dev)
            if need_w and self.per_atom_stresses
            else None
        )
        call(
            "tsx_system_reduce", ptr(e_atom), ptr(w_atom), ptr(cell), ptr(si.sys_start),
            si.max_atoms, s, ptr(energy), ptr(stress), ptr(stresses), ptr(si.reduce_ws),
            stream(),
        )
        dt = self._dtype
        out = {"energy": energy.to(dt)}
        if self.per_atom_energies:
            out["energies"] = (2.0 * e_atom).to(dt)
        if self._compute_forces:
            out["forces"] = forces.to(dt)
        if need_w:
            out["stress"] = stress.to(dt)
            if stresses is not None:
                out["stresses"] = stresses.to(dt)
        if self.retain_graph and n:
            self._attached(state, out)
        return out


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
