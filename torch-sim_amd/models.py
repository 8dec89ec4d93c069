"""Pair-potential models behind the reference ModelInterface surface.

``forward(state) -> {"energy" [S], "forces" [N,3], "stress" [S,3,3],
"energies" [N], "stresses" [N,3,3]}`` as torch_sim/models/interface.py:169-201
and pair_potential.py:351-464, with the constructor keywords of
LennardJonesModel (lennard_jones.py:92-106), MorseModel (morse.py:96-111),
SoftSphereModel (soft_sphere.py:192-208) and SoftSphereMultiModel
(soft_sphere.py:265-283).

Each call rebuilds the neighbor list (as the reference does,
pair_potential.py:168-174) with the HIP cell list and evaluates the pairs
with the HIP pair kernel (FP32 pair arithmetic, FP64 accumulation).  The
kernel walks full lists (each pair from both ends, no atomics); the result
equals the reference's half-list sum.

Reference quirk kept: per-atom ``energies`` give every atom the full energy
of each of its pairs (pair_potential.py:411-417), so they sum to twice the
system energy.
"""

from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream
from .neighbors import NeighborList, SystemIndex, TensorKey, VerletList, torchsim_nl

LJ, MORSE, SOFT_SPHERE = 0, 1, 2


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
            "No device setter has been defined for this model"
            " so the device cannot be changed after initialization."
        )

    @property
    def dtype(self) -> torch.dtype:
        return self._dtype

    @dtype.setter
    def dtype(self, dtype: torch.dtype) -> None:
        raise NotImplementedError(
            "No dtype setter has been defined for this model"
            " so the dtype cannot be changed after initialization."
        )

    @property
    def compute_stress(self) -> bool:
        return self._compute_stress

    @compute_stress.setter
    def compute_stress(self, compute_stress: bool) -> None:
        raise NotImplementedError(
            "No compute_stress setter has been defined for this model"
            " so compute_stress cannot be set after initialization."
        )

    @property
    def compute_forces(self) -> bool:
        return self._compute_forces

    @compute_forces.setter
    def compute_forces(self, compute_forces: bool) -> None:
        raise NotImplementedError(
            "No compute_forces setter has been defined for this model"
            " so compute_forces cannot be set after initialization."
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


class PairPotentialModel(ModelInterface):
    """Closed-form pair potential evaluated by the HIP pair kernel.

    ``kind`` is one of LJ / MORSE / SOFT_SPHERE; ``sigma``, ``epsilon`` and
    ``alpha`` are symmetric [n_species, n_species] matrices (or scalars for a
    single species).  ``species_z`` lists the atomic numbers of the species
    rows; atoms are mapped to rows through a lookup as in
    MultiSoftSpherePairFn (soft_sphere.py:90-167).
    """

    def __init__(
        self,
        kind: int,
        *,
        sigma,
        epsilon,
        alpha=2.0,
        cutoff: float,
        species_z=None,
        device: torch.device | None = None,
        dtype: torch.dtype = torch.float64,
        compute_forces: bool = True,
        compute_stress: bool = False,
        per_atom_energies: bool = False,
        per_atom_stresses: bool = False,
        neighbor_list_fn=torchsim_nl,
        reduce_to_half_list: bool = True,
        retain_graph: bool = False,
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        super().__init__()
        self._device = _default_device(device)
        self._dtype = dtype
        self._compute_forces = compute_forces
        self._compute_stress = compute_stress
        self.per_atom_energies = per_atom_energies
        self.per_atom_stresses = per_atom_stresses
        if neighbor_list_fn is not torchsim_nl and neighbor_list_fn is not None:
            raise NotImplementedError(
                "torch_sim_amd models build their neighbor list with the HIP cell list; "
                "a custom neighbor_list_fn is not supported"
            )
        self.neighbor_list_fn = torchsim_nl
        self.reduce_to_half_list = reduce_to_half_list
        self.retain_graph = retain_graph
        self.kind = int(kind)
        self.cutoff = float(cutoff)
        sig = torch.as_tensor(sigma, dtype=torch.float64)
        eps = torch.as_tensor(epsilon, dtype=torch.float64)
        alp = torch.as_tensor(alpha, dtype=torch.float64)
        if sig.ndim == 0:
            n = 1
            sig, eps, alp = sig.view(1, 1), eps.view(1, 1), alp.view(1, 1)
        else:
            n = sig.shape[0]
            eps = eps.expand(n, n) if eps.ndim == 0 else eps
            alp = alp.expand(n, n) if alp.ndim == 0 else alp
        for name, m in (("sigma", sig), ("epsilon", eps), ("alpha", alp)):
            if m.shape != (n, n):
                raise ValueError(f"{name} matrix must have shape ({n}, {n})")
            if not torch.allclose(m, m.T):
                raise ValueError(f"{name}_matrix is not symmetric")
        if n > 8:
            raise ValueError("at most 8 species are supported")
        self.n_species = n
        self._params = (
            torch.stack([sig, eps, alp], dim=-1).reshape(n * n, 3).contiguous().to(self._device)
        )
        self._z_to_idx = None
        if n > 1:
            if species_z is None:
                raise ValueError("species_z (atomic numbers of the species) is required")
            z = torch.as_tensor(species_z, dtype=torch.long).flatten()
            table = torch.full((int(z.max()) + 1,), -1, dtype=torch.int32)
            for k, zz in enumerate(z.tolist()):
                table[int(zz)] = k
            self._z_to_idx = table.to(self._device)
        self._nl = NeighborList(self.cutoff, tpa=tpa)
        # skin > 0: keep a Verlet list between calls and rebuild only the systems
        # that left the skin (results equal a rebuild at every call)
        self.skin = float(skin)
        self._vl = VerletList(self.cutoff, self.skin, tpa=tpa) if self.skin > 0 else None
        self._sysidx: SystemIndex | None = None
        self._species_key = None
        self._species = None
        self.tpa = tpa

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

    def forward(self, state, **_kwargs) -> dict[str, torch.Tensor]:
        if state.dtype != self._dtype:
            raise TypeError(
                f"SimState dtype {state.dtype} does not match model dtype {self._dtype}. "
                f"Either set the model dtype to {state.dtype} or convert the SimState "
                f"to {self._dtype} using sim_state.to(dtype={self._dtype})."
            )
        if self.retain_graph:
            raise NotImplementedError(
                "retain_graph=True: use torch_sim_amd.autograd for differentiable pair terms"
            )
        _lib.require_gpu(state.positions, "state.positions")
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
            LJ,
            sigma=sigma,
            epsilon=epsilon,
            cutoff=cutoff if cutoff is not None else 2.5 * sigma,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
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
            MORSE,
            sigma=sigma,
            epsilon=epsilon,
            alpha=alpha,
            cutoff=cutoff if cutoff is not None else 2.5 * sigma,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
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
            SOFT_SPHERE,
            sigma=sigma,
            epsilon=epsilon,
            alpha=alpha,
            cutoff=cutoff if cutoff is not None else sigma,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


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
        ones = torch.ones((n, n), dtype=torch.float64)
        sig = sigma_matrix.detach().cpu().double() if sigma_matrix is not None else ones
        eps = epsilon_matrix.detach().cpu().double() if epsilon_matrix is not None else ones
        alp = alpha_matrix.detach().cpu().double() if alpha_matrix is not None else 2.0 * ones
        for name, m in (("sigma_matrix", sig), ("epsilon_matrix", eps), ("alpha_matrix", alp)):
            if m.shape != (n, n):
                raise ValueError(f"{name} must have shape ({n}, {n})")
        self.sigma_matrix, self.epsilon_matrix, self.alpha_matrix = sig, eps, alp
        super().__init__(
            SOFT_SPHERE,
            sigma=sig,
            epsilon=eps,
            alpha=alp,
            cutoff=cutoff or float(sig.max()),
            species_z=unique_z,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            neighbor_list_fn=neighbor_list_fn,
            retain_graph=retain_graph,
            tpa=tpa,
            skin=skin,
        )


class MultiLennardJonesModel(PairPotentialModel):
    """Species-pair LJ model (per-pair sigma / epsilon matrices).

    The reference has no multi-species LJ class; its counterpart is a
    PairPotentialModel with a species-lookup pair function in the style of
    MultiSoftSpherePairFn (soft_sphere.py:90-167).
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
        tpa: int = 4,
        skin: float = 0.0,
    ) -> None:
        unique_z = torch.unique(torch.as_tensor(atomic_numbers).cpu()).sort().values.long()
        self.sigma_matrix = sigma_matrix.detach().cpu().double()
        self.epsilon_matrix = epsilon_matrix.detach().cpu().double()
        super().__init__(
            LJ,
            sigma=self.sigma_matrix,
            epsilon=self.epsilon_matrix,
            alpha=torch.zeros_like(self.sigma_matrix),
            cutoff=cutoff,
            species_z=unique_z,
            device=device,
            dtype=dtype,
            compute_forces=compute_forces,
            compute_stress=compute_stress,
            per_atom_energies=per_atom_energies,
            per_atom_stresses=per_atom_stresses,
            tpa=tpa,
            skin=skin,
        )
