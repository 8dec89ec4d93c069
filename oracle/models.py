"""Oracle pair potentials (CPU, float64).  Test infrastructure only.

Restates torch_sim/models/{lennard_jones,morse,soft_sphere,pair_potential,
particle_life}.py: the closed-form pair energies, the particle-life force,
the half-list forward pass with autograd forces and the pair-virial stress,
and the force-defined forward pass of PairForcesModel.
"""

from __future__ import annotations

import torch

from oracle import nl as onl


def lennard_jones_pair(dr, zi, zj, sigma=1.0, epsilon=1.0):
    """V = 4 eps [(s/r)^12 - (s/r)^6], 0 at r <= 0 (lennard_jones.py:23-47)."""
    idr6 = (sigma / dr).pow(6)
    e = 4.0 * epsilon * (idr6 * idr6 - idr6)
    return torch.where(dr > 0, e, torch.zeros_like(e))


def morse_pair(dr, zi, zj, sigma=1.0, epsilon=5.0, alpha=5.0):
    """V = eps (1 - exp(-a (r - s)))^2 - eps, 0 at r <= 0 (morse.py:23-47)."""
    e = epsilon * (1.0 - torch.exp(-alpha * (dr - sigma))).pow(2) - epsilon
    return torch.where(dr > 0, e, torch.zeros_like(e))


def soft_sphere_pair(dr, zi, zj, sigma=1.0, epsilon=1.0, alpha=2.0):
    """V = eps/a (1 - r/s)^a for r < s else 0 (soft_sphere.py:35-59)."""
    e = epsilon / alpha * (1.0 - dr / sigma).pow(alpha)
    return torch.where(dr < sigma, e, torch.zeros_like(e))


class MultiPairFn:
    """Species-pair parameter lookup (soft_sphere.py:90-167).

    ``kind`` selects the closed form ("soft_sphere" or "lennard_jones"); the
    per-pair parameters come from symmetric matrices indexed by the species of
    the two atoms (atomic number -> row via ``z_to_idx``).
    """

    def __init__(self, atomic_numbers, sigma, epsilon, alpha=None, kind="soft_sphere"):
        self.kind = kind
        self.sigma = sigma
        self.epsilon = epsilon
        n = len(atomic_numbers)
        self.alpha = alpha if alpha is not None else torch.full((n, n), 2.0)
        zmax = int(max(atomic_numbers)) + 1
        self.z_to_idx = torch.full((zmax,), -1, dtype=torch.long)
        for k, z in enumerate(list(atomic_numbers)):
            self.z_to_idx[int(z)] = k

    def __call__(self, dr, zi, zj):
        a, b = self.z_to_idx[zi], self.z_to_idx[zj]
        s, e = self.sigma[a, b], self.epsilon[a, b]
        if self.kind == "soft_sphere":
            al = self.alpha[a, b]
            out = e / al * (1.0 - dr / s).pow(al)
            return torch.where(dr < s, out, torch.zeros_like(out))
        idr6 = (s / dr).pow(6)
        out = 4.0 * e * (idr6 * idr6 - idr6)
        return torch.where(dr > 0, out, torch.zeros_like(out))


def pair_forward(
    positions,
    cell_col,
    pbc,
    atomic_numbers,
    system_idx,
    pair_fn,
    cutoff,
    *,
    half=True,
    compute_stress=True,
    per_atom=False,
    create_graph=False,
    nl=None,
):
    """Half-list pair-potential forward (pair_potential.py:130-197, 351-464).

    ``cell_col`` is the column-vector SimState cell [S,3,3]; the neighbor
    list receives the row-vector cell (pair_potential.py:154).  Returns a dict
    with energy [S], forces [N,3], stress [S,3,3] and optionally per-atom
    energies / stresses.  With ``create_graph`` the forces keep their graph so
    force-matching losses can be differentiated (the reference uses
    create_graph=False, pair_potential.py:419-425).
    """
    n_sys = cell_col.shape[0]
    row_cell = cell_col.transpose(1, 2)
    pbc_b = pbc.unsqueeze(0).expand(n_sys, -1) if pbc.ndim == 1 else pbc
    if nl is None:
        nl = onl.brute_force_nl
    with torch.no_grad():
        mapping, sysm, shifts = nl(
            positions.detach(), row_cell.detach(), pbc_b, cutoff, system_idx
        )
    if half:
        mapping, sysm, shifts = onl.full_to_half_list(mapping, sysm, shifts)
    cell_shifts = torch.einsum("jn,jnm->jm", shifts.to(positions.dtype), row_cell[sysm])
    dr = positions[mapping[1]] - positions[mapping[0]] + cell_shifts
    dist = dr.norm(dim=1)
    zi, zj = atomic_numbers[mapping[0]], atomic_numbers[mapping[1]]
    mask = dist < cutoff
    d_req = dist if create_graph else dist.detach().requires_grad_()
    e_pair = pair_fn(d_req, zi, zj)
    e_pair = torch.where(mask, e_pair, torch.zeros_like(e_pair))
    ew = 1.0 if half else 0.5
    energy = torch.zeros(n_sys, dtype=positions.dtype).index_add(0, sysm, ew * e_pair)
    out = {"energy": energy}
    if per_atom:
        ea = torch.zeros(positions.shape[0], dtype=positions.dtype)
        ea = ea.index_add(0, mapping[0], ew * e_pair).index_add(0, mapping[1], ew * e_pair)
        out["energies"] = ea
    (dvdr,) = torch.autograd.grad(
        e_pair.sum(), d_req, create_graph=create_graph, retain_graph=create_graph
    )
    safe = torch.where(dist > 0, dist, torch.ones_like(dist))
    fvec = (-dvdr / safe)[:, None] * dr
    forces = torch.zeros_like(positions)
    if half:
        forces = forces.index_add(0, mapping[0], -fvec).index_add(0, mapping[1], fvec)
    else:
        forces = forces.index_add(0, mapping[0], -fvec)
    out["forces"] = forces
    if compute_stress:
        vol = torch.abs(torch.linalg.det(row_cell))
        spp = torch.einsum("...i,...j->...ij", dr, fvec)
        stress = torch.zeros((n_sys, 3, 3), dtype=positions.dtype)
        stress = stress.index_add(0, sysm, -spp) / vol[:, None, None]
        out["stress"] = stress * (1.0 if half else 0.5)
        if per_atom:
            w = 0.5 if half else 0.25
            sa = torch.zeros((positions.shape[0], 3, 3), dtype=positions.dtype)
            sa = sa.index_add(0, mapping[0], -w * spp).index_add(0, mapping[1], -w * spp)
            out["stresses"] = sa / vol[system_idx, None, None]
    if not create_graph:
        out = {k: v.detach() for k, v in out.items()}
    return out


def particle_life_force(dr, zi, zj, A=1.0, beta=0.3, sigma=1.0):  # noqa: N803
    """Particle-life force magnitude (particle_life.py:24-56)."""
    zero = torch.zeros_like(dr)
    inner = torch.where(dr < beta, dr / beta - 1.0, zero)
    outer = A * (1.0 - torch.abs(2.0 * dr - 1.0 - beta) / (1.0 - beta))
    return inner + torch.where((dr >= beta) & (dr < sigma), outer, zero)


def pair_forces_forward(positions, cell_col, pbc, atomic_numbers, system_idx, force_fn, cutoff,
                        *, half=False, compute_stress=True, per_atom=False, nl=None):
    """PairForcesModel.forward (pair_potential.py:539-612): energies zero;
    every list entry adds -f r_hat to its first atom and +f r_hat to its
    second (on a full list each pair twice); the stress weights of
    _accumulate_stress (pair_potential.py:236-274)."""
    n_sys = cell_col.shape[0]
    row_cell = cell_col.transpose(1, 2)
    pbc_b = pbc.unsqueeze(0).expand(n_sys, -1) if pbc.ndim == 1 else pbc
    if nl is None:
        nl = onl.brute_force_nl
    mapping, sysm, shifts = nl(positions, row_cell, pbc_b, cutoff, system_idx)
    if half:
        mapping, sysm, shifts = onl.full_to_half_list(mapping, sysm, shifts)
    cell_shifts = torch.einsum("jn,jnm->jm", shifts.to(positions.dtype), row_cell[sysm])
    dr = positions[mapping[1]] - positions[mapping[0]] + cell_shifts
    dist = dr.norm(dim=1)
    f = force_fn(dist, atomic_numbers[mapping[0]], atomic_numbers[mapping[1]])
    f = torch.where(dist < cutoff, f, torch.zeros_like(f))
    safe = torch.where(dist > 0, dist, torch.ones_like(dist))
    fvec = (f / safe)[:, None] * dr
    forces = torch.zeros_like(positions).index_add(0, mapping[0], -fvec)
    forces = forces.index_add(0, mapping[1], fvec)
    out = {"energy": torch.zeros(n_sys, dtype=positions.dtype), "forces": forces}
    if compute_stress:
        vol = torch.abs(torch.linalg.det(row_cell))
        spp = torch.einsum("...i,...j->...ij", dr, fvec)
        stress = torch.zeros((n_sys, 3, 3), dtype=positions.dtype).index_add(0, sysm, -spp)
        out["stress"] = stress / vol[:, None, None] * (1.0 if half else 0.5)
        if per_atom:
            w = 0.5 if half else 0.25
            sa = torch.zeros((positions.shape[0], 3, 3), dtype=positions.dtype)
            sa = sa.index_add(0, mapping[0], -w * spp).index_add(0, mapping[1], -w * spp)
            out["stresses"] = sa / vol[system_idx, None, None]
    return out
