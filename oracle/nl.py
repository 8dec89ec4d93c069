"""Oracle neighbor lists (CPU, float64).  Test infrastructure only.

Restates the reference neighbor-list semantics:

* wrap into the primary cell and record integer lattice shifts
  (torch_sim/transforms.py:160-210),
* number of image repeats per axis (torch_sim/transforms.py:435-467),
* the strict FP64 cutoff ``d^2 < rc^2`` on wrapped coordinates
  (torch_sim/neighbors/torch_nl.py:25-91) and the shift correction that makes
  shifts valid for unwrapped input (torch_sim/neighbors/torch_nl.py:214),
* the serial linked-cell search (torch_sim/transforms.py:843-1006),
* full -> half reduction (torch_sim/models/pair_potential.py:86-127).

The pair set returned by every function here is the full ordered list
{(i, j, S)} with i, j in the same system, |S_k| <= n_rep_k on periodic axes
(S_k = 0 otherwise), (i, i, 0) excluded unless ``self_interaction``.
"""

from __future__ import annotations

import itertools

import torch


def normalize_inputs(cell: torch.Tensor, pbc: torch.Tensor, n_systems: int):
    """Cell -> [S,3,3], pbc -> [S,3] (torch_sim/neighbors/utils.py:7-34)."""
    if cell.ndim == 2:
        cell = (
            cell.unsqueeze(0).expand(n_systems, -1, -1)
            if cell.shape[0] == 3
            else cell.reshape(n_systems, 3, 3)
        )
    if pbc.ndim == 1:
        pbc = (
            pbc.unsqueeze(0).expand(n_systems, -1)
            if pbc.shape[0] == 3
            else pbc.reshape(n_systems, 3)
        )
    return cell.contiguous(), pbc.contiguous()


def wrap_positions(positions, row_cell, system_idx, pbc):
    """Wrapped positions and integer shifts (torch_sim/transforms.py:160-210).

    ``row_cell`` rows are lattice vectors.  frac = inv(cell^T) r; on periodic
    axes of systems with an invertible cell, shift = floor(frac) and
    wrapped = cell^T (frac - shift).  Other atoms are left unchanged.
    """
    cell_t = row_cell.transpose(1, 2)
    dets = torch.linalg.det(cell_t)
    invertible = torch.isfinite(dets) & (dets.abs() > 1e-12)
    active = pbc.any(dim=1) & invertible
    if not bool(active.any()):
        return positions.clone(), torch.zeros_like(positions)
    inv = torch.zeros_like(row_cell)
    inv[active] = torch.linalg.inv(cell_t[active])
    frac = torch.bmm(inv[system_idx], positions.unsqueeze(2)).squeeze(2)
    act_atom = active[system_idx].unsqueeze(1)
    mask = pbc[system_idx] & act_atom
    ishift = torch.where(mask, torch.floor(frac), torch.zeros_like(frac))
    wrapped = torch.bmm(cell_t[system_idx], (frac - ishift).unsqueeze(2)).squeeze(2)
    out = torch.where(act_atom, wrapped, positions)
    shifts = torch.where(act_atom, ishift, torch.zeros_like(ishift))
    return out, shifts


def num_cell_repeats(cutoff: float, row_cell, pbc):
    """ceil(rc * |reciprocal row|) on periodic axes (transforms.py:435-467)."""
    row_cell = row_cell.view(-1, 3, 3)
    pbc = pbc.view(-1, 3)
    has_pbc = pbc.any(dim=1)
    recip = torch.zeros_like(row_cell)
    if bool(has_pbc.any()):
        recip[has_pbc] = torch.linalg.inv(row_cell[has_pbc]).transpose(2, 1)
    n = torch.ceil(cutoff * recip.norm(2, dim=-1)).to(torch.long)
    return torch.where(pbc, n, torch.zeros_like(n))


def _cell_shift_vectors(row_cell_s, shifts):
    """S . L with the reference einsum "jn,jnm->jm" (transforms.py:540-566)."""
    return torch.einsum("jn,nm->jm", shifts, row_cell_s)


def brute_force_nl(
    positions, cell, pbc, cutoff, system_idx, self_interaction=False
):
    """Exhaustive FP64 enumeration of the reference pair set.

    Returns (mapping [2,P] int64, system_mapping [P] int64, shifts [P,3] f64)
    sorted lexicographically by (i, j, Sx, Sy, Sz).  Shifts are valid for the
    *unwrapped* input positions (torch_nl.py:214).
    """
    dtype = positions.dtype
    n_sys = int(system_idx.max().item()) + 1 if system_idx.numel() else 0
    cell, pbc = normalize_inputs(cell, pbc, n_sys)
    wrapped, wshift = wrap_positions(positions, cell, system_idx, pbc)
    reps = num_cell_repeats(cutoff, cell, pbc)
    counts = torch.bincount(system_idx, minlength=n_sys)
    offs = torch.zeros(n_sys + 1, dtype=torch.long)
    offs[1:] = torch.cumsum(counts, 0)
    rc2 = cutoff * cutoff
    rows = []
    for s in range(n_sys):
        a, b = int(offs[s]), int(offs[s + 1])
        if b == a:
            continue
        n = reps[s].tolist()
        rng = [range(-k, k + 1) for k in n]
        svec = torch.tensor(list(itertools.product(*rng)), dtype=dtype)
        cart = _cell_shift_vectors(cell[s], svec)  # [ns,3]
        p = wrapped[a:b]
        # diff[s, i, j] = p_i - p_j - S.L  (torch_nl.py:87 ordering)
        diff = p[None, :, None, :] - p[None, None, :, :] - cart[:, None, None, :]
        d2 = diff.square().sum(-1)
        ok = d2 < rc2
        if not self_interaction:
            zero = (svec == 0).all(dim=1)
            eye = torch.eye(b - a, dtype=torch.bool)
            ok = ok & ~(zero[:, None, None] & eye[None])
        si, ii, jj = torch.nonzero(ok, as_tuple=True)
        gi, gj = ii + a, jj + a
        sh = svec[si] + wshift[gi] - wshift[gj]
        rows.append((gi, gj, sh))
    if not rows:
        return (
            torch.zeros((2, 0), dtype=torch.long),
            torch.zeros(0, dtype=torch.long),
            torch.zeros((0, 3), dtype=dtype),
        )
    gi = torch.cat([r[0] for r in rows])
    gj = torch.cat([r[1] for r in rows])
    sh = torch.cat([r[2] for r in rows])
    return sort_pairs(torch.stack([gi, gj]), system_idx[gi], sh)


def sort_pairs(mapping, system_mapping, shifts):
    """Canonical lexicographic order (i, j, Sx, Sy, Sz) for set comparison."""
    if mapping.shape[1] == 0:
        return mapping, system_mapping, shifts
    keys = torch.cat([mapping.T.to(torch.float64), shifts.to(torch.float64)], dim=1)
    order = torch.arange(keys.shape[0])
    for col in range(4, -1, -1):
        idx = torch.argsort(keys[order, col], stable=True)
        order = order[idx]
    return mapping[:, order], system_mapping[order], shifts[order]


def linked_cell_single(pos, row_cell, cutoff, n_rep, self_interaction=False):
    """Serial linked-cell search for one wrapped system (transforms.py:843-1006).

    Replicates the cell over the image shifts, bins every image into cubic
    bins of edge ``cutoff`` over the bounding box, and gathers candidates from
    the 27 surrounding bins of each original atom.  Returns candidate pairs
    (i, j, S) before the strict cutoff filter.
    """
    n = pos.shape[0]
    dtype = pos.dtype
    rng = [range(-k, k + 1) for k in n_rep.tolist()]
    svec = torch.tensor(list(itertools.product(*rng)), dtype=dtype)
    # put the zero shift first so image index == atom index for S = 0
    zero_first = torch.argsort(svec.abs().sum(1), stable=True)
    svec = svec[zero_first]
    n_img = svec.shape[0]
    images = (pos[None] + (svec @ row_cell)[:, None]).reshape(-1, 3)
    img_shift = svec.repeat_interleave(n, dim=0)
    bmin = images.min(0).values
    bmax = images.max(0).values
    shifted = images - (bmin - 1e-5)
    nb = torch.clamp(torch.ceil((bmax - bmin + 1e-3) / cutoff), min=1).to(torch.long)
    b3 = torch.floor(shifted / cutoff).to(torch.long)
    b3 = torch.minimum(torch.clamp(b3, min=0), nb - 1)
    blin = (b3[:, 0] * nb[1] + b3[:, 1]) * nb[2] + b3[:, 2]
    nbins = int(nb.prod())
    # bin table [nbins, max_apb] of image indices (sentinel = n_images)
    order = torch.argsort(blin, stable=True)
    sb = blin[order]
    cnt = torch.bincount(blin, minlength=nbins)
    start = torch.zeros(nbins + 1, dtype=torch.long)
    start[1:] = torch.cumsum(cnt, 0)
    within = torch.arange(sb.numel()) - start[sb]
    max_apb = int(cnt.max())
    n_images = images.shape[0]
    table = torch.full((nbins, max_apb), n_images, dtype=torch.long)
    table[sb, within] = order
    # 27-bin stencil around the bin of every original atom
    d = torch.tensor([0, 1, -1])
    stencil = torch.cartesian_prod(d, d, d)
    nbr3 = b3[:n, None, :] + stencil[None]  # [n, 27, 3]
    ok = ((nbr3 >= 0) & (nbr3 < nb)).all(dim=-1)
    nbr3 = torch.minimum(torch.clamp(nbr3, min=0), nb - 1)
    lin = (nbr3[..., 0] * nb[1] + nbr3[..., 1]) * nb[2] + nbr3[..., 2]
    cand = table[lin]  # [n, 27, max_apb]
    valid = ok[..., None] & (cand != n_images)
    ii = torch.arange(n)[:, None, None].expand_as(cand)[valid]
    jj = cand[valid]
    if not self_interaction:
        keep = ii != jj
        ii, jj = ii[keep], jj[keep]
    del n_img
    return ii, jj % n, img_shift[jj]


def linked_cell_nl(positions, cell, pbc, cutoff, system_idx, self_interaction=False):
    """Reference linked-cell neighbor list (torch_nl.py:156-215), per system.

    Uses :func:`linked_cell_single` (transforms.py:843-1006) for each system,
    then the strict FP64 filter (torch_nl.py:25-91) and the unwrapped-shift
    correction (torch_nl.py:214).  Output order is the search order (the
    reference order is unspecified; compare sets).
    """
    dtype = positions.dtype
    n_sys = int(system_idx.max().item()) + 1
    cell, pbc = normalize_inputs(cell, pbc, n_sys)
    wrapped, wshift = wrap_positions(positions, cell, system_idx, pbc)
    reps = num_cell_repeats(cutoff, cell, pbc)
    counts = torch.bincount(system_idx, minlength=n_sys)
    offs = torch.zeros(n_sys + 1, dtype=torch.long)
    offs[1:] = torch.cumsum(counts, 0)
    mi, mj, ms = [], [], []
    for s in range(n_sys):
        a, b = int(offs[s]), int(offs[s + 1])
        if b == a:
            continue
        ii, jj, sh = linked_cell_single(
            wrapped[a:b], cell[s], cutoff, reps[s], self_interaction
        )
        mi.append(ii + a)
        mj.append(jj + a)
        ms.append(sh)
    gi, gj, sh = torch.cat(mi), torch.cat(mj), torch.cat(ms)
    sysm = system_idx[gi]
    cart = torch.einsum("jn,jnm->jm", sh, cell[sysm])
    d2 = (wrapped[gi] - wrapped[gj] - cart).square().sum(dim=1)
    keep = d2 < cutoff * cutoff
    gi, gj, sh, sysm = gi[keep], gj[keep], sh[keep], sysm[keep]
    sh = sh + wshift[gi] - wshift[gj]
    return torch.stack([gi, gj]), sysm, sh.to(dtype)


def full_to_half_list(mapping, system_mapping, shifts):
    """Keep i<j; for i==j keep the lexicographically positive shift.

    Restates torch_sim/models/pair_potential.py:86-127.
    """
    i, j = mapping[0], mapping[1]
    keep = i < j
    same = i == j
    if bool(same.any()):
        s = shifts[same]
        first = torch.zeros(s.shape[0], dtype=s.dtype)
        done = torch.zeros(s.shape[0], dtype=torch.bool)
        for dim in range(3):
            col = s[:, dim]
            nz = (col != 0) & ~done
            first = torch.where(nz, col, first)
            done = done | nz
        keep = keep.clone()
        keep[same] = first > 0
    return mapping[:, keep], system_mapping[keep], shifts[keep]
