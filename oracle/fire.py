"""Oracle FIRE relaxation with the Frechet cell filter (CPU, float64).

Test infrastructure only.  Restates torch_sim/optimizers/fire.py:21-499 and
torch_sim/optimizers/cell_filters.py:137-178, 230-284, 353-399 on a plain
dict state.  ``model_fn(positions, cell_col) -> {"energy","forces","stress"}``.
"""

from __future__ import annotations

import torch

from oracle import mat33

MAX_LOG_DEFORM = 2.0


def batched_vdot(x, y, idx, n):
    """Per-system sum of x*y (math.py:748-778)."""
    return torch.zeros(n, dtype=x.dtype).index_add(0, idx, (x * y).sum(dim=1))


def deform_grad(reference_row, current_row):
    """solve(ref, cur)^T (cell_filters.py:137-139)."""
    return torch.linalg.solve(reference_row, current_row).transpose(-2, -1)


def fire_init(
    positions,
    cell_col,
    masses,
    system_idx,
    model_fn,
    *,
    dt_start=0.1,
    alpha_start=0.1,
    frechet=False,
    unit=False,
    scalar_pressure=0.0,
    hydrostatic_strain=False,
    constant_volume=False,
):
    """fire_init + frechet / unit cell-filter init (fire.py:21-114,
    cell_filters.py:180-284)."""
    n_sys = cell_col.shape[0]
    out = model_fn(positions, cell_col)
    st = {
        "positions": positions.clone(),
        "cell": cell_col.clone(),
        "masses": masses,
        "system_idx": system_idx,
        "n_sys": n_sys,
        "forces": out["forces"],
        "energy": out["energy"],
        "stress": out.get("stress"),
        "velocities": torch.full_like(positions, float("nan")),
        "dt": torch.full((n_sys,), float(dt_start), dtype=positions.dtype),
        "alpha": torch.full((n_sys,), float(alpha_start), dtype=positions.dtype),
        "n_pos": torch.zeros(n_sys, dtype=torch.int32),
        "frechet": frechet or unit,
        "filter": "frechet" if frechet else ("unit" if unit else None),
        "hydrostatic_strain": hydrostatic_strain,
        "constant_volume": constant_volume,
    }
    if frechet or unit:
        counts = torch.bincount(system_idx, minlength=n_sys).to(positions.dtype)
        st["cell_factor"] = counts.view(n_sys, 1, 1)
        st["pressure"] = (scalar_pressure * torch.eye(3, dtype=positions.dtype)).expand(
            n_sys, 3, 3
        )
        st["reference_cell"] = cell_col.clone()
        vol = torch.linalg.det(cell_col).view(n_sys, 1, 1)
        virial = _apply_constraints(-vol * (out["stress"] + st["pressure"]), st)
        f = deform_grad(cell_col.mT, cell_col.mT)
        if frechet:
            st["cell_positions"] = torch.zeros((n_sys, 3, 3), dtype=positions.dtype)
            ucf = torch.bmm(virial, torch.linalg.inv(f.transpose(1, 2)))
            st["cell_forces"] = ucf / st["cell_factor"]
        else:  # cell_filters.py:180-227
            st["cell_positions"] = f * st["cell_factor"]
            st["cell_forces"] = virial / st["cell_factor"]
        st["cell_velocities"] = torch.full_like(st["cell_forces"], float("nan"))
    return st


def _apply_constraints(virial, st):
    """Hydrostatic-strain / constant-volume virial projections (cell_filters.py:117-134)."""
    eye = torch.eye(3, dtype=virial.dtype)
    if st.get("hydrostatic_strain"):
        virial = torch.diagonal(virial, dim1=1, dim2=2).mean(dim=1).view(-1, 1, 1) * eye
    if st.get("constant_volume"):
        virial = virial - torch.diagonal(virial, dim1=1, dim2=2).mean(dim=1).view(-1, 1, 1) * eye
    return virial


def compute_cell_forces(st, out):
    """compute_cell_forces, Frechet and unit branches (cell_filters.py:353-399)."""
    n = st["n_sys"]
    vol = torch.linalg.det(st["cell"]).view(n, 1, 1)
    virial = _apply_constraints(-vol * (out["stress"] + st["pressure"]), st)
    f = deform_grad(st["reference_cell"].mT, st["cell"].mT)
    if st["filter"] == "unit":
        vt = torch.linalg.solve(f, virial.transpose(-2, -1)).transpose(-2, -1)
        st["cell_forces"] = vt / st["cell_factor"]
        return
    ucf = torch.bmm(virial, torch.linalg.inv(f.transpose(1, 2)))
    log_f = mat33.logm(f).clamp(-MAX_LOG_DEFORM, MAX_LOG_DEFORM)
    st["cell_forces"] = mat33.frechet_cell_forces(log_f, ucf) / st["cell_factor"]


def ase_fire_step(
    st,
    model_fn,
    *,
    dt_max=1.0,
    n_min=5,
    f_inc=1.1,
    f_dec=0.5,
    alpha_start=0.1,
    f_alpha=0.99,
    max_step=0.2,
):
    """One ASE-flavoured FIRE step (fire.py:288-499)."""
    eps = 1e-16
    idx, n = st["system_idx"], st["n_sys"]
    cellf = st["frechet"]
    nan_v = st["velocities"].isnan().any(dim=1)
    st["velocities"] = torch.nan_to_num(st["velocities"], nan=0.0)
    if cellf:
        st["cell_velocities"] = torch.nan_to_num(st["cell_velocities"], nan=0.0)
    if bool(nan_v.all()):
        forces = st["forces"]
    else:
        if cellf:
            row = st["cell"].mT
            cur = deform_grad(row, row)  # fire.py:322-329 reference quirk: ~identity
            forces = torch.bmm(st["forces"].unsqueeze(1), cur[idx]).squeeze(1)
        else:
            forces = st["forces"]
        power = batched_vdot(forces, st["velocities"], idx, n)
        if cellf:
            power = power + (st["cell_forces"] * st["cell_velocities"]).sum(dim=(1, 2))
        pos = power > 0.0
        neg = ~pos
        inc = (st["n_pos"] > n_min) & pos
        st["dt"][inc] = torch.minimum(st["dt"][inc] * f_inc, torch.tensor(dt_max, dtype=st["dt"].dtype))
        st["alpha"][inc] *= f_alpha
        st["n_pos"][pos] += 1
        st["dt"][neg] *= f_dec
        st["alpha"][neg] = alpha_start
        st["n_pos"][neg] = 0
        vv = batched_vdot(st["velocities"], st["velocities"], idx, n)
        ff = batched_vdot(forces, forces, idx, n)
        if cellf:
            vv = vv + st["cell_velocities"].pow(2).sum(dim=(1, 2))
            ff = ff + st["cell_forces"].pow(2).sum(dim=(1, 2))
            vs = torch.sqrt(vv.view(n, 1, 1))
            fs = torch.sqrt(ff.view(n, 1, 1))
            mix_c = st["cell_forces"] / (fs + eps) * vs
            a = st["alpha"].view(n, 1, 1)
            st["cell_velocities"] = torch.where(
                pos.view(n, 1, 1),
                (1.0 - a) * st["cell_velocities"] + a * mix_c,
                torch.zeros_like(st["cell_velocities"]),
            )
        vs = torch.sqrt(vv[idx].unsqueeze(-1))
        fs = torch.sqrt(ff[idx].unsqueeze(-1))
        mix = forces * (vs / (fs + eps))
        a = st["alpha"][idx].unsqueeze(-1)
        st["velocities"] = torch.where(
            pos[idx].unsqueeze(-1),
            (1.0 - a) * st["velocities"] + a * mix,
            torch.zeros_like(st["velocities"]),
        )
    dt_a = st["dt"][idx].unsqueeze(-1)
    st["velocities"] = st["velocities"] + forces * dt_a
    dr = st["velocities"] * dt_a
    drs = batched_vdot(dr, dr, idx, n)
    if cellf:
        dtc = st["dt"].view(n, 1, 1)
        st["cell_velocities"] = st["cell_velocities"] + st["cell_forces"] * dtc
        drc = st["cell_velocities"] * dtc
        drs = drs + drc.pow(2).sum(dim=(1, 2))
        drsc = torch.sqrt(drs).view(n, 1, 1)
        drc = torch.where(drsc > max_step, max_step * drc / (drsc + eps), drc)
    drsa = torch.sqrt(drs)[idx].unsqueeze(-1)
    dr = torch.where(drsa > max_step, max_step * dr / (drsa + eps), dr)
    if cellf:
        ref = st["reference_cell"]
        cur = deform_grad(ref.mT, st["cell"].mT)
        frac = torch.linalg.solve(cur[idx], st["positions"].unsqueeze(-1)).squeeze(-1)
        new_frac = frac + dr
        cp = st["cell_positions"] + drc
        if st["filter"] == "unit":
            f_new = cp / st["cell_factor"]
        else:
            log_new = cp / st["cell_factor"]
            if bool((log_new.abs() > MAX_LOG_DEFORM).any()):
                log_new = log_new.clamp(-MAX_LOG_DEFORM, MAX_LOG_DEFORM)
                cp = log_new * st["cell_factor"]
            f_new = mat33.expm(log_new)
        st["cell"] = torch.bmm(f_new, ref)
        adj = deform_grad(ref.mT, st["cell"].mT)
        if st["filter"] == "unit":
            st["cell_positions"] = adj * st["cell_factor"]
        else:
            st["cell_positions"] = mat33.logm(adj) * st["cell_factor"]
        st["positions"] = torch.bmm(new_frac.unsqueeze(1), adj[idx].transpose(-2, -1)).squeeze(1)
    else:
        st["positions"] = st["positions"] + dr
    out = model_fn(st["positions"], st["cell"])
    st["forces"], st["energy"], st["stress"] = out["forces"], out["energy"], out.get("stress")
    if cellf:
        compute_cell_forces(st, out)
    return st


def vv_fire_step(
    st,
    model_fn,
    *,
    dt_max=1.0,
    n_min=5,
    f_inc=1.1,
    f_dec=0.5,
    alpha_start=0.1,
    f_alpha=0.99,
):
    """One velocity-Verlet FIRE step without a cell filter (fire.py:179-285)."""
    eps = 1e-16
    idx, n = st["system_idx"], st["n_sys"]
    nan_v = st["velocities"].isnan().any(dim=1)
    if bool(nan_v.any()):
        st["velocities"][nan_v] = 0
    m = st["masses"].unsqueeze(-1)
    dt_a = st["dt"][idx].unsqueeze(-1)
    st["velocities"] = st["velocities"] + 0.5 * dt_a * st["forces"] / m
    st["positions"] = st["positions"] + dt_a * st["velocities"]
    out = model_fn(st["positions"], st["cell"])
    st["forces"], st["energy"], st["stress"] = out["forces"], out["energy"], out.get("stress")
    st["velocities"] = st["velocities"] + 0.5 * dt_a * st["forces"] / m
    power = batched_vdot(st["forces"], st["velocities"], idx, n)
    pos = power > 0.0
    neg = ~pos
    st["n_pos"][pos] += 1
    inc = (st["n_pos"] > n_min) & pos
    st["dt"][inc] = torch.minimum(st["dt"][inc] * f_inc, torch.tensor(dt_max, dtype=st["dt"].dtype))
    st["alpha"][inc] *= f_alpha
    st["dt"][neg] *= f_dec
    st["alpha"][neg] = alpha_start
    st["n_pos"][neg] = 0
    vv = batched_vdot(st["velocities"], st["velocities"], idx, n)
    ff = batched_vdot(st["forces"], st["forces"], idx, n)
    vs = torch.sqrt(vv[idx].unsqueeze(-1))
    fs = torch.sqrt(ff[idx].unsqueeze(-1))
    mix = st["forces"] * (vs / (fs + eps))
    a = st["alpha"][idx].unsqueeze(-1)
    st["velocities"] = torch.where(
        pos[idx].unsqueeze(-1),
        (1.0 - a) * st["velocities"] + a * mix,
        torch.zeros_like(st["velocities"]),
    )
    return st
