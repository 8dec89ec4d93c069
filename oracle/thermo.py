"""Oracle thermostats and barostats (CPU, float64).  Test infrastructure only.

Restates the Nose-Hoover chain (torch_sim/integrators/md.py:249-565), the
NVT Nose-Hoover and velocity-rescaling integrators (nvt.py:243-800), the
anisotropic NPT Langevin step (npt.py:62-650), the isotropic MTK Nose-Hoover
barostat (npt.py:1110-2060) and the stochastic cell-rescaling barostats
(npt.py:2062-2986).  Random draws are caller-supplied tensors in the
reference's draw order, so trajectories can be compared with identical
noise.  ``force_fn(x, cell) -> (energy [S], forces [N,3], stress [S,3,3])``.
"""

from __future__ import annotations

import torch

from .md import gjf_particle_velocity

# Suzuki-Yoshida weights (md.py:309-333) as the reference's tensors hold
# them: torch.tensor of Python floats takes the default dtype, float32.
_SY = {
    1: [1.0],
    3: [0.828981543588751, -0.657963087177502, 0.828981543588751],
    5: [0.2967324292201065, 0.2967324292201065, -0.186929716880426, 0.2967324292201065,
        0.2967324292201065],
    7: [0.784513610477560, 0.235573213359357, -1.17767998417887, 1.31518632068391,
        -1.17767998417887, 0.235573213359357, 0.784513610477560],
}
SY_WEIGHTS = {k: torch.tensor(v, dtype=torch.float32) for k, v in _SY.items()}


def per_system_sum(v, system_idx, n_sys):
    return torch.zeros((n_sys,) + v.shape[1:], dtype=v.dtype).index_add(0, system_idx, v)


def kinetic_energy(p, m, system_idx, n_sys):
    """0.5 sum p^2 / m per system (quantities.py:104-139)."""
    return 0.5 * per_system_sum((p.square() / m.unsqueeze(-1)).sum(-1), system_idx, n_sys)


# ------------------------------------------------------- Nose-Hoover chain
def nhc_init(dof, ke, kT, tau, chain_length):
    """Chain state of md.py:386-417: zero positions and momenta, masses
    kT tau^2 with the first scaled by dof."""
    n = ke.shape[0]
    q = (kT.unsqueeze(-1) * torch.square(tau).unsqueeze(-1)
         * torch.ones((n, chain_length), dtype=ke.dtype))
    q[:, 0] *= dof
    return {"xi": torch.zeros((n, chain_length), dtype=ke.dtype),
            "pxi": torch.zeros((n, chain_length), dtype=ke.dtype),
            "Q": q, "tau": tau.clone(), "ke": ke.clone(), "dof": dof.clone()}


def nhc_update_mass(chain, kT):
    """md.py:537-565."""
    n, L = chain["Q"].shape
    q = kT.unsqueeze(-1) * torch.square(chain["tau"]).unsqueeze(-1) * torch.ones(
        (n, L), dtype=chain["Q"].dtype)
    q[:, 0] *= chain["dof"]
    return dict(chain, Q=q)


def _nhc_substep(delta, P, chain, kT, system_idx):
    """One substep (md.py:419-494); P scaled by the per-system factor."""
    xi, pxi, Q = chain["xi"].clone(), chain["pxi"].clone(), chain["Q"]
    KE, dof = chain["ke"], chain["dof"]
    d2 = delta / 2.0
    d4 = d2 / 2.0
    d8 = d4 / 2.0
    M = Q.shape[1] - 1
    if M > 0:
        G = torch.square(pxi[:, M - 1]) / Q[:, M - 1] - kT
        pxi[:, M] += d4 * G
    for m in range(M - 1, 0, -1):
        G = torch.square(pxi[:, m - 1]) / Q[:, m - 1] - kT
        sc = torch.exp(-d8 * pxi[:, m + 1] / Q[:, m + 1])
        pxi[:, m] = sc * (sc * pxi[:, m] + d4 * G)
    G = 2.0 * KE - dof * kT
    sc = torch.exp(-d8 * pxi[:, 1] / Q[:, 1]) if M > 0 else 1.0
    pxi[:, 0] = sc * (sc * pxi[:, 0] + d4 * G)
    sc = torch.exp(-d2 * pxi[:, 0] / Q[:, 0])
    KE = KE * torch.square(sc)
    P = P * sc[system_idx].unsqueeze(-1)
    xi = xi + d2 * pxi / Q
    G = 2.0 * KE - dof * kT
    for m in range(M):
        sc = torch.exp(-d8 * pxi[:, m + 1] / Q[:, m + 1])
        pxi[:, m] = sc * (sc * pxi[:, m] + d4 * G)
        G = torch.square(pxi[:, m]) / Q[:, m] - kT
    pxi[:, M] += d4 * G
    return P, dict(chain, xi=xi, pxi=pxi, ke=KE)


def nhc_half_step(P, chain, kT, system_idx, dt, chain_steps, sy_steps):
    """md.py:496-535: chain_steps x sy_steps substeps of dt / chain_steps
    times the (float32) Suzuki-Yoshida weights."""
    dt = torch.as_tensor(dt, dtype=torch.float64)
    if chain_steps == 1 and sy_steps == 1:
        return _nhc_substep(dt, P, chain, kT, system_idx)
    delta = dt / chain_steps
    w = SY_WEIGHTS[sy_steps]
    for k in range(chain_steps * sy_steps):
        P, chain = _nhc_substep(delta * w[k % sy_steps], P, chain, kT, system_idx)
    return P, chain


def chain_energy(chain, kT, dof):
    """Thermostat energy of a chain (npt.py:1966-1994, nvt.py:535-560)."""
    e = torch.square(chain["pxi"][:, 0]) / (2 * chain["Q"][:, 0]) + dof * kT * chain["xi"][:, 0]
    for i in range(1, chain["xi"].shape[1]):
        e = e + torch.square(chain["pxi"][:, i]) / (2 * chain["Q"][:, i]) + kT * chain["xi"][:, i]
    return e


def nvt_nose_hoover_step(x, p, f, m, system_idx, force_fn, cell, dt, kT, chain, chain_steps,
                         sy_steps):
    """nvt.py:381-488: masses, chain half step, velocity Verlet, chain
    kinetic energy, chain half step.  Returns (x, p, f, energy, chain)."""
    n_sys = kT.shape[0]
    chain = nhc_update_mass(chain, kT)
    p, chain = nhc_half_step(p, chain, kT, system_idx, dt, chain_steps, sy_steps)
    p = p + f * (dt / 2)
    x = x + (p / m.unsqueeze(-1)) * dt
    e, f, _ = force_fn(x, cell)
    p = p + f * (dt / 2)
    chain = dict(chain, ke=kinetic_energy(p, m, system_idx, n_sys))
    p, chain = nhc_half_step(p, chain, kT, system_idx, dt, chain_steps, sy_steps)
    return x, p, f, e, chain


def nvt_nose_hoover_invariant(energy, p, m, system_idx, chain, kT):
    """nvt.py:491-562 with dof = 3n - 3."""
    n_sys = kT.shape[0]
    dof = 3 * torch.bincount(system_idx, minlength=n_sys).to(p.dtype) - 3
    return energy + kinetic_energy(p, m, system_idx, n_sys) + chain_energy(chain, kT, dof)


# ------------------------------------------------------------------ CSVR
def vrescale_update(p, m, system_idx, dof, kT, tau, dt, r1, r2):
    """nvt.py:594-654 with the draws supplied: r1 [S] standard normals,
    r2 [S] = 2 Gamma((dof - 1) / 2)."""
    n_sys = kT.shape[0]
    current_kT = per_system_sum((p.square() / m.unsqueeze(-1)).sum(-1), system_idx, n_sys) / dof
    ke_old = dof * current_kT / 2
    ke_new = dof * kT / 2
    c1 = torch.exp(-dt / tau)
    c2 = (1 - c1) * ke_new / ke_old / dof
    scale = c1 + (c2 * (torch.square(r1) + r2)) + (2 * r1 * torch.sqrt(c1 * c2))
    lam = torch.sqrt(scale)
    return p * lam[system_idx].unsqueeze(-1)


def nvt_vrescale_step(x, p, f, m, system_idx, force_fn, cell, dt, kT, tau, r1, r2):
    """nvt.py:715-800: CSVR rescale (dof = 3n - 3), then velocity Verlet."""
    n_sys = kT.shape[0]
    dof = 3 * torch.bincount(system_idx, minlength=n_sys).to(p.dtype) - 3
    p = vrescale_update(p, m, system_idx, dof, kT, tau, dt, r1, r2)
    p = p + f * (dt / 2)
    x = x + (p / m.unsqueeze(-1)) * dt
    e, f, stress = force_fn(x, cell)
    p = p + f * (dt / 2)
    return x, p, f, e, stress


# ------------------------------------------------ anisotropic NPT Langevin
def aniso_strain_force(cell, stress, n_atoms, kT, p_ext):
    """F_k = V (P_kk - P_ext,k), P_kk = -sigma_kk + n kT / V (npt.py:370-406)."""
    vol = torch.linalg.det(cell)
    p_diag = -torch.diagonal(stress, dim1=-2, dim2=-1) + (n_atoms * kT / vol).unsqueeze(-1)
    return vol.unsqueeze(-1) * (p_diag - p_ext)


def npt_langevin_anisotropic_step(x, p, f, m, system_idx, ref_cell, cell, eps, eps_dot, stress,
                                  force_fn, dt, kT, p_ext, alpha, cell_alpha, b_tau, cell_noise,
                                  noise):
    """npt.py:513-650 with the draws supplied: cell_noise [S,3], then noise
    [N,3].  eps, eps_dot, p_ext [S,3].  Returns
    (x, p, f, energy, stress, cell, eps, eps_dot)."""
    n_sys = ref_cell.shape[0]
    cnt = torch.bincount(system_idx, minlength=n_sys).to(x.dtype)
    q = (cnt + 1) * kT * torch.square(b_tau)

    def cell_force(cell_now, stress_now):
        return aniso_strain_force(cell_now, stress_now, cnt, kT, p_ext)

    f_n = cell_force(cell, stress)
    cell_beta = torch.sqrt(2.0 * cell_alpha * kT * dt).unsqueeze(-1) * cell_noise
    beta = torch.sqrt(2 * alpha[system_idx] * kT[system_idx] * dt).unsqueeze(-1) * noise
    q2 = (2 * q).unsqueeze(-1)
    cb = 1 / (1 + (cell_alpha.unsqueeze(-1) * dt) / q2)
    eps_new = eps + cb * dt * eps_dot + cb * dt * dt * f_n / q2 + cb * dt * cell_beta / q2
    cell_new = torch.exp(eps_new).unsqueeze(-1) * ref_cell
    scale = torch.exp(eps_new - eps)[system_idx]
    b = 1 / (1 + ((alpha[system_idx] * dt) / (2 * m)))
    m2 = 2 * m.unsqueeze(-1)
    v = p / m.unsqueeze(-1)
    c2 = (2 * scale / (scale + 1)) * b.unsqueeze(-1) * dt
    x = scale * x + c2 * (v + dt * f / m2 + beta / m2)
    e, f_new, stress_new = force_fn(x, cell_new)
    f_new_eps = cell_force(cell_new, stress_new)
    qq = q.unsqueeze(-1)
    ca = cell_alpha.unsqueeze(-1)
    a = (1 - (ca * dt) / (2 * qq)) / (1 + (ca * dt) / (2 * qq))
    bb = 1 / (1 + (ca * dt) / (2 * qq))
    eps_dot = a * eps_dot + dt * ((a * f_n) + f_new_eps) / (2 * qq) + bb * cell_beta / qq
    v = gjf_particle_velocity(v, f, f_new, m, dt, alpha[system_idx], beta)
    return x, v * m.unsqueeze(-1), f_new, e, stress_new, cell_new, eps_new, eps_dot


# ------------------------------------------------- MTK Nose-Hoover NPT
def sinhx_x(x):
    """npt.py:1312-1343."""
    return (1 + torch.pow(x, 2) / 6 + torch.pow(x, 4) / 120 + torch.pow(x, 6) / 5040
            + torch.pow(x, 8) / 362_880 + torch.pow(x, 10) / 39_916_800)


def npt_nose_hoover_isotropic_step(x, p, f, m, system_idx, ref_cell, stress, force_fn, dt, kT,
                                   p_ext, eps, cell_mom, thermostat, barostat, chain_steps,
                                   sy_steps, b_tau):
    """npt.py:1812-1963.  eps (cell_position) [S], cell_mom [S].  Returns
    (x, p, f, energy, stress, cell, eps, cell_mom, cell_mass, thermostat,
    barostat)."""
    n_sys = ref_cell.shape[0]
    cnt = torch.bincount(system_idx, minlength=n_sys)
    dof = 3 * cnt.to(x.dtype) - 3
    barostat = nhc_update_mass(barostat, kT)
    thermostat = nhc_update_mass(thermostat, kT)
    w = 3 * (cnt + 1) * kT * torch.square(b_tau)
    cell_idx = torch.arange(n_sys)
    cm, barostat = nhc_half_step(cell_mom.unsqueeze(-1), barostat, kT, cell_idx, dt, chain_steps,
                                 sy_steps)
    cell_mom = cm.squeeze(-1)
    p, thermostat = nhc_half_step(p, thermostat, kT, system_idx, dt, chain_steps, sy_steps)
    # inner step (npt.py:1532-1640)
    dt2 = dt / 2
    v0 = torch.linalg.det(ref_cell)
    alpha = 1 + 3 / dof

    def cell_force(mom, stress_now, vol):
        ke = kinetic_energy(mom, m, system_idx, n_sys)
        tr = torch.diagonal(stress_now, dim1=-2, dim2=-1).sum(-1)
        return (alpha * 2 * ke) - (tr * vol) - (p_ext * vol * 3)

    vol = v0 * torch.exp(3 * eps)
    cell_mom = cell_mom + dt2 * cell_force(p, stress, vol)
    v_eps = cell_mom / w

    def il2(mom, force):
        xx = (alpha * v_eps * dt2)[system_idx]
        return (mom * torch.exp(-xx).unsqueeze(-1)
                + dt2 * force * sinhx_x(xx / 2).unsqueeze(-1) * torch.exp(-xx / 2).unsqueeze(-1))

    p = il2(p, f)
    eps = eps + v_eps * dt
    vol = v0 * torch.exp(3 * eps)
    cell = torch.pow(vol / v0, 1.0 / 3)[:, None, None] * ref_cell
    xx = (v_eps * dt)[system_idx].unsqueeze(-1)
    x = x + (x * (torch.exp(xx) - 1) + dt * (p / m.unsqueeze(-1)) * torch.exp(xx / 2)
             * sinhx_x(xx / 2))
    e, f, stress = force_fn(x, cell)
    p = il2(p, f)
    cell_mom = cell_mom + dt2 * cell_force(p, stress, vol)
    thermostat = dict(thermostat, ke=kinetic_energy(p, m, system_idx, n_sys))
    barostat = dict(barostat, ke=torch.square(cell_mom) / (2 * w))
    p, thermostat = nhc_half_step(p, thermostat, kT, system_idx, dt, chain_steps, sy_steps)
    cm, barostat = nhc_half_step(cell_mom.unsqueeze(-1), barostat, kT, cell_idx, dt, chain_steps,
                                 sy_steps)
    return (x, p, f, e, stress, cell, eps, cm.squeeze(-1), w, thermostat, barostat)


def npt_nose_hoover_isotropic_invariant(energy, p, m, system_idx, ref_cell, eps, cell_mom,
                                        cell_mass, thermostat, barostat, kT, p_ext):
    """npt.py:1997-2059."""
    n_sys = kT.shape[0]
    dof = 3 * torch.bincount(system_idx, minlength=n_sys).to(p.dtype) - 3
    vol = torch.linalg.det(torch.exp(eps)[:, None, None] * ref_cell)
    e = energy + kinetic_energy(p, m, system_idx, n_sys)
    e = e + chain_energy(thermostat, kT, dof)
    e = e + chain_energy(barostat, kT, torch.ones_like(dof))
    return e + p_ext * vol + torch.square(cell_mom) / (2 * cell_mass)


# ----------------------------------------------- stochastic cell rescaling
def rotate_gram_schmidt(box):
    """npt.py:2093-2126: columns a, b, c of each box to rows of a
    lower-triangular matrix."""
    out = torch.zeros_like(box)
    a, b, c = box[:, :, 0], box[:, :, 1], box[:, :, 2]
    out[:, 0, 0] = torch.norm(a, dim=1)
    out[:, 1, 0] = torch.sum(a * b, dim=1) / out[:, 0, 0]
    out[:, 1, 1] = torch.sqrt(torch.sum(b * b, dim=1) - out[:, 1, 0] ** 2)
    out[:, 2, 0] = torch.sum(a * c, dim=1) / out[:, 0, 0]
    out[:, 2, 1] = (torch.sum(b * c, dim=1) - out[:, 2, 0] * out[:, 1, 0]) / out[:, 1, 1]
    out[:, 2, 2] = torch.sqrt(torch.sum(c * c, dim=1) - out[:, 2, 0] ** 2 - out[:, 2, 1] ** 2)
    return out


def instantaneous_pressure(p, m, system_idx, stress, vol):
    """quantities.py:156-198: sum p p^T / m / V - stress."""
    n_sys = stress.shape[0]
    k = per_system_sum(torch.einsum("bi,bj,b->bij", p, p, 1 / m), system_idx, n_sys)
    return k / vol.view(-1, 1, 1) - stress


def crescale_barostat(mode, x, p, m, system_idx, cell, stress, kT, beta_t, tau_p, dt, p_ext,
                      noise_a, noise_r, noise_b, p_ext_tensor=None, cell0_inv=None, vol0=None):
    """One barostat step: mode "isotropic" (npt.py:2455-2492), "anisotropic"
    (2282-2343), "triclinic" (2186-2279; with p_ext_tensor the deviatoric
    correction of 2145-2183) or "average" (2376-2452).  noise_a, noise_b [S]
    and noise_r ([S,3,3], anisotropic [S,3]) in the reference's draw order.
    Returns (x, p, cell)."""
    n_sys = cell.shape[0]
    vol = torch.det(cell)
    eye = torch.eye(3, dtype=cell.dtype).expand(n_sys, 3, 3)
    if mode == "average":
        dof = 3 * torch.bincount(system_idx, minlength=n_sys).to(x.dtype) - 3
        P = (dof / 3 * kT / vol)[:, None, None] * eye - stress
    else:
        P = instantaneous_pressure(p, m, system_idx, stress, vol)
    sqv = torch.sqrt(vol)
    trP = torch.einsum("bii->b", P)
    pr = torch.sqrt(kT * beta_t * dt / (4 * tau_p))
    pf = beta_t * sqv / (2 * tau_p)
    if mode == "isotropic":
        chg = -pf * (p_ext - trP / 3 - kT / (2 * vol)) * dt + torch.sqrt(
            2 * torch.ones_like(sqv)) * pr * noise_a
        mu = torch.pow((sqv + chg) / sqv, 2 / 3).unsqueeze(-1)
        x = mu[system_idx] * x + (mu + 1 / mu)[system_idx] * p * (0.5 * dt) / m.unsqueeze(-1)
        p = (1 / mu)[system_idx] * p
        return x, p, mu.unsqueeze(-1) * cell
    eff = p_ext
    P_drive = P
    if mode == "triclinic" and p_ext_tensor is not None:
        hyd = torch.einsum("bii->b", p_ext_tensor) / 3
        dev = p_ext_tensor - hyd[:, None, None] * eye
        sig = vol0[:, None, None] * (cell0_inv.transpose(-2, -1) @ dev @ cell0_inv)
        sig = 0.5 * (sig + sig.transpose(-2, -1))
        pdev = cell.transpose(-2, -1) @ sig @ cell / vol[:, None, None]
        eff = hyd + torch.einsum("bii->b", pdev) / 3
        P_drive = P - pdev
    nsv = sqv + (-pf * (eff - trP / 3 - kT / (2 * vol)) * dt / 2 + pr * noise_a)
    prm = torch.sqrt(2 * beta_t * kT * dt / (3 * tau_p)) / nsv
    if mode == "anisotropic":
        pd = torch.diagonal(P, dim1=-2, dim2=-1)
        at = (beta_t / (3 * tau_p))[:, None] * (pd - trP[:, None] / 3)
        rt = noise_r - torch.mean(noise_r, dim=1, keepdim=True)
        D = torch.exp(at * dt + prm[:, None] * rt)
        nsv = nsv + (-pf * (eff - trP / 3 - kT / (2 * vol)) * dt / 2 + pr * noise_b)
        rs = D * torch.pow(nsv / sqv, 2 / 3).unsqueeze(-1)
        x = rs[system_idx] * x + (rs + 1 / rs)[system_idx] * p * dt / (2 * m.unsqueeze(-1))
        p = (1 / rs)[system_idx] * p
        return x, p, torch.diag_embed(rs) @ cell
    trd = torch.einsum("bii->b", P_drive)
    at = (beta_t / (3 * tau_p))[:, None, None] * (P_drive - trd[:, None, None] / 3 * eye)
    rt = noise_r - torch.einsum("bii->b", noise_r)[:, None, None] / 3 * eye
    D = rotate_gram_schmidt(torch.matrix_exp(at * dt + prm[:, None, None] * rt))
    nsv = nsv + (-pf * (eff - trP / 3 - kT / (2 * vol)) * dt / 2 + pr * noise_b)
    R = D * torch.pow(nsv / sqv, 2 / 3).view(-1, 1, 1)
    Ri = R[system_idx]
    if mode == "triclinic":
        V = torch.inverse(R).transpose(-2, -1)
        x = (Ri @ x.unsqueeze(-1)).squeeze(-1) + ((V + R)[system_idx] @ p.unsqueeze(-1)).squeeze(
            -1) * dt / (2 * m.unsqueeze(-1))
        p = (V[system_idx] @ p.unsqueeze(-1)).squeeze(-1)
    else:
        x = (Ri @ x.unsqueeze(-1)).squeeze(-1) + ((eye + R)[system_idx] @ p.unsqueeze(-1)).squeeze(
            -1) * dt / (2 * m.unsqueeze(-1))
    return x, p, cell @ R.mT


def npt_crescale_step(mode, x, p, f, m, system_idx, cell, stress, force_fn, dt, kT, p_ext,
                      beta_t, tau_p, tau, noise, p_ext_tensor=None, cell0_inv=None, vol0=None):
    """npt.py:2517-2901: CSVR(dt/2), B(dt/2), barostat, forces, B(dt/2),
    CSVR(dt/2) with dof = 3n - 3.  ``noise`` is the dict of draws (r1a, r2a,
    ra, R, rb, r1b, r2b).  Returns (x, p, f, energy, stress, cell)."""
    n_sys = cell.shape[0]
    dof = 3 * torch.bincount(system_idx, minlength=n_sys).to(x.dtype) - 3
    p = vrescale_update(p, m, system_idx, dof, kT, tau, dt / 2, noise["r1a"], noise["r2a"])
    p = p + f * (dt / 2)
    x, p, cell = crescale_barostat(mode, x, p, m, system_idx, cell, stress, kT, beta_t, tau_p, dt,
                                   p_ext, noise.get("ra"), noise.get("R"), noise.get("rb"),
                                   p_ext_tensor, cell0_inv, vol0)
    e, f, stress = force_fn(x, cell)
    p = p + f * (dt / 2)
    p = vrescale_update(p, m, system_idx, dof, kT, tau, dt / 2, noise["r1b"], noise["r2b"])
    return x, p, f, e, stress, cell
