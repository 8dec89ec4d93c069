// NPT Langevin (gfx950): the Gronbech-Jensen-Farago integrator on the
// particles and on the logarithmic strain of each system's cell -- one strain
// eps = ln(V / V0) / 3 (isotropic, torch_sim/integrators/npt.py:653-1107) or
// one per Cartesian axis (anisotropic, npt.py:62-650).
//
// A step is two launches around the model call (the anisotropic pair
// tsx_npt_langevin_aniso_pre / _post does the same per axis):
//   tsx_npt_langevin_iso_pre   per system: barostat mass Q = (n + 1) kT b_tau^2,
//                              strain force F_eps = 3 V (P - P_ext) with
//                              P = -tr(sigma) / 3 + n kT / V, the strain's GJF
//                              position step, the cell exp(eps) h0; per atom:
//                              the noise beta = sqrt(2 alpha kT dt) xi and the
//                              GJF position step scaled by exp(eps' - eps)
//   tsx_npt_langevin_iso_post  per system: the strain force at the new cell and
//                              the strain's GJF velocity step; per atom: the
//                              GJF velocity step with the same beta
// The arithmetic follows the reference's operation order in FP64; the noise
// (standard normals, cell first, then particles) comes from the caller.
#include "tsx_common.h"

namespace tsx {

// per-system scratch: isotropic F_eps(n), beta_c, eps(n), exp(eps' - eps);
// anisotropic F_eps(n)[3], beta_c[3], exp(eps' - eps)[3]
constexpr int NPT_SW = 9;

__global__ void k_npt_iso_sys_pre(int n_sys, const int32_t* __restrict__ sys_start,
                                  const double* __restrict__ kT, const double* __restrict__ p_ext,
                                  const double* __restrict__ cell_alpha,
                                  const double* __restrict__ b_tau, double dt,
                                  const double* __restrict__ stress,
                                  const double* __restrict__ ref_cell, double* __restrict__ eps,
                                  const double* __restrict__ eps_dot,
                                  double* __restrict__ cell_masses, double* __restrict__ cell_col,
                                  const double* __restrict__ cell_noise,
                                  double* __restrict__ sw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double n = (double)(sys_start[s + 1] - sys_start[s]);
  const double kt = kT[s];
  const double q = (n + 1.0) * kt * (b_tau[s] * b_tau[s]);
  cell_masses[s] = q;
  const double* h0 = ref_cell + 9 * (int64_t)s;
  const double v0 = det33(h0);
  const double e0 = eps[s];
  const double vol = v0 * exp(3.0 * e0);
  const double* sg = stress + 9 * (int64_t)s;
  const double p_vir = -((sg[0] + sg[4]) + sg[8]) / 3.0;
  const double f_n = 3.0 * vol * ((p_vir + n * kt / vol) - p_ext[s]);
  const double ca = cell_alpha[s];
  const double beta_c = sqrt(2.0 * ca * kt * dt) * cell_noise[s];
  const double q2 = 2.0 * q;
  const double b = 1.0 / (1.0 + (ca * dt) / q2);
  const double e1 = ((e0 + b * dt * eps_dot[s]) + b * dt * dt * f_n / q2) + b * dt * beta_c / q2;
  eps[s] = e1;
  const double sc = exp(e1);
  for (int k = 0; k < 9; ++k) cell_col[9 * (int64_t)s + k] = sc * h0[k];
  sw[NPT_SW * s + 0] = f_n;
  sw[NPT_SW * s + 1] = beta_c;
  sw[NPT_SW * s + 2] = e0;
  sw[NPT_SW * s + 3] = exp(e1 - e0);
}

__global__ void k_npt_iso_atoms_pre(int64_t n_atoms, const int32_t* __restrict__ atom_sys,
                                    double* __restrict__ pos, const double* __restrict__ mom,
                                    const double* __restrict__ forces,
                                    const double* __restrict__ masses,
                                    const double* __restrict__ kT,
                                    const double* __restrict__ alpha, double dt,
                                    const double* __restrict__ noise,
                                    const double* __restrict__ sw, double* __restrict__ beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  const int s = atom_sys[i];
  const double al = alpha[s];
  const double mi = masses[i];
  const double pre = sqrt(2.0 * al * kT[s] * dt);
  const double sc = sw[NPT_SW * s + 3];
  const double m2 = 2.0 * mi;
  const double b = 1.0 / (1.0 + (al * dt) / (2.0 * mi));
  const double c2 = (2.0 * sc / (sc + 1.0)) * b * dt;
  for (int c = 0; c < 3; ++c) {
    const int64_t k = 3 * i + c;
    const double bt = pre * noise[k];
    beta[k] = bt;
    const double v = mom[k] / mi;
    const double c3 = (v + dt * forces[k] / m2) + bt / m2;
    pos[k] = sc * pos[k] + c2 * c3;
  }
}

__global__ void k_npt_iso_sys_post(int n_sys, const int32_t* __restrict__ sys_start,
                                   const double* __restrict__ kT,
                                   const double* __restrict__ p_ext,
                                   const double* __restrict__ cell_alpha, double dt,
                                   const double* __restrict__ stress,
                                   const double* __restrict__ ref_cell,
                                   const double* __restrict__ eps, double* __restrict__ eps_dot,
                                   const double* __restrict__ cell_masses,
                                   const double* __restrict__ sw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double n = (double)(sys_start[s + 1] - sys_start[s]);
  const double kt = kT[s];
  const double vol = det33(ref_cell + 9 * (int64_t)s) * exp(3.0 * eps[s]);
  const double* sg = stress + 9 * (int64_t)s;
  const double p_vir = -((sg[0] + sg[4]) + sg[8]) / 3.0;
  const double f_new = 3.0 * vol * ((p_vir + n * kt / vol) - p_ext[s]);
  const double q = cell_masses[s];
  const double ca = cell_alpha[s];
  const double a = (1.0 - (ca * dt) / (2.0 * q)) / (1.0 + (ca * dt) / (2.0 * q));
  const double b = 1.0 / (1.0 + (ca * dt) / (2.0 * q));
  const double f_n = sw[NPT_SW * s + 0];
  const double beta_c = sw[NPT_SW * s + 1];
  eps_dot[s] = (a * eps_dot[s] + dt * ((a * f_n) + f_new) / (2.0 * q)) + b * beta_c / q;
}

__global__ void k_npt_iso_atoms_post(int64_t n_atoms, const int32_t* __restrict__ atom_sys,
                                     double* __restrict__ mom, const double* __restrict__ f_old,
                                     const double* __restrict__ f_new,
                                     const double* __restrict__ masses,
                                     const double* __restrict__ alpha, double dt,
                                     const double* __restrict__ beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  const double al = alpha[atom_sys[i]];
  const double mi = masses[i];
  const double m2 = 2.0 * mi;
  const double a = (1.0 - (al * dt) / m2) / (1.0 + (al * dt) / m2);
  const double b = 1.0 / (1.0 + (al * dt) / m2);
  for (int c = 0; c < 3; ++c) {
    const int64_t k = 3 * i + c;
    const double v = mom[k] / mi;
    const double vn = (a * v + dt * ((a * f_old[k]) + f_new[k]) / m2) + b * beta[k] / mi;
    mom[k] = vn * mi;
  }
}

// Anisotropic variant (npt.py:62-131, 409-650): one strain eps_k = ln(L_k /
// L_k0) per Cartesian axis, driven by P_kk.  Per system: Q as above, the
// strain forces F_k = V (P_kk - P_ext,k) with P_kk = -sigma_kk + n kT / V
// (V = det of the current cell), GJF strain positions, the cell rows scaled
// by exp(eps_k) (row k of a column-vector cell holds the k-th Cartesian
// component of every lattice vector).
__global__ void k_npt_aniso_sys_pre(int n_sys, const int32_t* __restrict__ sys_start,
                                    const double* __restrict__ kT,
                                    const double* __restrict__ p_ext,
                                    const double* __restrict__ cell_alpha,
                                    const double* __restrict__ b_tau, double dt,
                                    const double* __restrict__ stress,
                                    const double* __restrict__ ref_cell,
                                    double* __restrict__ cell_col, double* __restrict__ eps,
                                    const double* __restrict__ eps_dot,
                                    double* __restrict__ cell_masses,
                                    const double* __restrict__ cell_noise,
                                    double* __restrict__ sw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double n = (double)(sys_start[s + 1] - sys_start[s]);
  const double kt = kT[s];
  const double q = (n + 1.0) * kt * (b_tau[s] * b_tau[s]);
  cell_masses[s] = q;
  const double vol = det33(cell_col + 9 * (int64_t)s);
  const double* sg = stress + 9 * (int64_t)s;
  const double ca = cell_alpha[s];
  const double q2 = 2.0 * q;
  const double b = 1.0 / (1.0 + (ca * dt) / q2);
  const double pre = sqrt(2.0 * ca * kt * dt);
  const double* h0 = ref_cell + 9 * (int64_t)s;
  double* sws = sw + NPT_SW * (int64_t)s;
  for (int k = 0; k < 3; ++k) {
    const double f_n = vol * ((-sg[4 * k] + n * kt / vol) - p_ext[3 * s + k]);
    const double beta_c = pre * cell_noise[3 * s + k];
    const double e0 = eps[3 * s + k];
    const double e1 =
        ((e0 + b * dt * eps_dot[3 * s + k]) + b * dt * dt * f_n / q2) + b * dt * beta_c / q2;
    eps[3 * s + k] = e1;
    const double sc = exp(e1);
    for (int c = 0; c < 3; ++c) cell_col[9 * (int64_t)s + 3 * k + c] = sc * h0[3 * k + c];
    sws[k] = f_n;
    sws[3 + k] = beta_c;
    sws[6 + k] = exp(e1 - e0);
  }
}

__global__ void k_npt_aniso_atoms_pre(int64_t n_atoms, const int32_t* __restrict__ atom_sys,
                                      double* __restrict__ pos, const double* __restrict__ mom,
                                      const double* __restrict__ forces,
                                      const double* __restrict__ masses,
                                      const double* __restrict__ kT,
                                      const double* __restrict__ alpha, double dt,
                                      const double* __restrict__ noise,
                                      const double* __restrict__ sw, double* __restrict__ beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  const int s = atom_sys[i];
  const double al = alpha[s];
  const double mi = masses[i];
  const double pre = sqrt(2.0 * al * kT[s] * dt);
  const double m2 = 2.0 * mi;
  const double b = 1.0 / (1.0 + (al * dt) / (2.0 * mi));
  for (int c = 0; c < 3; ++c) {
    const int64_t k = 3 * i + c;
    const double sc = sw[NPT_SW * s + 6 + c];
    const double bt = pre * noise[k];
    beta[k] = bt;
    const double v = mom[k] / mi;
    const double c3 = (v + dt * forces[k] / m2) + bt / m2;
    pos[k] = sc * pos[k] + ((2.0 * sc / (sc + 1.0)) * b * dt) * c3;
  }
}

__global__ void k_npt_aniso_sys_post(int n_sys, const int32_t* __restrict__ sys_start,
                                     const double* __restrict__ kT,
                                     const double* __restrict__ p_ext,
                                     const double* __restrict__ cell_alpha, double dt,
                                     const double* __restrict__ stress,
                                     const double* __restrict__ cell_col,
                                     double* __restrict__ eps_dot,
                                     const double* __restrict__ cell_masses,
                                     const double* __restrict__ sw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double n = (double)(sys_start[s + 1] - sys_start[s]);
  const double kt = kT[s];
  const double vol = det33(cell_col + 9 * (int64_t)s);
  const double* sg = stress + 9 * (int64_t)s;
  const double q = cell_masses[s];
  const double ca = cell_alpha[s];
  const double a = (1.0 - (ca * dt) / (2.0 * q)) / (1.0 + (ca * dt) / (2.0 * q));
  const double b = 1.0 / (1.0 + (ca * dt) / (2.0 * q));
  const double* sws = sw + NPT_SW * (int64_t)s;
  for (int k = 0; k < 3; ++k) {
    const double f_new = vol * ((-sg[4 * k] + n * kt / vol) - p_ext[3 * s + k]);
    eps_dot[3 * s + k] =
        (a * eps_dot[3 * s + k] + dt * ((a * sws[k]) + f_new) / (2.0 * q)) + b * sws[3 + k] / q;
  }
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int64_t tsx_npt_ws_bytes(int64_t n_atoms, int32_t n_systems) {
  return (int64_t)sizeof(double) * (NPT_SW * (int64_t)(n_systems > 0 ? n_systems : 1) +
                                    3 * (n_atoms > 0 ? n_atoms : 1));
}

int tsx_npt_langevin_iso_pre(int32_t n_systems, const int32_t* sys_start,
                             const int32_t* atom_sys, int64_t n_atoms, double* pos,
                             const double* mom, const double* forces, const double* masses,
                             const double* kT, const double* p_ext, const double* alpha,
                             const double* cell_alpha, const double* b_tau, double dt,
                             const double* stress, const double* ref_cell, double* eps,
                             const double* eps_dot, double* cell_masses, double* cell_col,
                             const double* cell_noise, const double* noise, void* ws,
                             void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(ws && kT && p_ext && alpha && cell_alpha && b_tau && stress && ref_cell && eps &&
                  eps_dot && cell_masses && cell_col && cell_noise,
              "every per-system array and the workspace are required");
  hipStream_t st = (hipStream_t)stream;
  double* sw = (double*)ws;
  double* beta = sw + NPT_SW * (int64_t)n_systems;
  k_npt_iso_sys_pre<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, sys_start, kT, p_ext, cell_alpha, b_tau, dt, stress, ref_cell, eps, eps_dot,
      cell_masses, cell_col, cell_noise, sw);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  TSX_REQUIRE(pos && mom && forces && masses && noise && atom_sys, "per-atom arrays required");
  k_npt_iso_atoms_pre<<<blocks_for(n_atoms, 256), 256, 0, st>>>(
      n_atoms, atom_sys, pos, mom, forces, masses, kT, alpha, dt, noise, sw, beta);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_npt_langevin_iso_post(int32_t n_systems, const int32_t* sys_start,
                              const int32_t* atom_sys, int64_t n_atoms, double* mom,
                              const double* f_old, const double* f_new, const double* masses,
                              const double* kT, const double* p_ext, const double* alpha,
                              const double* cell_alpha, double dt, const double* stress,
                              const double* ref_cell, const double* eps, double* eps_dot,
                              const double* cell_masses, const void* ws, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(ws && stress && eps_dot, "workspace, stress and eps_dot are required");
  hipStream_t st = (hipStream_t)stream;
  const double* sw = (const double*)ws;
  const double* beta = sw + NPT_SW * (int64_t)n_systems;
  k_npt_iso_sys_post<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, sys_start, kT, p_ext, cell_alpha, dt, stress, ref_cell, eps, eps_dot,
      cell_masses, sw);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  k_npt_iso_atoms_post<<<blocks_for(n_atoms, 256), 256, 0, st>>>(
      n_atoms, atom_sys, mom, f_old, f_new, masses, alpha, dt, beta);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_npt_langevin_aniso_pre(int32_t n_systems, const int32_t* sys_start,
                               const int32_t* atom_sys, int64_t n_atoms, double* pos,
                               const double* mom, const double* forces, const double* masses,
                               const double* kT, const double* p_ext, const double* alpha,
                               const double* cell_alpha, const double* b_tau, double dt,
                               const double* stress, const double* ref_cell, double* cell_col,
                               double* eps, const double* eps_dot, double* cell_masses,
                               const double* cell_noise, const double* noise, void* ws,
                               void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(ws && kT && p_ext && alpha && cell_alpha && b_tau && stress && ref_cell &&
                  cell_col && eps && eps_dot && cell_masses && cell_noise,
              "every per-system array and the workspace are required");
  hipStream_t st = (hipStream_t)stream;
  double* sw = (double*)ws;
  double* beta = sw + NPT_SW * (int64_t)n_systems;
  k_npt_aniso_sys_pre<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, sys_start, kT, p_ext, cell_alpha, b_tau, dt, stress, ref_cell, cell_col, eps,
      eps_dot, cell_masses, cell_noise, sw);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  TSX_REQUIRE(pos && mom && forces && masses && noise && atom_sys, "per-atom arrays required");
  k_npt_aniso_atoms_pre<<<blocks_for(n_atoms, 256), 256, 0, st>>>(
      n_atoms, atom_sys, pos, mom, forces, masses, kT, alpha, dt, noise, sw, beta);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_npt_langevin_aniso_post(int32_t n_systems, const int32_t* sys_start,
                                const int32_t* atom_sys, int64_t n_atoms, double* mom,
                                const double* f_old, const double* f_new, const double* masses,
                                const double* kT, const double* p_ext, const double* alpha,
                                const double* cell_alpha, double dt, const double* stress,
                                const double* cell_col, double* eps_dot,
                                const double* cell_masses, const void* ws, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(ws && stress && cell_col && eps_dot, "workspace, stress, cell and eps_dot required");
  hipStream_t st = (hipStream_t)stream;
  const double* sw = (const double*)ws;
  const double* beta = sw + NPT_SW * (int64_t)n_systems;
  k_npt_aniso_sys_post<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, sys_start, kT, p_ext, cell_alpha, dt, stress, cell_col, eps_dot, cell_masses,
      sw);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  k_npt_iso_atoms_post<<<blocks_for(n_atoms, 256), 256, 0, st>>>(
      n_atoms, atom_sys, mom, f_old, f_new, masses, alpha, dt, beta);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
