// Thermostats and barostats beyond Langevin (gfx950): the Nose-Hoover chain
// (torch_sim/integrators/md.py:249-565), NVT velocity rescaling
// (nvt.py:565-800), the isotropic MTK Nose-Hoover barostat (npt.py:1110-2060)
// and the stochastic cell-rescaling barostats (npt.py:2062-2986).
//
// Per-atom work is one pass over momenta (and positions) per update, HBM
// bound; per-system work (chains, volumes, 3x3 cell maps) runs one thread
// per system in FP64.  The reductions a step needs (sum p^2 / m, or the six
// components of sum p p^T / m) are fused into the pass that updates the
// momenta they read: chunks of 2048 atoms counted from each system's first
// atom, one workgroup per (system, chunk), a fixed-order tree, and a second
// kernel adding the chunk partials when a system spans several chunks -- so
// a system sums bitwise alike alone or inside any batch.
//
// Random numbers are the caller's (drawn from the state's generator in the
// reference's order); the kernels only consume them.
#include "tsx_common.h"
#include "tsx_mat33.h"

namespace tsx {

constexpr int TH_CH = 2048;
constexpr int TH_THREADS = 256;
constexpr int TH_WAVES = TH_THREADS / 64;
constexpr int NHC_MAXM = 16;
constexpr int NPTNH_F = 8;   // per-system factors: a2 s2 e2 em1 e1 s1 v alpha
constexpr int CR_XF = 27;    // per-system maps: R, Vp, Mx (row-major 3x3 each)
constexpr int CR_NOISE = 11; // per-system normals: ra, R[9] (or R[3]), rb

// p <- a_s p + h F (either part optional) or the MTK momentum map
// p <- p a2 + ((h F) s2) e2 (npt.py:1402-1457), then per-system sums.
enum { UPD_SCALE_KICK = 0, UPD_IL2 = 1 };

template <int UPD, int V>
__global__ __launch_bounds__(TH_THREADS) void k_mom_reduce(
    int chunks, const int32_t* __restrict__ sys_start, double* __restrict__ mom,
    const double* __restrict__ masses, const double* __restrict__ scale,
    const double* __restrict__ forces, double h, const double* __restrict__ fac,
    double* __restrict__ part, double* __restrict__ out, double factor) {
  __shared__ double sh[TH_WAVES][V];
  const int s = blockIdx.x / chunks, c = blockIdx.x % chunks;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t lo = (int64_t)sys_start[s] + (int64_t)c * TH_CH;
  const int64_t end = sys_start[s + 1];
  const int64_t hi = lo + TH_CH < end ? lo + TH_CH : end;
  const bool write = UPD == UPD_IL2 || scale || forces;
  double as = 1.0, s2 = 0.0, e2 = 0.0;
  if (UPD == UPD_SCALE_KICK && scale) as = scale[s];
  if (UPD == UPD_IL2) {
    as = fac[NPTNH_F * s + 0];
    s2 = fac[NPTNH_F * s + 1];
    e2 = fac[NPTNH_F * s + 2];
  }
  double acc[V];
  for (int q = 0; q < V; ++q) acc[q] = 0.0;
  for (int64_t k = lo + threadIdx.x; k < hi; k += TH_THREADS) {
    double p[3];
    for (int d = 0; d < 3; ++d) {
      double v = mom[3 * k + d];
      if (UPD == UPD_SCALE_KICK) {
        if (scale) v = as * v;
        if (forces) v = v + forces[3 * k + d] * h;
      } else {
        v = v * as + ((h * forces[3 * k + d]) * s2) * e2;
      }
      p[d] = v;
    }
    if (write)
      for (int d = 0; d < 3; ++d) mom[3 * k + d] = p[d];
    if (!out) continue;
    const double m = masses[k];
    if (V == 1) {
      acc[0] += (p[0] * p[0] / m + p[1] * p[1] / m) + p[2] * p[2] / m;
    } else {
      const double im = 1.0 / m;
      acc[0] += p[0] * p[0] * im;
      acc[1] += p[1] * p[1] * im;
      acc[2] += p[2] * p[2] * im;
      acc[3] += p[1] * p[2] * im;
      acc[4] += p[0] * p[2] * im;
      acc[5] += p[0] * p[1] * im;
    }
  }
  if (!out) return;  // uniform over the launch
  for (int q = 0; q < V; ++q) {
    const double t = wave_sum(acc[q]);
    if (lane == 0) sh[wid][q] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < V; ++q) {
      double t = 0.0;
      for (int w = 0; w < TH_WAVES; ++w) t += sh[w][q];
      if (chunks == 1)
        out[(int64_t)V * s + q] = factor * t;
      else
        part[((int64_t)s * chunks + c) * V + q] = t;
    }
  }
}

template <int V>
__global__ __launch_bounds__(TH_THREADS) void k_mom_finish(const int32_t* __restrict__ sys_start,
                                                           int chunks,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out,
                                                           double factor) {
  __shared__ double sh[TH_WAVES][V];
  const int s = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n_s = (int64_t)sys_start[s + 1] - sys_start[s];
  int nc = (int)((n_s + TH_CH - 1) / TH_CH);
  nc = nc < chunks ? nc : chunks;
  double acc[V];
  for (int q = 0; q < V; ++q) acc[q] = 0.0;
  for (int c = threadIdx.x; c < nc; c += TH_THREADS)
    for (int q = 0; q < V; ++q) acc[q] += part[((int64_t)s * chunks + c) * V + q];
  for (int q = 0; q < V; ++q) {
    const double t = wave_sum(acc[q]);
    if (lane == 0) sh[wid][q] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < V; ++q) {
      double t = 0.0;
      for (int w = 0; w < TH_WAVES; ++w) t += sh[w][q];
      out[(int64_t)V * s + q] = factor * t;
    }
}

// p <- a_s p + h F, then x <- x + (p / m) dt (md.py:169-208 after a
// per-system momentum scale)
__global__ void k_scale_kick_drift(int64_t n, const int32_t* __restrict__ atom_sys,
                                   double* __restrict__ x, double* __restrict__ p,
                                   const double* __restrict__ scale,
                                   const double* __restrict__ f, double h,
                                   const double* __restrict__ m, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double as = scale ? scale[atom_sys[i]] : 1.0;
  const double mi = m[i];
  for (int d = 0; d < 3; ++d) {
    double v = p[3 * i + d];
    if (scale) v = as * v;
    v = v + f[3 * i + d] * h;
    p[3 * i + d] = v;
    x[3 * i + d] = x[3 * i + d] + (v / mi) * dt;
  }
}

// Suzuki-Yoshida composition weights for 1, 3, 5 and 7 steps (Yoshida,
// Phys. Lett. A 150, 262 (1990); Suzuki, Proc. Japan Acad. 69, 161 (1993)),
// held at float32 precision as the reference's default-dtype weight tensors
// hold them (md.py:309-333; the product with the float64 step promotes the
// rounded weight).
__device__ __forceinline__ double sy_weight(int nsy, int k) {
  float w = 1.0f;
  if (nsy == 3) {
    w = k == 1 ? -0.657963087177502f : 0.828981543588751f;
  } else if (nsy == 5) {
    w = k == 2 ? -0.186929716880426f : 0.2967324292201065f;
  } else if (nsy == 7) {
    const float t[7] = {0.784513610477560f, 0.235573213359357f, -1.17767998417887f,
                        1.31518632068391f,  -1.17767998417887f, 0.235573213359357f,
                        0.784513610477560f};
    w = t[k];
  }
  return (double)w;
}

// Half a step of a Nose-Hoover chain per system (md.py:419-535): chain_steps
// x sy_steps substeps of the Martyna et al. direct translation.  ke holds the
// coupled kinetic energy and is scaled with the momenta; scale_out receives
// the product of the substeps' momentum scales; a per-system momentum array
// (the barostat's) is scaled substep by substep as the reference scales it.
__global__ void k_nhc_half(int n_sys, int L, int nc, int nsy, double dt,
                           const double* __restrict__ kT, const double* __restrict__ tau,
                           const double* __restrict__ dof, double* __restrict__ xi,
                           double* __restrict__ pxi, double* __restrict__ Q,
                           double* __restrict__ ke, int update_mass,
                           double* __restrict__ scale_out, double* __restrict__ sys_mom) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  double x[NHC_MAXM], p[NHC_MAXM], q[NHC_MAXM];
  const int64_t o = (int64_t)L * s;
  for (int m = 0; m < L; ++m) {
    x[m] = xi[o + m];
    p[m] = pxi[o + m];
    q[m] = Q[o + m];
  }
  const double kt = kT[s], nf = dof[s];
  if (update_mass) {  // md.py:537-565
    const double t2 = tau[s] * tau[s];
    for (int m = 0; m < L; ++m) q[m] = kt * t2;
    q[0] = q[0] * nf;
  }
  const int M = L - 1;
  double K = ke[s], lam = 1.0;
  double pm = sys_mom ? sys_mom[s] : 0.0;
  auto substep = [&](double delta) {
    const double d2 = delta / 2.0, d4 = d2 / 2.0, d8 = d4 / 2.0;
    double G;
    if (M > 0) {
      G = p[M - 1] * p[M - 1] / q[M - 1] - kt;
      p[M] = p[M] + d4 * G;
    }
    for (int m = M - 1; m > 0; --m) {
      G = p[m - 1] * p[m - 1] / q[m - 1] - kt;
      const double sc = exp(-d8 * p[m + 1] / q[m + 1]);
      p[m] = sc * (sc * p[m] + d4 * G);
    }
    G = 2.0 * K - nf * kt;
    {
      const double sc = M > 0 ? exp(-d8 * p[1] / q[1]) : 1.0;
      p[0] = sc * (sc * p[0] + d4 * G);
    }
    const double sc0 = exp(-d2 * p[0] / q[0]);
    K = K * (sc0 * sc0);
    lam = lam * sc0;
    pm = pm * sc0;
    for (int m = 0; m < L; ++m) x[m] = x[m] + d2 * p[m] / q[m];
    G = 2.0 * K - nf * kt;
    for (int m = 0; m < M; ++m) {
      const double sc = exp(-d8 * p[m + 1] / q[m + 1]);
      p[m] = sc * (sc * p[m] + d4 * G);
      G = p[m] * p[m] / q[m] - kt;
    }
    p[M] = p[M] + d4 * G;
  };
  if (nc == 1 && nsy == 1) {
    substep(dt);
  } else {
    const double delta = dt / (double)nc;
    for (int k = 0; k < nc * nsy; ++k) substep(delta * sy_weight(nsy, k % nsy));
  }
  for (int m = 0; m < L; ++m) {
    xi[o + m] = x[m];
    pxi[o + m] = p[m];
    Q[o + m] = q[m];
  }
  ke[s] = K;
  if (scale_out) scale_out[s] = lam;
  if (sys_mom) sys_mom[s] = pm;
}

// CSVR momentum scale per system (nvt.py:594-654): ke2 = sum p^2 / m.
__global__ void k_vrescale(int n_sys, const double* __restrict__ ke2,
                           const double* __restrict__ dof, const double* __restrict__ kT,
                           const double* __restrict__ tau, double dt,
                           const double* __restrict__ r1, const double* __restrict__ r2,
                           double* __restrict__ lam) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double nf = dof[s];
  const double ckt = ke2[s] / nf;
  const double ke_old = nf * ckt / 2.0;
  const double ke_new = nf * kT[s] / 2.0;
  const double c1 = exp(-dt / tau[s]);
  const double c2 = (1.0 - c1) * ke_new / ke_old / nf;
  const double a = r1[s];
  const double sc = (c1 + c2 * (a * a + r2[s])) + 2.0 * a * sqrt(c1 * c2);
  lam[s] = sqrt(sc);
}

// ----------------------------------------------------- MTK Nose-Hoover NPT
// sinh(x)/x by its Taylor series to x^10 (npt.py:1312-1343)
__device__ __forceinline__ double sinhx_x(double x) {
  return ((((1.0 + pow(x, 2.0) / 6.0) + pow(x, 4.0) / 120.0) + pow(x, 6.0) / 5040.0) +
          pow(x, 8.0) / 362880.0) +
         pow(x, 10.0) / 39916800.0;
}

// cell force alpha 2 K - tr(sigma) V - 3 P V (npt.py:1460-1529)
__device__ __forceinline__ double nh_cell_force(double alpha, double ke, const double* sg,
                                                double vol, double pext) {
  const double tr = (sg[0] + sg[4]) + sg[8];
  return ((alpha * 2.0) * ke - tr * vol) - pext * vol * 3.0;
}

// First half of the inner step (npt.py:1532-1590) per system: cell mass
// 3 (n + 1) kT b_tau^2, cell force on the thermostatted momenta (ke2 =
// sum p^2 / m), p_eps += dt/2 G, the strain eps += v dt and its cell, and
// the per-system factors of the two particle maps.
__global__ void k_nptnh_pre_sys(int n_sys, const int32_t* __restrict__ sys_start,
                                const double* __restrict__ kT, const double* __restrict__ b_tau,
                                const double* __restrict__ dof, const double* __restrict__ ke2,
                                const double* __restrict__ stress,
                                const double* __restrict__ p_ext,
                                const double* __restrict__ ref_cell, double* __restrict__ eps,
                                double* __restrict__ cell_mom, double* __restrict__ cell_mass,
                                double* __restrict__ cell_col, double dt,
                                double* __restrict__ fac) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double n = (double)(sys_start[s + 1] - sys_start[s]);
  const double w = (3.0 * (n + 1.0)) * kT[s] * (b_tau[s] * b_tau[s]);
  cell_mass[s] = w;
  const double* h0 = ref_cell + 9 * (int64_t)s;
  const double v0 = det33(h0);
  const double e0 = eps[s];
  const double vol = v0 * exp(3.0 * e0);
  const double alpha = 1.0 + 3.0 / dof[s];
  const double dt2 = dt / 2.0;
  const double g = nh_cell_force(alpha, 0.5 * ke2[s], stress + 9 * (int64_t)s, vol, p_ext[s]);
  const double pm = cell_mom[s] + dt2 * g;
  cell_mom[s] = pm;
  const double v = pm / w;
  const double x2 = alpha * v * dt2, x2h = x2 / 2.0;
  const double x1 = v * dt, x1h = x1 / 2.0;
  double* f = fac + NPTNH_F * (int64_t)s;
  f[0] = exp(-x2);
  f[1] = sinhx_x(x2h);
  f[2] = exp(-x2h);
  f[3] = exp(x1) - 1.0;
  f[4] = exp(x1h);
  f[5] = sinhx_x(x1h);
  f[6] = v;
  f[7] = alpha;
  const double e1 = e0 + v * dt;
  eps[s] = e1;
  const double sc = pow((v0 * exp(3.0 * e1)) / v0, 1.0 / 3.0);
  for (int k = 0; k < 9; ++k) cell_col[9 * (int64_t)s + k] = sc * h0[k];
}

// per atom: p <- exp(iL2 dt/2) p, then x <- exp(iL1 dt) x with v = p / m
__global__ void k_nptnh_pre_atoms(int64_t n, const int32_t* __restrict__ atom_sys,
                                  double* __restrict__ x, double* __restrict__ p,
                                  const double* __restrict__ f, const double* __restrict__ m,
                                  const double* __restrict__ fac, double dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = fac + NPTNH_F * (int64_t)atom_sys[i];
  const double dt2 = dt / 2.0;
  const double mi = m[i];
  for (int d = 0; d < 3; ++d) {
    const int64_t k = 3 * i + d;
    const double pn = p[k] * c[0] + ((dt2 * f[k]) * c[1]) * c[2];
    p[k] = pn;
    const double vel = pn / mi;
    x[k] = x[k] + (x[k] * c[3] + ((dt * vel) * c[4]) * c[5]);
  }
}

// Second half of the inner step per system (npt.py:1600-1640 and the
// kinetic energies of npt.py:1940-1960): the cell force at the new stress,
// volume and momenta, p_eps += dt/2 G, the thermostat's kinetic energy and
// the barostat's p_eps^2 / (2 W).
__global__ void k_nptnh_post_sys(int n_sys, const double* __restrict__ ke2,
                                 const double* __restrict__ stress,
                                 const double* __restrict__ p_ext,
                                 const double* __restrict__ ref_cell,
                                 const double* __restrict__ eps, double* __restrict__ cell_mom,
                                 const double* __restrict__ cell_mass,
                                 const double* __restrict__ fac, double dt,
                                 double* __restrict__ ke_therm, double* __restrict__ ke_baro) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double vol = det33(ref_cell + 9 * (int64_t)s) * exp(3.0 * eps[s]);
  const double ke = 0.5 * ke2[s];
  const double g =
      nh_cell_force(fac[NPTNH_F * (int64_t)s + 7], ke, stress + 9 * (int64_t)s, vol, p_ext[s]);
  const double pm = cell_mom[s] + (dt / 2.0) * g;
  cell_mom[s] = pm;
  ke_therm[s] = ke;
  ke_baro[s] = pm * pm / (2.0 * cell_mass[s]);
}

// ------------------------------------------------- stochastic cell rescaling
// lower-triangular form of a box whose columns are its vectors
// (npt.py:2093-2126): row k of the result is vector k in the rotated frame
__device__ M33 gram_schmidt_rows(const M33& b) {
  const double a0 = b.a[0], a1 = b.a[3], a2 = b.a[6];
  const double b0 = b.a[1], b1 = b.a[4], b2 = b.a[7];
  const double c0 = b.a[2], c1 = b.a[5], c2 = b.a[8];
  M33 o = m33_zero();
  o.a[0] = sqrt((a0 * a0 + a1 * a1) + a2 * a2);
  o.a[3] = ((a0 * b0 + a1 * b1) + a2 * b2) / o.a[0];
  o.a[4] = sqrt(((b0 * b0 + b1 * b1) + b2 * b2) - o.a[3] * o.a[3]);
  o.a[6] = ((a0 * c0 + a1 * c1) + a2 * c2) / o.a[0];
  o.a[7] = (((b0 * c0 + b1 * c1) + b2 * c2) - o.a[6] * o.a[3]) / o.a[4];
  o.a[8] = sqrt((((c0 * c0 + c1 * c1) + c2 * c2) - o.a[6] * o.a[6]) - o.a[7] * o.a[7]);
  return o;
}

// One barostat step per system (npt.py:2186-2492).  mode 0 isotropic,
// 1 anisotropic (diagonal), 2 triclinic, 3 triclinic with the average
// (equipartition) kinetic pressure.  kin = sum p p^T / m of the kicked
// momenta [S,6] (xx yy zz yz xz xy); the cell is updated in place and xf
// receives R (positions), Vp (momenta) and Mx (the momentum term of the
// position update) for tsx_crescale_apply.
__global__ void k_crescale_sys(int n_sys, int mode, const double* __restrict__ kin,
                               const double* __restrict__ stress, double* __restrict__ cell,
                               const double* __restrict__ kT, const double* __restrict__ beta,
                               const double* __restrict__ tau_p, double dt,
                               const double* __restrict__ p_ext,
                               const double* __restrict__ p_ext_t,
                               const double* __restrict__ cell0_inv,
                               const double* __restrict__ vol0, const double* __restrict__ dof,
                               const double* __restrict__ noise, double* __restrict__ xf) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const int64_t o9 = 9 * (int64_t)s;
  const M33 C = m33_load(cell + o9);
  const M33 S = m33_load(stress + o9);
  const double vol = m33_det(C);
  const double kt = kT[s], bt = beta[s], tp = tau_p[s];
  const double* z = noise + CR_NOISE * (int64_t)s;
  // internal pressure tensor (quantities.py:156-198, npt.py:2346-2373)
  M33 P;
  if (mode == 3) {
    const double pre = dof[s] / 3.0 * kt / vol;
    for (int k = 0; k < 9; ++k) P.a[k] = (k % 4 == 0 ? pre : 0.0) - S.a[k];
  } else {
    const double* kn = kin + 6 * (int64_t)s;
    const double kk[9] = {kn[0], kn[5], kn[4], kn[5], kn[1], kn[3], kn[4], kn[3], kn[2]};
    for (int k = 0; k < 9; ++k) P.a[k] = kk[k] / vol - S.a[k];
  }
  const double sqv = sqrt(vol);
  const double trp = (P.a[0] + P.a[4]) + P.a[8];
  const double pr = sqrt(kt * bt * dt / (4.0 * tp));
  const double pf = bt * sqv / (2.0 * tp);
  M33 R = m33_eye(), Vp = m33_eye(), Mx = m33_eye();
  if (mode == 0) {
    const double chg = -pf * ((p_ext[s] - trp / 3.0) - kt / (2.0 * vol)) * dt +
                       sqrt(2.0) * pr * z[0];
    const double mu = pow((sqv + chg) / sqv, 2.0 / 3.0);
    R = m33_eye(mu);
    Vp = m33_eye(1.0 / mu);
    Mx = m33_eye(mu + 1.0 / mu);
    m33_store(m33_scale(C, mu), cell + o9);
  } else if (mode == 1) {
    const double pe = p_ext[s];
    double nsv = sqv + (-pf * ((pe - trp / 3.0) - kt / (2.0 * vol)) * dt / 2.0 + pr * z[0]);
    const double prm = sqrt(2.0 * bt * kt * dt / (3.0 * tp)) / nsv;
    const double mean = ((z[1] + z[2]) + z[3]) / 3.0;
    double dk[3];
    for (int k = 0; k < 3; ++k) {
      const double at = (bt / (3.0 * tp)) * (P.a[4 * k] - trp / 3.0);
      dk[k] = exp(at * dt + prm * (z[1 + k] - mean));
    }
    nsv += -pf * ((pe - trp / 3.0) - kt / (2.0 * vol)) * dt / 2.0 + pr * z[10];
    const double g = pow(nsv / sqv, 2.0 / 3.0);
    R = m33_zero();
    Vp = m33_zero();
    Mx = m33_zero();
    M33 Cn = C;
    for (int k = 0; k < 3; ++k) {
      const double r = dk[k] * g;
      R.a[4 * k] = r;
      Vp.a[4 * k] = 1.0 / r;
      Mx.a[4 * k] = r + 1.0 / r;
      for (int c = 0; c < 3; ++c) Cn.a[3 * k + c] = r * C.a[3 * k + c];
    }
    m33_store(Cn, cell + o9);
  } else {
    double eff = p_ext[s];
    M33 Pd = P;
    if (mode == 2 && p_ext_t) {  // deviatoric target stress (npt.py:2145-2183)
      const M33 E = m33_load(p_ext_t + o9);
      const double hyd = ((E.a[0] + E.a[4]) + E.a[8]) / 3.0;
      const M33 Sd = m33_sub(E, m33_eye(hyd));
      const M33 Ci = m33_load(cell0_inv + o9);
      const M33 sig = m33_scale(m33_mul(m33_mul(m33_t(Ci), Sd), Ci), vol0[s]);
      const M33 ssym = m33_scale(m33_add(sig, m33_t(sig)), 0.5);
      const M33 pdev = m33_scale(m33_mul(m33_mul(m33_t(C), ssym), C), 1.0 / vol);
      eff = hyd + ((pdev.a[0] + pdev.a[4]) + pdev.a[8]) / 3.0;
      Pd = m33_sub(P, pdev);
    }
    double nsv = sqv + (-pf * ((eff - trp / 3.0) - kt / (2.0 * vol)) * dt / 2.0 + pr * z[0]);
    const double prm = sqrt(2.0 * bt * kt * dt / (3.0 * tp)) / nsv;
    const double trd = (Pd.a[0] + Pd.a[4]) + Pd.a[8];
    const double trz = (z[1] + z[5]) + z[9];
    M33 A;
    for (int k = 0; k < 9; ++k) {
      const bool dg = k % 4 == 0;
      const double at = (bt / (3.0 * tp)) * (Pd.a[k] - (dg ? trd / 3.0 : 0.0));
      const double rt = z[1 + k] - (dg ? trz / 3.0 : 0.0);
      A.a[k] = at * dt + prm * rt;
    }
    const M33 D = gram_schmidt_rows(m33_expm(A));
    nsv += -pf * ((eff - trp / 3.0) - kt / (2.0 * vol)) * dt / 2.0 + pr * z[10];
    R = m33_scale(D, pow(nsv / sqv, 2.0 / 3.0));
    if (mode == 2) {
      Vp = m33_t(m33_inv(R));
      Mx = m33_add(Vp, R);
    } else {
      Vp = m33_eye();
      Mx = m33_add(m33_eye(), R);
    }
    m33_store(m33_mul(C, m33_t(R)), cell + o9);
  }
  double* x = xf + CR_XF * (int64_t)s;
  m33_store(R, x);
  m33_store(Vp, x + 9);
  m33_store(Mx, x + 18);
}

// per atom: x <- R x + (Mx p) dt / (2 m), p <- Vp p (npt.py:2266-2280)
__global__ void k_crescale_apply(int64_t n, const int32_t* __restrict__ atom_sys,
                                 double* __restrict__ x, double* __restrict__ p,
                                 const double* __restrict__ m, const double* __restrict__ xf,
                                 double dt, int scale_momenta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = xf + CR_XF * (int64_t)atom_sys[i];
  const double xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
  const double pi[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  const double m2 = 2.0 * m[i];
  for (int r = 0; r < 3; ++r) {
    const double rx = (c[3 * r] * xi[0] + c[3 * r + 1] * xi[1]) + c[3 * r + 2] * xi[2];
    const double mp = (c[18 + 3 * r] * pi[0] + c[18 + 3 * r + 1] * pi[1]) + c[18 + 3 * r + 2] * pi[2];
    x[3 * i + r] = rx + mp * dt / m2;
  }
  if (scale_momenta)
    for (int r = 0; r < 3; ++r)
      p[3 * i + r] = (c[9 + 3 * r] * pi[0] + c[9 + 3 * r + 1] * pi[1]) + c[9 + 3 * r + 2] * pi[2];
}

static int chunks_for(int32_t max_system_atoms) {
  return max_system_atoms > 0 ? (max_system_atoms + TH_CH - 1) / TH_CH : 1;
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int64_t tsx_mom_reduce_ws_bytes(int32_t max_system_atoms, int32_t n_systems) {
  return (int64_t)sizeof(double) * 6 * chunks_for(max_system_atoms) *
         (n_systems > 0 ? n_systems : 1);
}

int tsx_md_scale_kick_reduce(int32_t n_systems, const int32_t* sys_start,
                             int32_t max_system_atoms, double* mom, const double* masses,
                             const double* scale, const double* forces, double h,
                             int32_t tensor, double factor, double* out, void* ws,
                             void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(max_system_atoms >= 0, "max_system_atoms must be >= 0");
  TSX_REQUIRE(sys_start && mom, "sys_start and momenta are required");
  TSX_REQUIRE(!out || (masses && ws), "a reduction needs the masses and the workspace");
  hipStream_t st = (hipStream_t)stream;
  const int chunks = chunks_for(max_system_atoms);
  const int64_t nblk = (int64_t)n_systems * chunks;
  TSX_REQUIRE(nblk < (1ll << 31), "too many reduction blocks");
  double* part = (double*)ws;
  if (tensor)
    k_mom_reduce<UPD_SCALE_KICK, 6><<<(int)nblk, TH_THREADS, 0, st>>>(
        chunks, sys_start, mom, masses, scale, forces, h, nullptr, part, out, factor);
  else
    k_mom_reduce<UPD_SCALE_KICK, 1><<<(int)nblk, TH_THREADS, 0, st>>>(
        chunks, sys_start, mom, masses, scale, forces, h, nullptr, part, out, factor);
  TSX_CHECK_LAUNCH();
  if (out && chunks > 1) {
    if (tensor)
      k_mom_finish<6><<<n_systems, TH_THREADS, 0, st>>>(sys_start, chunks, part, out, factor);
    else
      k_mom_finish<1><<<n_systems, TH_THREADS, 0, st>>>(sys_start, chunks, part, out, factor);
    TSX_CHECK_LAUNCH();
  }
  return TSX_OK;
}

int tsx_md_scale_kick_drift(int64_t n_atoms, const int32_t* atom_sys, double* pos, double* mom,
                            const double* scale, const double* forces, double h,
                            const double* masses, double dt, void* stream) {
  if (n_atoms == 0) return TSX_OK;
  TSX_REQUIRE(pos && mom && forces && masses && (!scale || atom_sys),
              "positions, momenta, forces and masses are required");
  k_scale_kick_drift<<<blocks_for(n_atoms, 256), 256, 0, (hipStream_t)stream>>>(
      n_atoms, atom_sys, pos, mom, scale, forces, h, masses, dt);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_nhc_half_step(int32_t n_systems, int32_t chain_length, int32_t chain_steps,
                      int32_t sy_steps, double dt, const double* kT, const double* tau,
                      const double* dof, double* xi, double* pxi, double* masses, double* ke,
                      int32_t update_mass, double* scale_out, double* sys_mom, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(chain_length >= 1 && chain_length <= NHC_MAXM, "chain_length must be in [1, %d]",
              NHC_MAXM);
  TSX_REQUIRE(chain_steps >= 1, "chain_steps must be >= 1");
  TSX_REQUIRE(sy_steps == 1 || sy_steps == 3 || sy_steps == 5 || sy_steps == 7,
              "sy_steps must be 1, 3, 5 or 7");
  TSX_REQUIRE(kT && dof && xi && pxi && masses && ke && (!update_mass || tau),
              "chain arrays, kT, dof and ke are required");
  k_nhc_half<<<blocks_for(n_systems, 64), 64, 0, (hipStream_t)stream>>>(
      n_systems, chain_length, chain_steps, sy_steps, dt, kT, tau, dof, xi, pxi, masses, ke,
      update_mass, scale_out, sys_mom);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_vrescale_factor(int32_t n_systems, const double* ke2, const double* dof,
                        const double* kT, const double* tau, double dt, const double* r1,
                        const double* r2, double* lam, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(ke2 && dof && kT && tau && r1 && r2 && lam, "every per-system array is required");
  k_vrescale<<<blocks_for(n_systems, 64), 64, 0, (hipStream_t)stream>>>(n_systems, ke2, dof, kT,
                                                                        tau, dt, r1, r2, lam);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int64_t tsx_npt_nh_ws_bytes(int32_t n_systems) {
  return (int64_t)sizeof(double) * NPTNH_F * (n_systems > 0 ? n_systems : 1);
}

int tsx_npt_nh_pre(int32_t n_systems, const int32_t* sys_start, const int32_t* atom_sys,
                   int64_t n_atoms, double* pos, double* mom, const double* forces,
                   const double* masses, const double* kT, const double* b_tau,
                   const double* dof, const double* ke2, const double* stress,
                   const double* p_ext, const double* ref_cell, double* eps, double* cell_mom,
                   double* cell_mass, double* cell_col, double dt, void* ws, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(sys_start && kT && b_tau && dof && ke2 && stress && p_ext && ref_cell && eps &&
                  cell_mom && cell_mass && cell_col && ws,
              "every per-system array and the workspace are required");
  hipStream_t st = (hipStream_t)stream;
  double* fac = (double*)ws;
  k_nptnh_pre_sys<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, sys_start, kT, b_tau, dof, ke2, stress, p_ext, ref_cell, eps, cell_mom,
      cell_mass, cell_col, dt, fac);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  TSX_REQUIRE(atom_sys && pos && mom && forces && masses, "per-atom arrays required");
  k_nptnh_pre_atoms<<<blocks_for(n_atoms, 256), 256, 0, st>>>(n_atoms, atom_sys, pos, mom, forces,
                                                               masses, fac, dt);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_npt_nh_post(int32_t n_systems, const int32_t* sys_start, int32_t max_system_atoms,
                    double* mom, const double* forces, const double* masses,
                    const double* stress, const double* p_ext, const double* ref_cell,
                    const double* eps, double* cell_mom, const double* cell_mass, double dt,
                    double* ke2, double* ke_therm, double* ke_baro, const void* ws,
                    void* red_ws, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(max_system_atoms >= 0, "max_system_atoms must be >= 0");
  TSX_REQUIRE(sys_start && mom && forces && masses && stress && p_ext && ref_cell && eps &&
                  cell_mom && cell_mass && ke2 && ke_therm && ke_baro && ws && red_ws,
              "every array and both workspaces are required");
  hipStream_t st = (hipStream_t)stream;
  const double* fac = (const double*)ws;
  const int chunks = chunks_for(max_system_atoms);
  const int64_t nblk = (int64_t)n_systems * chunks;
  TSX_REQUIRE(nblk < (1ll << 31), "too many reduction blocks");
  double* part = (double*)red_ws;
  k_mom_reduce<UPD_IL2, 1><<<(int)nblk, TH_THREADS, 0, st>>>(
      chunks, sys_start, mom, masses, nullptr, forces, dt / 2.0, fac, part, ke2, 1.0);
  TSX_CHECK_LAUNCH();
  if (chunks > 1) {
    k_mom_finish<1><<<n_systems, TH_THREADS, 0, st>>>(sys_start, chunks, part, ke2, 1.0);
    TSX_CHECK_LAUNCH();
  }
  k_nptnh_post_sys<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, ke2, stress, p_ext, ref_cell, eps, cell_mom, cell_mass, fac, dt, ke_therm,
      ke_baro);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int64_t tsx_crescale_ws_bytes(int32_t n_systems) {
  return (int64_t)sizeof(double) * CR_XF * (n_systems > 0 ? n_systems : 1);
}

int tsx_crescale_step(int32_t n_systems, int32_t mode, const int32_t* atom_sys, int64_t n_atoms,
                      double* pos, double* mom, const double* masses, const double* kin,
                      const double* stress, double* cell_col, const double* kT,
                      const double* compressibility, const double* tau_p, double dt,
                      const double* p_ext, const double* p_ext_tensor, const double* cell0_inv,
                      const double* vol0, const double* dof, const double* noise, void* ws,
                      void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(mode >= 0 && mode <= 3, "mode must be 0 (isotropic), 1 (anisotropic), "
              "2 (triclinic) or 3 (triclinic, average kinetic energy)");
  TSX_REQUIRE(stress && cell_col && kT && compressibility && tau_p && p_ext && noise && ws,
              "stress, cell, kT, compressibility, tau_p, p_ext, noise and ws are required");
  TSX_REQUIRE(mode == 3 ? dof != nullptr : kin != nullptr,
              "the kinetic tensor (or, for the average mode, the dof) is required");
  TSX_REQUIRE(!p_ext_tensor || (mode == 2 && cell0_inv && vol0),
              "a target stress tensor needs the triclinic mode and the initial cell");
  hipStream_t st = (hipStream_t)stream;
  double* xf = (double*)ws;
  k_crescale_sys<<<blocks_for(n_systems, 64), 64, 0, st>>>(
      n_systems, mode, kin, stress, cell_col, kT, compressibility, tau_p, dt, p_ext,
      p_ext_tensor, cell0_inv, vol0, dof, noise, xf);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  TSX_REQUIRE(atom_sys && pos && mom && masses, "per-atom arrays required");
  k_crescale_apply<<<blocks_for(n_atoms, 256), 256, 0, st>>>(n_atoms, atom_sys, pos, mom, masses,
                                                              xf, dt, mode != 3);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
