// Backward kernels of the pair models: vector-Jacobian products of the forces
// (force-matching losses, Hessian-vector products for the NVE adjoint) and
// of the energy, with respect to positions, the cell and the per-species-pair
// parameters (sigma, epsilon, alpha).
//
// Notation, for the ordered pair (i, j, S) of the full list:
//   d = r_j - r_i + S.L,  r = |d|,  u(r) the pair energy,  g = u'(r) / r,
//   F_i = sum_j g d,  E = 1/2 sum_(i,j) u.
// Force VJP for a cotangent w [N,3]:
//   d<w,F>/dr_i     = sum_j H(d) (w_j - w_i),   H = g I + (u'' - g) / r^2 d d^T
//   d<w,F>/dL[a,c]  = sum_(i,j) S_a (H w_i)_c          (row-vector cell)
//   d<w,F>/dtheta   = sum_(i,j) (w_i . d) dg/dtheta
// Energy gradient (w == NULL), weighted per system by ebar (or 1):
//   dE/dr_i = -F_i,  dE/dL[a,c] = 1/2 sum_(i,j) S_a g d_c,  dE/dtheta = 1/2 sum du/dtheta.
// This is what autograd of the reference's half-list forward
// (torch_sim/models/pair_potential.py:351-464, create_graph=True) computes.
//
// Each atom's row is walked by tpa lanes; every sum is accumulated in FP64 in
// a fixed order (lane shuffles, block tree, block partials summed in order),
// so gradients are bitwise reproducible.  The pair derivatives are FP32.
#include "tsx_pair.h"

namespace tsx {

constexpr int GRAD_THREADS = 256;
constexpr int GRAD_MAX_NS = 4;        // species with parameter gradients (buckets per thread)
constexpr int MAX_SPECIES_GRAD = 8;   // species for position / cell gradients

// u, u', u'' and the parameter derivatives of u and u' for one pair type.
struct PairD {
  float u, d1, d2;
  float du[3];   // d u / d(sigma, eps, alpha)
  float dd1[3];  // d u' / d(sigma, eps, alpha)
};

template <int KIND>
__device__ __forceinline__ PairD pair_grad_terms(float r, float sig, float eps, float alp) {
  PairD o;
  if (KIND == TSX_LJ) {
    // u = 4 eps (s12 - s6), s6 = (sigma / r)^6
    const float ir = frcp(r);
    const float sr = sig * ir;
    const float s2 = sr * sr;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    o.u = 4.f * eps * (s12 - s6);
    o.d1 = 24.f * eps * ir * (s6 - 2.f * s12);
    o.d2 = 24.f * eps * ir * ir * (26.f * s12 - 7.f * s6);
    const float isg = frcp(sig);
    o.du[0] = 24.f * eps * isg * (2.f * s12 - s6);
    o.du[1] = 4.f * (s12 - s6);
    o.du[2] = 0.f;
    o.dd1[0] = 144.f * eps * isg * ir * (s6 - 4.f * s12);
    o.dd1[1] = 24.f * ir * (s6 - 2.f * s12);
    o.dd1[2] = 0.f;
  } else if (KIND == TSX_MORSE) {
    // u = eps (x^2 - 2x), x = exp(-alpha (r - sigma))
    const float rm = r - sig;
    const float x = __expf(-alp * rm);
    const float omx = 1.f - x;
    o.u = eps * (x * x - 2.f * x);
    o.d1 = 2.f * eps * alp * x * omx;
    o.d2 = 2.f * eps * alp * alp * x * (2.f * x - 1.f);
    o.du[0] = -o.d1;
    o.du[1] = x * x - 2.f * x;
    o.du[2] = 2.f * eps * rm * x * omx;
    o.dd1[0] = -o.d2;
    o.dd1[1] = 2.f * alp * x * omx;
    o.dd1[2] = 2.f * eps * x * omx - 2.f * eps * alp * rm * x * (1.f - 2.f * x);
  } else {
    // u = eps / alpha t^alpha for r < sigma, t = 1 - r / sigma
    if (r < sig) {
      const float isg = frcp(sig);
      const float t = 1.f - r * isg;
      const float lt = t > 0.f ? __logf(t) : 0.f;
      const float tam2 = t > 0.f ? __expf((alp - 2.f) * lt) : 0.f;  // t^(alpha-2)
      const float tam1 = tam2 * t;
      const float ta = tam1 * t;
      const float ia = frcp(alp);
      o.u = eps * ia * ta;
      o.d1 = -eps * isg * tam1;
      o.d2 = eps * isg * isg * (alp - 1.f) * tam2;
      o.du[0] = eps * tam1 * r * isg * isg;
      o.du[1] = ta * ia;
      o.du[2] = eps * ia * ta * (lt - ia);
      o.dd1[0] = eps * isg * isg * tam2 * (t - (alp - 1.f) * r * isg);
      o.dd1[1] = -tam1 * isg;
      o.dd1[2] = -eps * isg * tam1 * lt;
    } else {
      o.u = o.d1 = o.d2 = 0.f;
      for (int k = 0; k < 3; ++k) o.du[k] = o.dd1[k] = 0.f;
    }
  }
  return o;
}

template <int KIND, int TPA, bool VJP, bool CELL, bool THETA, int PS>
__global__ __launch_bounds__(GRAD_THREADS) void k_pair_grad(
    const double* __restrict__ params, int ns, const int32_t* __restrict__ species,
    const double* __restrict__ pos, const double* __restrict__ cell_col,
    const int32_t* __restrict__ sys_start, const int32_t* __restrict__ atom_sys,
    const uint32_t* __restrict__ nbr, const int32_t* __restrict__ count,
    const int32_t* __restrict__ order, int64_t n_atoms, int jbits, int64_t ld, double rc2,
    const double* __restrict__ w, const double* __restrict__ ebar, double* __restrict__ dpos,
    double* __restrict__ dcell_atom, double* __restrict__ dtheta_part) {
  __shared__ float spar[MAX_SPECIES_GRAD * MAX_SPECIES_GRAD * 3];
  __shared__ double red[GRAD_THREADS / 64][GRAD_MAX_NS * GRAD_MAX_NS * 3];
  for (int k = threadIdx.x; k < ns * ns * 3; k += blockDim.x) spar[k] = (float)params[k];
  __syncthreads();

  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = gid / TPA;
  const int lane = (int)(gid % TPA);
  const bool valid = t < n_atoms;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  double gc[9];
  for (int k = 0; k < 9; ++k) gc[k] = 0.0;
  double th[GRAD_MAX_NS][3];
  for (int b = 0; b < GRAD_MAX_NS; ++b)
    for (int k = 0; k < 3; ++k) th[b][k] = 0.0;
  int ti = 0, i = 0;
  if (valid) {
    i = order[t];
    const int s = atom_sys[i];
    const int base = sys_start[s];
    const double sw = ebar ? ebar[s] : 1.0;
    const double* C = cell_col + 9 * (int64_t)s;
    const double L00 = C[0], L01 = C[3], L02 = C[6];
    const double L10 = C[1], L11 = C[4], L12 = C[7];
    const double L20 = C[2], L21 = C[5], L22 = C[8];
    const double* pi = pos + PS * (int64_t)i;
    const double xi = pi[0], yi = pi[1], zi = pi[2];
    ti = ns > 1 ? species[i] : 0;
    double wix = 0.0, wiy = 0.0, wiz = 0.0;
    if (VJP) {
      wix = w[3 * (int64_t)i];
      wiy = w[3 * (int64_t)i + 1];
      wiz = w[3 * (int64_t)i + 2];
    }
    const int sbits = (32 - jbits) / 3;
    const uint32_t jmask = (1u << jbits) - 1u;
    const int cnt = count[t];
    for (int n = lane; n < cnt; n += TPA) {
      const uint32_t word = nbr[ell_idx(t, n, TPA, ld)];
      const int j = base + (int)(word & jmask);
      const int sx = nl_shift(word, jbits, sbits, 0);
      const int sy = nl_shift(word, jbits, sbits, 1);
      const int sz = nl_shift(word, jbits, sbits, 2);
      const double* pj = pos + PS * (int64_t)j;
      double dx = pj[0] - xi, dy = pj[1] - yi, dz = pj[2] - zi;
      const bool shifted = (sx | sy | sz) != 0;
      if (shifted) {
        const double fxs = (double)sx, fys = (double)sy, fzs = (double)sz;
        dx += fxs * L00 + fys * L10 + fzs * L20;
        dy += fxs * L01 + fys * L11 + fzs * L21;
        dz += fxs * L02 + fys * L12 + fzs * L22;
      }
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < rc2)) continue;
      const int tj = ns > 1 ? species[j] : 0;
      const float* q = spar + 3 * (ti * ns + tj);
      const float r = fsqrt((float)d2);
      if (!(r > 0.f)) continue;
      const PairD pd = pair_grad_terms<KIND>(r, q[0], q[1], q[2]);
      const float ir = frcp(r);
      const double g = (double)(pd.d1 * ir);
      if (VJP) {
        // H (w_j - w_i) and H w_i with H = g I + h d d^T
        const double h = (double)((pd.d2 - pd.d1 * ir) * ir * ir);
        const double ux = w[3 * (int64_t)j] - wix, uy = w[3 * (int64_t)j + 1] - wiy,
                     uz = w[3 * (int64_t)j + 2] - wiz;
        const double du = dx * ux + dy * uy + dz * uz;
        gx += g * ux + h * du * dx;
        gy += g * uy + h * du * dy;
        gz += g * uz + h * du * dz;
        const double wd = dx * wix + dy * wiy + dz * wiz;
        if (CELL && shifted) {
          const double hx = g * wix + h * wd * dx, hy = g * wiy + h * wd * dy,
                       hz = g * wiz + h * wd * dz;
          const double sa[3] = {(double)sx, (double)sy, (double)sz};
          for (int a = 0; a < 3; ++a) {
            gc[3 * a + 0] += sa[a] * hx;
            gc[3 * a + 1] += sa[a] * hy;
            gc[3 * a + 2] += sa[a] * hz;
          }
        }
        if (THETA) {
          const double c = wd * (double)ir;
#pragma unroll
          for (int b = 0; b < GRAD_MAX_NS; ++b)
            if (b == tj)
              for (int k = 0; k < 3; ++k) th[b][k] += c * (double)pd.dd1[k];
        }
      } else {
        // energy: dE/dr_i = -F_i = -sum_j g d
        gx -= sw * g * dx;
        gy -= sw * g * dy;
        gz -= sw * g * dz;
        if (CELL && shifted) {
          const double hg = 0.5 * sw * g;
          const double sa[3] = {(double)sx, (double)sy, (double)sz};
          for (int a = 0; a < 3; ++a) {
            gc[3 * a + 0] += sa[a] * hg * dx;
            gc[3 * a + 1] += sa[a] * hg * dy;
            gc[3 * a + 2] += sa[a] * hg * dz;
          }
        }
        if (THETA) {
#pragma unroll
          for (int b = 0; b < GRAD_MAX_NS; ++b)
            if (b == tj)
              for (int k = 0; k < 3; ++k) th[b][k] += 0.5 * sw * (double)pd.du[k];
        }
      }
    }
  }
  // combine the TPA lanes of each atom
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) {
    gx += __shfl_xor(gx, off, 64);
    gy += __shfl_xor(gy, off, 64);
    gz += __shfl_xor(gz, off, 64);
    if (CELL)
      for (int k = 0; k < 9; ++k) gc[k] += __shfl_xor(gc[k], off, 64);
  }
  if (valid && lane == 0) {
    if (dpos) {
      dpos[3 * (int64_t)i] = gx;
      dpos[3 * (int64_t)i + 1] = gy;
      dpos[3 * (int64_t)i + 2] = gz;
    }
    if (CELL)
      for (int k = 0; k < 9; ++k) dcell_atom[9 * (int64_t)i + k] = gc[k];
  }
  if (!THETA) return;
  // parameter buckets: row ti of [ns, ns, 3]; block sum in a fixed order
  const int nb = ns * ns * 3;
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < GRAD_MAX_NS; ++a) {
    if (a >= ns) break;
#pragma unroll
    for (int b = 0; b < GRAD_MAX_NS; ++b) {
      if (b >= ns) break;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double v = (valid && ti == a) ? th[b][k] : 0.0;
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) red[wid][(a * ns + b) * 3 + k] = v;
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    double v = 0.0;
    for (int q = 0; q < GRAD_THREADS / 64; ++q) v += red[q][k];
    dtheta_part[(int64_t)blockIdx.x * nb + k] = v;
  }
}

// Sum the block partials and symmetrise: out[a,b] = out[b,a] =
// (P[a,b] + P[b,a]) / 2, so that the sum of the two entries is the derivative
// with respect to the shared parameter of the unordered species pair.  One
// workgroup: thread t adds blocks t, t + 256, ... for every entry, then a
// fixed-order tree over the workgroup (deterministic).
constexpr int THETA_T = 1024;

// Parameter gradient from the per-block partials: one workgroup per
// component k (and its symmetric partner ks, the (b, a) entry), 1024
// threads striding over the blocks, fixed-order wave and block trees, so
// repeated calls give the same bits.  out[k] = (sum_k + sum_ks) / 2.
__global__ __launch_bounds__(THETA_T) void k_theta_reduce(const double* __restrict__ part,
                                                          int n_blocks, int ns,
                                                          double* __restrict__ out,
                                                          int accumulate) {
  const int nb = ns * ns * 3;
  const int k = blockIdx.x;
  if (k >= nb) return;
  const int a = k / (3 * ns), b = (k / 3) % ns, p = k % 3;
  const int ks = (b * ns + a) * 3 + p;
  __shared__ double red[2][THETA_T / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int blk = threadIdx.x; blk < n_blocks; blk += THETA_T) {
    const double* pb = part + (int64_t)blk * nb;
    s0 += pb[k];
    s1 += pb[ks];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (lane == 0) {
    red[0][wid] = s0;
    red[1][wid] = s1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int q = 0; q < THETA_T / 64; ++q) {
      t0 += red[0][q];
      t1 += red[1][q];
    }
    const double v = 0.5 * (t0 + t1);
    out[k] = accumulate ? out[k] + v : v;
  }
}

// Per-system fixed-order sum of per-atom [N,9] row-cell gradients into the
// column-cell gradient [S,3,3] (the transpose).
__global__ __launch_bounds__(256) void k_cell_reduce(const double* __restrict__ dcell_atom,
                                                     const int32_t* __restrict__ sys_start,
                                                     double* __restrict__ out, int accumulate) {
  __shared__ double sh[4];
  const int s = blockIdx.x;
  const int a0 = sys_start[s], a1 = sys_start[s + 1];
  for (int k = 0; k < 9; ++k) {
    double v = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) v += dcell_atom[9 * (int64_t)i + k];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
      const int a = k / 3, c = k % 3;  // row-cell entry (a, c) -> column-cell (c, a)
      double* o = out + 9 * (int64_t)s + 3 * c + a;
      *o = accumulate ? *o + v : v;
    }
  }
}

// One NVE adjoint update between two force VJPs (momentum form of velocity
// Verlet, see torch_sim/integrators/nve.py:74-130):
//   xbar += vjp;  pbar += (dt / m) xbar;  w = wscale * pbar
__global__ void k_nve_adjoint(double* __restrict__ xbar, double* __restrict__ pbar,
                              const double* __restrict__ vjp, const double* __restrict__ m,
                              double dt, double wscale, double* __restrict__ w, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * n) return;
  const double xb = xbar[k] + vjp[k];
  const double pb = pbar[k] + dt / m[k / 3] * xb;
  xbar[k] = xb;
  pbar[k] = pb;
  w[k] = wscale * pb;
}

template <int KIND, int TPA, bool VJP, bool CELL, bool THETA>
static void launch_grad(int ps, hipStream_t st, int nblk, const double* params, int ns,
                        const int32_t* species, const double* pos, const double* cell,
                        const int32_t* sys_start, const int32_t* atom_sys, const uint32_t* nbr,
                        const int32_t* count, const int32_t* order, int64_t n, int jbits,
                        int64_t ld, double rc2, const double* w, const double* ebar,
                        double* dpos, double* dcell, double* dth) {
  if (ps == 3)
    k_pair_grad<KIND, TPA, VJP, CELL, THETA, 3><<<nblk, GRAD_THREADS, 0, st>>>(
        params, ns, species, pos, cell, sys_start, atom_sys, nbr, count, order, n, jbits, ld,
        rc2, w, ebar, dpos, dcell, dth);
  else
    k_pair_grad<KIND, TPA, VJP, CELL, THETA, 4><<<nblk, GRAD_THREADS, 0, st>>>(
        params, ns, species, pos, cell, sys_start, atom_sys, nbr, count, order, n, jbits, ld,
        rc2, w, ebar, dpos, dcell, dth);
}

#define TSX_GARGS ps, st, nblk, params, ns, species, pos, cell, sys_start, atom_sys, nbr, \
                  count, order, n, jbits, ld, rc2, w, ebar, dpos, dcell, dth

template <int KIND, int TPA>
static void dispatch_mode(bool vjp, bool cellg, bool theta, int ps, hipStream_t st, int nblk,
                          const double* params, int ns, const int32_t* species,
                          const double* pos, const double* cell, const int32_t* sys_start,
                          const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                          const int32_t* order, int64_t n, int jbits, int64_t ld, double rc2,
                          const double* w, const double* ebar, double* dpos, double* dcell,
                          double* dth) {
  if (vjp) {
    if (cellg) {
      if (theta) launch_grad<KIND, TPA, true, true, true>(TSX_GARGS);
      else launch_grad<KIND, TPA, true, true, false>(TSX_GARGS);
    } else {
      if (theta) launch_grad<KIND, TPA, true, false, true>(TSX_GARGS);
      else launch_grad<KIND, TPA, true, false, false>(TSX_GARGS);
    }
  } else {
    if (cellg) {
      if (theta) launch_grad<KIND, TPA, false, true, true>(TSX_GARGS);
      else launch_grad<KIND, TPA, false, true, false>(TSX_GARGS);
    } else {
      if (theta) launch_grad<KIND, TPA, false, false, true>(TSX_GARGS);
      else launch_grad<KIND, TPA, false, false, false>(TSX_GARGS);
    }
  }
}

template <int KIND>
static int dispatch_tpa_grad(int tpa, bool vjp, bool cellg, bool theta, int ps, hipStream_t st,
                             int nblk, const double* params, int ns, const int32_t* species,
                             const double* pos, const double* cell, const int32_t* sys_start,
                             const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                             const int32_t* order, int64_t n, int jbits, int64_t ld, double rc2,
                             const double* w, const double* ebar, double* dpos, double* dcell,
                             double* dth) {
#define TSX_MARGS vjp, cellg, theta, TSX_GARGS
  switch (tpa) {
    case 1: dispatch_mode<KIND, 1>(TSX_MARGS); return 0;
    case 2: dispatch_mode<KIND, 2>(TSX_MARGS); return 0;
    case 4: dispatch_mode<KIND, 4>(TSX_MARGS); return 0;
    case 8: dispatch_mode<KIND, 8>(TSX_MARGS); return 0;
    default: return -1;
  }
#undef TSX_MARGS
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int64_t tsx_grad_ws_bytes(int64_t n_atoms, int32_t tpa, int32_t n_species) {
  if (tpa < 1) tpa = 1;
  const int64_t nblk = (n_atoms * tpa + GRAD_THREADS - 1) / GRAD_THREADS;
  const int64_t nb = (int64_t)n_species * n_species * 3;
  return (int64_t)sizeof(double) * ((nblk > 0 ? nblk : 1) * nb + 9 * n_atoms);
}

int tsx_pair_grad(int32_t kind, const double* params, int32_t n_species, const int32_t* species,
                  const double* pos, int32_t pos_stride, const double* cell_col,
                  const int32_t* sys_start, const int32_t* atom_sys, int32_t n_systems,
                  const uint32_t* nbr, const int32_t* count, const int32_t* order,
                  int64_t n_atoms, int32_t jbits, int32_t tpa, int64_t ld, double cutoff,
                  const double* w, const double* ebar, double* dpos, double* dcell_out,
                  double* dparams, int32_t accumulate, void* ws, void* stream) {
  TSX_REQUIRE(kind >= TSX_LJ && kind <= TSX_SOFT_SPHERE, "unknown pair kind %d", kind);
  TSX_REQUIRE(n_species >= 1 && n_species <= MAX_SPECIES_GRAD, "n_species must be in [1, %d]",
              MAX_SPECIES_GRAD);
  TSX_REQUIRE(!dparams || n_species <= GRAD_MAX_NS,
              "parameter gradients support at most %d species", GRAD_MAX_NS);
  TSX_REQUIRE(n_species == 1 || species != nullptr, "species required for n_species > 1");
  TSX_REQUIRE(pos_stride == 3 || pos_stride == 4, "pos_stride must be 3 or 4");
  TSX_REQUIRE(tpa == 1 || tpa == 2 || tpa == 4 || tpa == 8, "tpa must be 1, 2, 4 or 8");
  TSX_REQUIRE(jbits >= 1 && jbits <= 26, "jbits must be in [1, 26]");
  TSX_REQUIRE(!(dcell_out || dparams) || ws, "workspace of tsx_grad_ws_bytes required");
  TSX_REQUIRE(!dcell_out || n_systems > 0, "n_systems required for the cell gradient");
  hipStream_t st = (hipStream_t)stream;
  if (n_atoms == 0) return TSX_OK;
  const int nblk = blocks_for(n_atoms * tpa, GRAD_THREADS);
  const int nb = n_species * n_species * 3;
  // kernel outputs: block partials of dparams and per-atom [N,9] row-cell
  // gradients, both in the workspace; reduced below into dparams / dcell_out
  double* dth = dparams ? (double*)ws : nullptr;
  double* dcell = dcell_out ? (double*)ws + (int64_t)nblk * nb : nullptr;
  const double rc2 = cutoff * cutoff;
  const int ps = pos_stride;
  const int ns = n_species;
  const int64_t n = n_atoms;
  const double* cell = cell_col;
#define TSX_DARGS tpa, w != nullptr, dcell != nullptr, dparams != nullptr, TSX_GARGS
  int rc;
  if (kind == TSX_LJ)
    rc = dispatch_tpa_grad<TSX_LJ>(TSX_DARGS);
  else if (kind == TSX_MORSE)
    rc = dispatch_tpa_grad<TSX_MORSE>(TSX_DARGS);
  else
    rc = dispatch_tpa_grad<TSX_SOFT_SPHERE>(TSX_DARGS);
#undef TSX_DARGS
  TSX_REQUIRE(rc == 0, "unsupported launch configuration");
  TSX_CHECK_LAUNCH();
  if (dparams) {
    k_theta_reduce<<<ns * ns * 3, THETA_T, 0, st>>>(dth, nblk, ns, dparams, accumulate);
    TSX_CHECK_LAUNCH();
  }
  if (dcell_out) {
    k_cell_reduce<<<n_systems, 256, 0, st>>>(dcell, sys_start, dcell_out, accumulate);
    TSX_CHECK_LAUNCH();
  }
  return TSX_OK;
}

int tsx_nve_adjoint_update(double* xbar, double* pbar, const double* vjp, const double* masses,
                           double dt, double wscale, double* w, int64_t n_atoms, void* stream) {
  if (n_atoms == 0) return TSX_OK;
  k_nve_adjoint<<<blocks_for(3 * n_atoms, 256), 256, 0, (hipStream_t)stream>>>(
      xbar, pbar, vjp, masses, dt, wscale, w, n_atoms);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
