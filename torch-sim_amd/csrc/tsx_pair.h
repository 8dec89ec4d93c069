// FP32 pair functions on FP64 displacement vectors, shared by the model,
// fused-step and backward kernels.
#pragma once

#include "tsx_common.h"

namespace tsx {

// Derived per-species-pair constants in FP32:
//   LJ:          a = sigma^2, b = 4 eps,  c = 24 eps
//   Morse:       a = sigma,   b = eps,    c = alpha,  d = 2 eps alpha,
//                e = sigma alpha log2(e), f = -alpha log2(e)
//                (exp(-alpha (r - sigma)) = exp2(f r + e): one FMA and v_exp)
//   soft sphere: a = sigma,   b = eps,    c = alpha,  d = eps / sigma
//   particle life (sigma, eps, alpha = outer radius s, amplitude A, inner
//   radius b): a = b, b = A, c = s, d = 1 / b, e = 1 / (1 - b)
struct PairP {
  float a, b, c, d, e, f;
};

__host__ __device__ inline PairP make_pair_params(int kind, double sigma, double eps,
                                                  double alpha) {
  PairP p;
  p.e = (float)(1.0 / sigma);
  p.f = 0.f;
  if (kind == TSX_PARTICLE_LIFE) {
    p.a = (float)alpha;
    p.b = (float)eps;
    p.c = (float)sigma;
    p.d = (float)(1.0 / alpha);
    p.e = (float)(1.0 / (1.0 - alpha));
  } else if (kind == TSX_LJ) {
    p.a = (float)(sigma * sigma);
    p.b = (float)(4.0 * eps);
    p.c = (float)(24.0 * eps);
    p.d = 0.f;
  } else if (kind == TSX_MORSE) {
    p.a = (float)sigma;
    p.b = (float)eps;
    p.c = (float)alpha;
    p.d = (float)(2.0 * eps * alpha);
    p.e = (float)(sigma * alpha * 1.4426950408889634);
    p.f = (float)(-alpha * 1.4426950408889634);
  } else {
    p.a = (float)sigma;
    p.b = (float)eps;
    p.c = (float)alpha;
    p.d = (float)(eps / sigma);
  }
  return p;
}

// Energy e and scalar fs with force vector f_ij = fs * dr_ij = -(dV/dr / r) dr_ij
// (the force on j due to i, pair_potential.py:426-428).
// Hardware reciprocal / square root (1 ulp): the pair arithmetic is FP32 by
// design, and r^2 > 0 is never denormal here.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

template <int KIND>
__device__ __forceinline__ void pair_eval(float r2, const PairP& p, float& e, float& fs) {
  if (KIND == TSX_LJ) {
    if (r2 > 0.f) {
      const float ir2 = frcp(r2);
      const float sr2 = p.a * ir2;
      const float sr6 = sr2 * sr2 * sr2;
      e = p.b * (sr6 * sr6 - sr6);
      fs = p.c * (2.f * sr6 * sr6 - sr6) * ir2;
    } else {
      e = 0.f;
      fs = 0.f;
    }
  } else if (KIND == TSX_MORSE) {
    if (r2 > 0.f) {
      // one reciprocal square root gives r = r2 / sqrt(r2) and 1 / r
      const float ir = frsq(r2);
      const float r = r2 * ir;
      const float x = __builtin_amdgcn_exp2f(fmaf(p.f, r, p.e));
      const float omx = 1.f - x;
      e = p.b * omx * omx - p.b;
      fs = -p.d * x * omx * ir;
    } else {
      e = 0.f;
      fs = 0.f;
    }
  } else if (KIND == TSX_PARTICLE_LIFE) {
    e = 0.f;
    if (r2 > 0.f) {
      const float ir = frsq(r2);
      const float r = r2 * ir;
      const float f = r < p.a ? r * p.d - 1.f
                    : r < p.c ? p.b * (1.f - fabsf(2.f * r - 1.f - p.a) * p.e)
                              : 0.f;
      fs = f * ir;
    } else {
      fs = 0.f;
    }
  } else {
    const float ir = r2 > 0.f ? frsq(r2) : 0.f;
    const float r = r2 * ir;
    if (r < p.a) {
      const float t = 1.f - r * p.e;
      if (p.c == 2.f) {
        e = 0.5f * p.b * t * t;
        fs = p.d * t * ir;
      } else {
        const float tm1 = powf(t, p.c - 1.f);
        e = p.b / p.c * tm1 * t;
        fs = p.d * tm1 * ir;
      }
    } else {
      e = 0.f;
      fs = 0.f;
    }
  }
}

// pair_eval for r2 > 0 without branches: the same operations, hence the same
// bits, as pair_eval where pair_eval finds r2 > 0.  Loops that evaluate every
// entry of a batch and discard some with a select call it with a positive
// stand-in r2 for those.
template <int KIND>
__device__ __forceinline__ void pair_eval_pos(float r2, const PairP& p, float& e, float& fs) {
  if (KIND == TSX_LJ) {
    const float ir2 = frcp(r2);
    const float sr2 = p.a * ir2;
    const float sr6 = sr2 * sr2 * sr2;
    e = p.b * (sr6 * sr6 - sr6);
    fs = p.c * (2.f * sr6 * sr6 - sr6) * ir2;
  } else if (KIND == TSX_MORSE) {
    const float ir = frsq(r2);
    const float r = r2 * ir;
    const float x = __builtin_amdgcn_exp2f(fmaf(p.f, r, p.e));
    const float omx = 1.f - x;
    e = p.b * omx * omx - p.b;
    fs = -p.d * x * omx * ir;
  } else if (KIND == TSX_PARTICLE_LIFE) {
    const float ir = frsq(r2);
    const float r = r2 * ir;
    const float f = r < p.a ? r * p.d - 1.f
                  : r < p.c ? p.b * (1.f - fabsf(2.f * r - 1.f - p.a) * p.e)
                            : 0.f;
    e = 0.f;
    fs = f * ir;
  } else {
    const float ir = frsq(r2);
    const float r = r2 * ir;
    const float t = 1.f - r * p.e;
    float ee, ff;
    if (p.c == 2.f) {
      ee = 0.5f * p.b * t * t;
      ff = p.d * t * ir;
    } else {
      const float tm1 = powf(t, p.c - 1.f);
      ee = p.b / p.c * tm1 * t;
      ff = p.d * tm1 * ir;
    }
    const bool inside = r < p.a;
    e = inside ? ee : 0.f;
    fs = inside ? ff : 0.f;
  }
}

// Second radial derivative helpers for the backward kernels:
//   phi'(r) and phi''(r) in FP32 together with dphi'/dsigma, dphi'/deps.
template <int KIND>
__device__ __forceinline__ void pair_derivs(float r, const PairP& p, float& d1, float& d2) {
  if (KIND == TSX_LJ) {
    // V = 4e (s^12 r^-12 - s^6 r^-6)
    const float ir = 1.f / r;
    const float sr2 = p.a * ir * ir;
    const float sr6 = sr2 * sr2 * sr2;
    const float e4 = p.b;
    d1 = e4 * (-12.f * sr6 * sr6 + 6.f * sr6) * ir;
    d2 = e4 * (156.f * sr6 * sr6 - 42.f * sr6) * ir * ir;
  } else if (KIND == TSX_MORSE) {
    const float x = __expf(-p.c * (r - p.a));
    d1 = 2.f * p.b * p.c * x * (1.f - x);
    d2 = 2.f * p.b * p.c * p.c * x * (2.f * x - 1.f);
  } else {
    if (r < p.a) {
      const float t = 1.f - r / p.a;
      const float tm2 = (p.c == 2.f) ? 1.f : powf(t, p.c - 2.f);
      d1 = -p.d * tm2 * t;
      d2 = p.d / p.a * (p.c - 1.f) * tm2;
    } else {
      d1 = 0.f;
      d2 = 0.f;
    }
  }
}

}  // namespace tsx

namespace tsx {

__device__ __forceinline__ int64_t ell_idx(int64_t t, int n, int tpa, int64_t ld) {
  return (int64_t)(n / tpa) * (ld * tpa) + t * tpa + (n % tpa);
}

struct RowAcc {
  double fx, fy, fz, e;
  double w[6];
};

// Walk this lane's share of row t (atom i) of the ELL list and accumulate the
// FP64 force on i, half the pair energies and half the pair virials.  The
// displacement is (r_j - r_i) + S.L in FP64 (pair_potential.py:181-183), the
// mask is d^2 < rc^2 (pair_potential.py:194), the pair function FP32.
// Species of atom j come from ``species`` (stride 1) or, when ``species`` is
// NULL and ns > 1, from the packed fourth coordinate of a stride-4 position.
// Entries are taken UNR at a time: all list words, then all positions are
// requested before any is used, so each lane keeps UNR gathers in flight.
constexpr int UNR = 4;

template <int KIND, int TPA, bool VIRIAL, int PS, bool MULTI>
__device__ __forceinline__ RowAcc accumulate_row_impl(
    const PairP* sp, int ns, const int32_t* __restrict__ species,
    const double* __restrict__ pos, const double* C, int base, double xi, double yi, double zi,
    int ti, const uint32_t* __restrict__ nbr, int64_t t, int cnt, int lane, int jbits,
    int64_t ld, double rc2) {
  RowAcc a;
  a.fx = a.fy = a.fz = a.e = 0.0;
  for (int c = 0; c < 6; ++c) a.w[c] = 0.0;
  const double L00 = C[0], L01 = C[3], L02 = C[6];
  const double L10 = C[1], L11 = C[4], L12 = C[7];
  const double L20 = C[2], L21 = C[5], L22 = C[8];
  const int sbits = (32 - jbits) / 3;
  const uint32_t jmask = (1u << jbits) - 1u;
  // the row's system starts at pbase: a neighbour is pbase + PS * j_local
  const double* __restrict__ pbase = pos + (int64_t)PS * base;
  const int32_t* __restrict__ sbase = species ? species + base : nullptr;
  const PairP p0 = sp[MULTI ? 0 : ti * ns];  // single species: parameters in registers
  // entry n = lane + TPA * k of this row lives at wp[k * stride]
  const uint32_t* wp = nbr + t * TPA + lane;
  const int64_t stride = ld * TPA;
  const int iters = cnt > lane ? (cnt - lane + TPA - 1) / TPA : 0;
  for (int k0 = 0; k0 < iters; k0 += UNR) {
    uint32_t w[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) w[u] = (k0 + u < iters) ? wp[(int64_t)(k0 + u) * stride] : 0u;
    double px[UNR], py[UNR], pz[UNR];
    int tj[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t jl = w[u] & jmask;
      const double* pj = pbase + PS * jl;
      px[u] = pj[0];
      py[u] = pj[1];
      pz[u] = pj[2];
      tj[u] = 0;
      if (MULTI) tj[u] = sbase ? sbase[jl] : (int)__double_as_longlong(pj[3]);
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if (k0 + u >= iters) break;
      double dx = px[u] - xi;
      double dy = py[u] - yi;
      double dz = pz[u] - zi;
      if ((w[u] >> jbits) != 0u) {  // S' != 0
        const double fxs = (double)nl_shift(w[u], jbits, sbits, 0);
        const double fys = (double)nl_shift(w[u], jbits, sbits, 1);
        const double fzs = (double)nl_shift(w[u], jbits, sbits, 2);
        dx += fxs * L00 + fys * L10 + fzs * L20;
        dy += fxs * L01 + fys * L11 + fzs * L21;
        dz += fxs * L02 + fys * L12 + fzs * L22;
      }
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < rc2)) continue;
      float e, fs;
      pair_eval<KIND>((float)d2, MULTI ? sp[ti * ns + tj[u]] : p0, e, fs);
      const double f = (double)fs;
      a.fx = fma(-f, dx, a.fx);
      a.fy = fma(-f, dy, a.fy);
      a.fz = fma(-f, dz, a.fz);
      a.e = fma(0.5, (double)e, a.e);
      if (VIRIAL) {
        const double hf = 0.5 * f;
        a.w[0] += hf * dx * dx;
        a.w[1] += hf * dy * dy;
        a.w[2] += hf * dz * dz;
        a.w[3] += hf * dy * dz;
        a.w[4] += hf * dx * dz;
        a.w[5] += hf * dx * dy;
      }
    }
  }
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) {
    a.fx += __shfl_xor(a.fx, off, 64);
    a.fy += __shfl_xor(a.fy, off, 64);
    a.fz += __shfl_xor(a.fz, off, 64);
    a.e += __shfl_xor(a.e, off, 64);
    if (VIRIAL)
      for (int c = 0; c < 6; ++c) a.w[c] += __shfl_xor(a.w[c], off, 64);
  }
  return a;
}

// Row walk dispatched on the number of species: one species runs a loop with
// its parameters held in registers, several species look the pair up per
// neighbour.
template <int KIND, int TPA, bool VIRIAL, int PS>
__device__ __forceinline__ RowAcc accumulate_row(
    const PairP* sp, int ns, const int32_t* __restrict__ species,
    const double* __restrict__ pos, const double* C, int base, double xi, double yi, double zi,
    int ti, const uint32_t* __restrict__ nbr, int64_t t, int cnt, int lane, int jbits,
    int64_t ld, double rc2) {
  if (ns > 1)
    return accumulate_row_impl<KIND, TPA, VIRIAL, PS, true>(sp, ns, species, pos, C, base, xi,
                                                            yi, zi, ti, nbr, t, cnt, lane, jbits,
                                                            ld, rc2);
  return accumulate_row_impl<KIND, TPA, VIRIAL, PS, false>(sp, ns, species, pos, C, base, xi, yi,
                                                           zi, ti, nbr, t, cnt, lane, jbits, ld,
                                                           rc2);
}


}  // namespace tsx
