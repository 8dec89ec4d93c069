// Pair energies / forces / virials over the ELL neighbor list (gfx950).
//
// Replaces PairPotentialModel.forward (torch_sim/models/pair_potential.py:
// 130-464): instead of materialising dr [P,3], an autograd graph over P pairs
// and index_add scatters, each atom's row of the list is walked by ``tpa``
// lanes of one wave64; the displacement is formed in FP64 in the reference
// order (r_j - r_i) + S.L, the cutoff mask is the FP64 test d^2 < rc^2, the
// pair function runs in FP32 and the force, energy and virial accumulate in
// FP64 registers.  Full lists need no atomics (each row owns its atom);
// half lists add the j-side with FP64 atomics.
#include "tsx_pair.h"

namespace tsx {

constexpr int MAX_SPECIES = 8;

template <int KIND, int TPA, bool HALF, bool VIRIAL, int PS>
__global__ __launch_bounds__(256) void k_pair(
    const double* __restrict__ params, int ns, const int32_t* __restrict__ species,
    const double* __restrict__ pos, const double* __restrict__ cell_col,
    const int32_t* __restrict__ sys_start, const int32_t* __restrict__ atom_sys,
    const uint32_t* __restrict__ nbr, const int32_t* __restrict__ count,
    const int32_t* __restrict__ order, int64_t n_atoms, int jbits, int64_t ld, double rc2,
    double* __restrict__ forces, double* __restrict__ e_atom, double* __restrict__ w_atom) {
  __shared__ PairP sp[MAX_SPECIES * MAX_SPECIES];
  for (int k = threadIdx.x; k < ns * ns; k += blockDim.x)
    sp[k] = make_pair_params(KIND, params[3 * k], params[3 * k + 1], params[3 * k + 2]);
  __syncthreads();

  const int64_t gid = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t t = gid / TPA;
  const int lane = (int)(gid % TPA);
  const bool valid = t < n_atoms;
  if (!HALF) {
    // full list: the shared row walker (no atomics)
    const int64_t tt = valid ? t : 0;
    const int i = order[tt];
    const int s = atom_sys[i];
    const double* pi = pos + PS * (int64_t)i;
    const int ti = ns > 1 ? (species ? species[i] : (int)__double_as_longlong(pi[3])) : 0;
    const RowAcc a = accumulate_row<KIND, TPA, VIRIAL, PS>(
        sp, ns, species, pos, cell_col + 9 * (int64_t)s, sys_start[s], pi[0], pi[1], pi[2], ti,
        nbr, tt, valid ? count[tt] : 0, lane, jbits, ld, rc2);
    if (!valid || lane != 0) return;
    forces[3 * (int64_t)i] = a.fx;
    forces[3 * (int64_t)i + 1] = a.fy;
    forces[3 * (int64_t)i + 2] = a.fz;
    e_atom[i] = a.e;
    if (VIRIAL)
      for (int c = 0; c < 6; ++c) w_atom[6 * (int64_t)i + c] = a.w[c];
    return;
  }
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
  int i = 0;
  if (valid) {
    i = order[t];
    const int s = atom_sys[i];
    const int base = sys_start[s];
    const double* C = cell_col + 9 * (int64_t)s;
    // row vectors a_k = column k of the SimState cell
    const double L00 = C[0], L01 = C[3], L02 = C[6];
    const double L10 = C[1], L11 = C[4], L12 = C[7];
    const double L20 = C[2], L21 = C[5], L22 = C[8];
    const double xi = pos[PS * (int64_t)i], yi = pos[PS * (int64_t)i + 1],
                 zi = pos[PS * (int64_t)i + 2];
    const int ti = ns > 1 ? species[i] : 0;
    const int sbits = (32 - jbits) / 3;
    const uint32_t jmask = (1u << jbits) - 1u;
    const int cnt = count[t];
    for (int n = lane; n < cnt; n += TPA) {
      const uint32_t w = nbr[ell_idx(t, n, TPA, ld)];
      const int j = base + (int)(w & jmask);
      const int sx = nl_shift(w, jbits, sbits, 0);
      const int sy = nl_shift(w, jbits, sbits, 1);
      const int sz = nl_shift(w, jbits, sbits, 2);
      double dx = pos[PS * (int64_t)j] - xi;
      double dy = pos[PS * (int64_t)j + 1] - yi;
      double dz = pos[PS * (int64_t)j + 2] - zi;
      if ((sx | sy | sz) != 0) {
        const double fxs = (double)sx, fys = (double)sy, fzs = (double)sz;
        dx += fxs * L00 + fys * L10 + fzs * L20;
        dy += fxs * L01 + fys * L11 + fzs * L21;
        dz += fxs * L02 + fys * L12 + fzs * L22;
      }
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < rc2)) continue;
      const int tj = ns > 1 ? species[j] : 0;
      float e, fs;
      pair_eval<KIND>((float)d2, sp[ti * ns + tj], e, fs);
      const double f = (double)fs;
      const double fvx = f * dx, fvy = f * dy, fvz = f * dz;
      fx -= fvx;
      fy -= fvy;
      fz -= fvz;
      const double he = 0.5 * (double)e;
      ei += he;
      if (VIRIAL) {
        w0 += 0.5 * dx * fvx;
        w1 += 0.5 * dy * fvy;
        w2 += 0.5 * dz * fvz;
        w3 += 0.5 * dy * fvz;
        w4 += 0.5 * dx * fvz;
        w5 += 0.5 * dx * fvy;
      }
      if (HALF) {
        atomicAdd(&forces[3 * (int64_t)j], fvx);
        atomicAdd(&forces[3 * (int64_t)j + 1], fvy);
        atomicAdd(&forces[3 * (int64_t)j + 2], fvz);
        atomicAdd(&e_atom[j], he);
        if (VIRIAL) {
          atomicAdd(&w_atom[6 * (int64_t)j + 0], 0.5 * dx * fvx);
          atomicAdd(&w_atom[6 * (int64_t)j + 1], 0.5 * dy * fvy);
          atomicAdd(&w_atom[6 * (int64_t)j + 2], 0.5 * dz * fvz);
          atomicAdd(&w_atom[6 * (int64_t)j + 3], 0.5 * dy * fvz);
          atomicAdd(&w_atom[6 * (int64_t)j + 4], 0.5 * dx * fvz);
          atomicAdd(&w_atom[6 * (int64_t)j + 5], 0.5 * dx * fvy);
        }
      }
    }
  }
  // combine the TPA lanes of each atom (consecutive lanes of one wave)
#pragma unroll
  for (int off = TPA / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 64);
    fy += __shfl_xor(fy, off, 64);
    fz += __shfl_xor(fz, off, 64);
    ei += __shfl_xor(ei, off, 64);
    if (VIRIAL) {
      w0 += __shfl_xor(w0, off, 64);
      w1 += __shfl_xor(w1, off, 64);
      w2 += __shfl_xor(w2, off, 64);
      w3 += __shfl_xor(w3, off, 64);
      w4 += __shfl_xor(w4, off, 64);
      w5 += __shfl_xor(w5, off, 64);
    }
  }
  if (!valid || lane != 0) return;
  if (HALF) {
    atomicAdd(&forces[3 * (int64_t)i], fx);
    atomicAdd(&forces[3 * (int64_t)i + 1], fy);
    atomicAdd(&forces[3 * (int64_t)i + 2], fz);
    atomicAdd(&e_atom[i], ei);
    if (VIRIAL) {
      atomicAdd(&w_atom[6 * (int64_t)i + 0], w0);
      atomicAdd(&w_atom[6 * (int64_t)i + 1], w1);
      atomicAdd(&w_atom[6 * (int64_t)i + 2], w2);
      atomicAdd(&w_atom[6 * (int64_t)i + 3], w3);
      atomicAdd(&w_atom[6 * (int64_t)i + 4], w4);
      atomicAdd(&w_atom[6 * (int64_t)i + 5], w5);
    }
  } else {
    forces[3 * (int64_t)i] = fx;
    forces[3 * (int64_t)i + 1] = fy;
    forces[3 * (int64_t)i + 2] = fz;
    e_atom[i] = ei;
    if (VIRIAL) {
      double* wa = w_atom + 6 * (int64_t)i;
      wa[0] = w0;
      wa[1] = w1;
      wa[2] = w2;
      wa[3] = w3;
      wa[4] = w4;
      wa[5] = w5;
    }
  }
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

template <int KIND, int TPA, bool HALF, bool VIRIAL, int PS>
static void launch_pair(hipStream_t st, int nblk, const double* params, int ns,
                        const int32_t* species, const double* pos, const double* cell,
                        const int32_t* sys_start, const int32_t* atom_sys, const uint32_t* nbr,
                        const int32_t* count, const int32_t* order, int64_t n, int jbits,
                        int64_t ld, double rc2, double* f, double* e, double* w) {
  k_pair<KIND, TPA, HALF, VIRIAL, PS><<<nblk, 256, 0, st>>>(
      params, ns, species, pos, cell, sys_start, atom_sys, nbr, count, order, n, jbits, ld, rc2,
      f, e, w);
}

template <int KIND, int TPA, bool HALF, bool VIRIAL>
static int dispatch_ps(int ps, hipStream_t st, int nblk, const double* params, int ns,
                       const int32_t* species, const double* pos, const double* cell,
                       const int32_t* sys_start, const int32_t* atom_sys, const uint32_t* nbr,
                       const int32_t* count, const int32_t* order, int64_t n, int jbits,
                       int64_t ld, double rc2, double* f, double* e, double* w) {
  if (ps == 3)
    launch_pair<KIND, TPA, HALF, VIRIAL, 3>(st, nblk, params, ns, species, pos, cell, sys_start,
                                           atom_sys, nbr, count, order, n, jbits, ld, rc2, f, e,
                                           w);
  else
    launch_pair<KIND, TPA, HALF, VIRIAL, 4>(st, nblk, params, ns, species, pos, cell, sys_start,
                                           atom_sys, nbr, count, order, n, jbits, ld, rc2, f, e,
                                           w);
  return 0;
}

template <int KIND, int TPA>
static int dispatch_flags(bool half, bool vir, int ps, hipStream_t st, int nblk,
                          const double* params, int ns, const int32_t* species,
                          const double* pos, const double* cell, const int32_t* sys_start,
                          const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                          const int32_t* order, int64_t n, int jbits, int64_t ld, double rc2,
                          double* f, double* e, double* w) {
#define TSX_ARGS ps, st, nblk, params, ns, species, pos, cell, sys_start, atom_sys, nbr, count, \
                 order, n, jbits, ld, rc2, f, e, w
  if (half) {
    if (vir) return dispatch_ps<KIND, TPA, true, true>(TSX_ARGS);
    return dispatch_ps<KIND, TPA, true, false>(TSX_ARGS);
  }
  if (vir) return dispatch_ps<KIND, TPA, false, true>(TSX_ARGS);
  return dispatch_ps<KIND, TPA, false, false>(TSX_ARGS);
#undef TSX_ARGS
}

template <int KIND>
static int dispatch_tpa(int tpa, bool half, bool vir, int ps, hipStream_t st, int nblk,
                        const double* params, int ns, const int32_t* species, const double* pos,
                        const double* cell, const int32_t* sys_start, const int32_t* atom_sys,
                        const uint32_t* nbr, const int32_t* count, const int32_t* order,
                        int64_t n, int jbits, int64_t ld, double rc2, double* f, double* e,
                        double* w) {
#define TSX_ARGS half, vir, ps, st, nblk, params, ns, species, pos, cell, sys_start, atom_sys, \
                 nbr, count, order, n, jbits, ld, rc2, f, e, w
  switch (tpa) {
    case 1: return dispatch_flags<KIND, 1>(TSX_ARGS);
    case 2: return dispatch_flags<KIND, 2>(TSX_ARGS);
    case 4: return dispatch_flags<KIND, 4>(TSX_ARGS);
    case 8: return dispatch_flags<KIND, 8>(TSX_ARGS);
    default: return -1;
  }
#undef TSX_ARGS
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int tsx_pair_forces(int32_t kind, const double* params, int32_t n_species,
                    const int32_t* species, const double* pos, int32_t pos_stride,
                    const double* cell_col, const int32_t* sys_start, const int32_t* atom_sys,
                    const uint32_t* nbr, const int32_t* count, const int32_t* order,
                    int64_t n_atoms, int32_t jbits, int32_t tpa, int64_t ld, double cutoff,
                    int32_t half, double* forces, double* e_atom, double* w_atom,
                    void* stream) {
  TSX_REQUIRE(kind >= TSX_LJ && kind <= TSX_PARTICLE_LIFE, "unknown pair kind %d", kind);
  TSX_REQUIRE(n_species >= 1 && n_species <= MAX_SPECIES, "n_species must be in [1, %d]",
              MAX_SPECIES);
  TSX_REQUIRE(n_species == 1 || species != nullptr || (pos_stride == 4 && !half),
              "species required for n_species > 1 unless packed in stride-4 positions");
  TSX_REQUIRE(pos_stride == 3 || pos_stride == 4, "pos_stride must be 3 or 4");
  TSX_REQUIRE(tpa == 1 || tpa == 2 || tpa == 4 || tpa == 8, "tpa must be 1, 2, 4 or 8");
  TSX_REQUIRE(jbits >= 1 && jbits <= 26, "jbits must be in [1, 26]");
  TSX_REQUIRE(forces && e_atom, "forces and e_atom are required");
  if (n_atoms == 0) return TSX_OK;
  hipStream_t st = (hipStream_t)stream;
  if (half) {
    k_zero<<<blocks_for(3 * n_atoms, 256), 256, 0, st>>>(forces, 3 * n_atoms);
    k_zero<<<blocks_for(n_atoms, 256), 256, 0, st>>>(e_atom, n_atoms);
    if (w_atom) k_zero<<<blocks_for(6 * n_atoms, 256), 256, 0, st>>>(w_atom, 6 * n_atoms);
    TSX_CHECK_LAUNCH();
  }
  const int nblk = blocks_for(n_atoms * tpa, 256);
  const double rc2 = cutoff * cutoff;
  const bool vir = w_atom != nullptr;
#define TSX_ARGS tpa, half != 0, vir, pos_stride, st, nblk, params, n_species, species, pos, \
                 cell_col, sys_start, atom_sys, nbr, count, order, n_atoms, jbits, ld, rc2,  \
                 forces, e_atom, w_atom
  int rc;
  if (kind == TSX_LJ)
    rc = dispatch_tpa<TSX_LJ>(TSX_ARGS);
  else if (kind == TSX_MORSE)
    rc = dispatch_tpa<TSX_MORSE>(TSX_ARGS);
  else if (kind == TSX_SOFT_SPHERE)
    rc = dispatch_tpa<TSX_SOFT_SPHERE>(TSX_ARGS);
  else
    rc = dispatch_tpa<TSX_PARTICLE_LIFE>(TSX_ARGS);
#undef TSX_ARGS
  TSX_REQUIRE(rc == 0, "unsupported launch configuration");
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
