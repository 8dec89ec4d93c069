// Batched FIRE steps (ASE and velocity-Verlet flavours) with the unit and
// Frechet cell filters.
//
// One workgroup per system.  A step is
//     tsx_fire_pre  ->  forces / energy / stress (any model)  ->  tsx_fire_post
// The pre kernel runs the per-system power / norm reductions, the dt / alpha /
// n_pos update, velocity mixing, the max-step clamp and the position and cell
// update (ASE flavour), or the first half kick and drift (VV flavour).  The
// post kernel computes the cell forces from the new stress (Frechet: one
// Frechet derivative of exp through the adjoint identity
//     sum_ij L(A, E_mn)_ij G_ij = L(A^T, G)_mn ),
// then, for the VV flavour, the second half kick, power and mixing, and the
// per-system max force for convergence checks.  All sums over a system's atoms
// are fixed-order block reductions.
//
// Reference: torch_sim/optimizers/fire.py:21-114 (init), 179-285 (_vv_fire_step),
// 288-499 (_ase_fire_step); torch_sim/optimizers/cell_filters.py:26-60, 137-178,
// 180-284, 353-399; torch_sim/math.py:748-778 (batched_vdot).
#include <float.h>

#include "tsx_common.h"
#include "tsx_mat33.h"

namespace tsx {

constexpr int FIRE_THREADS = 256;
constexpr int FIRE_WAVES = FIRE_THREADS / 64;

// Sum N doubles over the block; every thread gets the totals (fixed order).
template <int N>
__device__ __forceinline__ void block_sum_n(double* v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k * FIRE_WAVES + wid] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double t = 0.0;
    for (int w = 0; w < FIRE_WAVES; ++w) t += sh[k * FIRE_WAVES + w];
    v[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ double nan_to_num(double x) {
  if (isnan(x)) return 0.0;
  if (isinf(x)) return x > 0 ? DBL_MAX : -DBL_MAX;
  return x;
}

// column-convention cell [3,3] (row-major storage of the column cell)
__device__ __forceinline__ M33 load_cell(const double* p, int s) { return m33_load(p + 9 * (int64_t)s); }

// deform_grad(ref_row, cur_row) = solve(ref_row, cur_row)^T = cur_col * ref_col^-1
__device__ __forceinline__ M33 deform_grad_col(const M33& ref_col, const M33& cur_col) {
  return m33_mul(cur_col, m33_inv(ref_col));
}

// Per-system workspace row written by k_fire_sys and read by the atom kernels.
constexpr int SYSW = 40;
// [0] |dr|^2 of the atoms, [1] clamp denominator (0: no clamp), [2] dt,
// [3..11] m1, [12..20] m2 (r <- m2 (m1 r + dr)), [21..29] pull-back to the
// build cell, [30] squared list threshold (< 0: rebuild), [31] has-cell flag

// Per-system list validity after a cell change: the list built at cell Lb with
// radius r_list stays valid for cutoff rc while every pulled-back atom
// (Lb L^-1 r) is within thr of its position at the build, where
//     thr = (r_list - rc / sigma) / 2,   sigma = 1 - ||L Lb^-1 - I||_F
// is a lower bound on the smallest singular value of the cell map.
__device__ inline void list_check_setup(const M33& cell_new, const M33& cell_build, double r_list,
                                        double cutoff, double* w) {
  const M33 a = m33_mul(cell_new, m33_inv(cell_build));
  const double sig = 1.0 - m33_frob(m33_sub(a, m33_eye()));
  m33_store(m33_mul(cell_build, m33_inv(cell_new)), w + 21);
  if (!(sig > 0.0) || !(sig * r_list > cutoff)) {
    w[30] = -1.0;
  } else {
    const double thr = 0.5 * (r_list - cutoff / sig);
    w[30] = thr * thr;
  }
}

struct FireArgs {
  tsx_fire_params p;
  int n_sys;
  int64_t n_atoms;
  const int32_t* sys_start;
  const int32_t* atom_sys;
  double* pos;
  int pos_stride;
  double* vel;
  const double* forces;
  const double* masses;
  double* dt;
  double* alpha;
  int32_t* n_pos;
  double* cell;
  const double* ref_cell;
  const double* cell_factor;
  double* cell_pos;
  double* cell_vel;
  const double* cell_forces;
  const int32_t* active;
  const double* cell_build;
  const double* xref;
  double r_list, cutoff;
  int32_t* flags;
  double* ws;
};

// Atom pass 1, one workgroup per system.
//   ASE: NaN -> 0, power / |v|^2 / |F|^2, dt / alpha / n_pos update and cell
//        velocity mixing (thread 0), then mixing, v += F dt and |dr|^2.
//   VV:  NaN rows -> 0, half kick with 1/m, drift.
__global__ __launch_bounds__(FIRE_THREADS) void k_fire_atoms1(FireArgs a) {
  __shared__ double sh[4 * FIRE_WAVES];
  __shared__ double coef[3];
  const int s = blockIdx.x;
  if (a.active && a.active[s] == 0) return;
  const int i0 = a.sys_start[s], i1 = a.sys_start[s + 1];
  const tsx_fire_params& p = a.p;
  const bool cellf = p.cell_filter != TSX_CELL_NONE;
  const int ps = a.pos_stride;
  double* w = a.ws + (int64_t)SYSW * s;

  if (p.flavor == TSX_FIRE_VV) {
    // fire.py:193-212
    const double dt = a.dt[s];
    if (threadIdx.x == 0) {
      if (cellf) {
        bool any_nan = false;
        for (int k = 0; k < 9; ++k) any_nan |= isnan(a.cell_vel[9 * s + k]);
        if (any_nan)
          for (int k = 0; k < 9; ++k) a.cell_vel[9 * s + k] = 0.0;
      }
      w[2] = dt;
    }
    for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS) {
      double v[3];
      for (int k = 0; k < 3; ++k) v[k] = a.vel[3 * (int64_t)i + k];
      if (isnan(v[0]) || isnan(v[1]) || isnan(v[2])) v[0] = v[1] = v[2] = 0.0;
      const double hm = 0.5 * dt / a.masses[i];
      for (int k = 0; k < 3; ++k) {
        v[k] += hm * a.forces[3 * (int64_t)i + k];
        a.vel[3 * (int64_t)i + k] = v[k];
        a.pos[(int64_t)ps * i + k] += dt * v[k];
      }
    }
    return;
  }

  // fire.py:306-311: NaN -> 0 element-wise (inf -> +-max)
  double acc[3] = {0.0, 0.0, 0.0};  // power, |v|^2, |F|^2
  for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS) {
    for (int k = 0; k < 3; ++k) {
      const double v = nan_to_num(a.vel[3 * (int64_t)i + k]);
      const double f = a.forces[3 * (int64_t)i + k];
      a.vel[3 * (int64_t)i + k] = v;
      acc[0] += f * v;
      acc[1] += v * v;
      acc[2] += f * f;
    }
  }
  block_sum_n<3>(acc, sh);
  if (threadIdx.x == 0) {
    double dt = a.dt[s];
    double c_keep = 1.0, c_mix = 0.0;
    if (cellf)
      for (int k = 0; k < 9; ++k) a.cell_vel[9 * s + k] = nan_to_num(a.cell_vel[9 * s + k]);
    if (!p.first_step) {
      // fire.py:331-380.  The cell-filter force transform of fire.py:322-329
      // evaluates deform_grad(row_cell, row_cell) = I (CellOptimState has no
      // reference_row_vector_cell), so forces enter untransformed.
      double power = acc[0], vv = acc[1], ff = acc[2];
      if (cellf)
        for (int k = 0; k < 9; ++k) {
          const double cf = a.cell_forces[9 * s + k], cv = a.cell_vel[9 * s + k];
          power += cf * cv;
          vv += cv * cv;
          ff += cf * cf;
        }
      const bool pos = power > 0.0;
      double alpha = a.alpha[s];
      int npos = a.n_pos[s];
      if (pos) {
        if (npos > p.n_min) {
          dt = fmin(dt * p.f_inc, p.dt_max);
          alpha *= p.f_alpha;
        }
        npos += 1;
      } else {
        dt *= p.f_dec;
        alpha = p.alpha_start;
        npos = 0;
      }
      a.alpha[s] = alpha;
      a.n_pos[s] = npos;
      const double vs = sqrt(vv), fsn = sqrt(ff);
      if (cellf)
        for (int k = 0; k < 9; ++k) {
          const double cf = a.cell_forces[9 * s + k], cv = a.cell_vel[9 * s + k];
          a.cell_vel[9 * s + k] = pos ? (1.0 - alpha) * cv + alpha * (cf / (fsn + p.eps) * vs) : 0.0;
        }
      c_keep = pos ? 1.0 - alpha : 0.0;
      c_mix = pos ? alpha * (vs / (fsn + p.eps)) : 0.0;
    }
    a.dt[s] = dt;
    w[2] = dt;
    coef[0] = c_keep;
    coef[1] = c_mix;
    coef[2] = dt;
  }
  __syncthreads();
  // fire.py:400-402: v <- mix(v); v += F dt; dr = v dt
  const double ck = coef[0], cm = coef[1], dt = coef[2];
  double dr2[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS) {
    for (int k = 0; k < 3; ++k) {
      const double f = a.forces[3 * (int64_t)i + k];
      double v = a.vel[3 * (int64_t)i + k];
      v = ck * v + f * cm;
      v += f * dt;
      a.vel[3 * (int64_t)i + k] = v;
      const double d = v * dt;
      dr2[0] += d * d;
    }
  }
  block_sum_n<1>(dr2, sh);
  if (threadIdx.x == 0) w[0] = dr2[0];
}

// Per-system 3x3 work, one thread per system (ASE: clamp, cell velocity and
// cell update with expm / logm, fire.py:404-481; both flavours: list-check
// matrices).
__global__ __launch_bounds__(64) void k_fire_sys(FireArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n_sys) return;
  if (a.active && a.active[s] == 0) return;
  const tsx_fire_params& p = a.p;
  const bool cellf = p.cell_filter != TSX_CELL_NONE;
  double* w = a.ws + (int64_t)SYSW * s;
  M33 cell_new = load_cell(a.cell, s);
  w[31] = 0.0;
  if (p.flavor == TSX_FIRE_ASE) {
    const double dt = w[2];
    double drc[9];
    double drs = w[0];
    if (cellf)
      for (int k = 0; k < 9; ++k) {
        const double cv = a.cell_vel[9 * s + k] + a.cell_forces[9 * s + k] * dt;
        a.cell_vel[9 * s + k] = cv;
        drc[k] = cv * dt;
        drs += drc[k] * drc[k];
      }
    const double nrm = sqrt(drs);
    const bool clamp = nrm > p.max_step;
    w[1] = clamp ? nrm + p.eps : 0.0;
    if (cellf) {
      const M33 ref = load_cell(a.ref_cell, s);
      const M33 f_old = deform_grad_col(ref, cell_new);
      const double cfac = a.cell_factor[s];
      M33 cp;
      for (int k = 0; k < 9; ++k)
        cp.a[k] = a.cell_pos[9 * s + k] + (clamp ? p.max_step * drc[k] / (nrm + p.eps) : drc[k]);
      M33 f_new;
      if (p.cell_filter == TSX_CELL_FRECHET) {
        M33 x = m33_scale(cp, 1.0 / cfac);
        bool exceeds = false;
        for (int k = 0; k < 9; ++k) exceeds |= fabs(x.a[k]) > TSX_MAX_LOG_DEFORM;
        if (exceeds) {  // cell_filters.py:26-60
          for (int k = 0; k < 9; ++k)
            x.a[k] = fmin(fmax(x.a[k], -TSX_MAX_LOG_DEFORM), TSX_MAX_LOG_DEFORM);
          cp = m33_scale(x, cfac);
        }
        f_new = m33_expm(x);
      } else {
        f_new = m33_scale(cp, 1.0 / cfac);
      }
      cell_new = m33_mul(f_new, ref);
      m33_store(cell_new, a.cell + 9 * (int64_t)s);
      const M33 f_adj = deform_grad_col(ref, cell_new);
      const M33 cp_new = p.cell_filter == TSX_CELL_FRECHET ? m33_scale(m33_logm(f_adj), cfac)
                                                            : m33_scale(f_adj, cfac);
      m33_store(cp_new, a.cell_pos + 9 * (int64_t)s);
      m33_store(m33_inv(f_old), w + 3);
      m33_store(f_adj, w + 12);
      w[31] = 1.0;
    }
  }
  if (a.flags) list_check_setup(cell_new, load_cell(a.cell_build, s), a.r_list, a.cutoff, w);
}

// Atom pass 2, one thread per atom: ASE position update (fire.py:404-415,
// 483-492) and, for both flavours, the list check.
__global__ __launch_bounds__(256) void k_fire_atoms2(FireArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_atoms) return;
  const int s = a.atom_sys[i];
  if (a.active && a.active[s] == 0) return;
  const double* w = a.ws + (int64_t)SYSW * s;
  const int ps = a.pos_stride;
  double r[3];
  for (int k = 0; k < 3; ++k) r[k] = a.pos[(int64_t)ps * i + k];
  if (a.p.flavor == TSX_FIRE_ASE) {
    const double dt = w[2], den = w[1];
    double d[3];
    for (int k = 0; k < 3; ++k) {
      const double dk = a.vel[3 * i + k] * dt;
      d[k] = den > 0.0 ? a.p.max_step * dk / den : dk;
    }
    if (w[31] != 0.0) {
      double fr[3];
      m33_apply(m33_load(w + 3), r, fr);
      for (int k = 0; k < 3; ++k) fr[k] += d[k];
      m33_apply(m33_load(w + 12), fr, r);
    } else {
      for (int k = 0; k < 3; ++k) r[k] += d[k];
    }
    for (int k = 0; k < 3; ++k) a.pos[(int64_t)ps * i + k] = r[k];
  }
  if (a.flags) {
    double rt[3];
    m33_apply(m33_load(w + 21), r, rt);
    const double* x0 = a.xref + 4 * i;
    const double dx = rt[0] - x0[0], dy = rt[1] - x0[1], dz = rt[2] - x0[2];
    if (!(dx * dx + dy * dy + dz * dz <= w[30])) a.flags[s] = 1;
  }
}

struct FirePostArgs {
  tsx_fire_params p;
  int n_sys;
  const int32_t* sys_start;
  double* vel;
  const double* forces;
  const double* masses;
  double* dt;
  double* alpha;
  int32_t* n_pos;
  const double* cell;
  const double* ref_cell;
  const double* cell_factor;
  const double* pressure;
  const double* stress;
  double* cell_vel;
  double* cell_forces;
  const double* cell_masses;
  const int32_t* active;
  double* fmax;
  double* cell_build;
  int32_t* flags;
  int init;
};

// Cell forces, one thread per system (cell_filters.py:353-399 and the init
// functions 180-284).
__global__ __launch_bounds__(64) void k_fire_cell_forces(FirePostArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n_sys) return;
  if (a.active && a.active[s] == 0) return;
  const tsx_fire_params& p = a.p;
  const M33 cell = load_cell(a.cell, s);
  const double vol = m33_det(cell);
  M33 vir;
  for (int k = 0; k < 9; ++k) vir.a[k] = -vol * (a.stress[9 * s + k] + a.pressure[9 * s + k]);
  if (p.hydrostatic_strain) {
    const double m = (vir.a[0] + vir.a[4] + vir.a[8]) / 3.0;
    vir = m33_eye(m);
  }
  if (p.constant_volume) {
    const double m = (vir.a[0] + vir.a[4] + vir.a[8]) / 3.0;
    vir.a[0] -= m;
    vir.a[4] -= m;
    vir.a[8] -= m;
  }
  const M33 f = deform_grad_col(load_cell(a.ref_cell, s), cell);
  M33 cf;
  if (p.cell_filter == TSX_CELL_FRECHET) {
    const M33 ucf = m33_mul(vir, m33_inv(m33_t(f)));
    M33 lg = m33_logm(f);
    for (int k = 0; k < 9; ++k)
      lg.a[k] = fmin(fmax(lg.a[k], -TSX_MAX_LOG_DEFORM), TSX_MAX_LOG_DEFORM);
    cf = m33_expm_frechet(m33_t(lg), ucf).y;
  } else {
    cf = m33_mul(vir, m33_t(m33_inv(f)));  // solve(F, virial^T)^T
  }
  const double inv_cf = a.cell_factor[s];
  for (int k = 0; k < 9; ++k) a.cell_forces[9 * s + k] = cf.a[k] / inv_cf;
}

// One workgroup per system: VV second half (fire.py:227-285), fmax, and the
// list bookkeeping of rebuilt systems.
__global__ __launch_bounds__(FIRE_THREADS) void k_fire_post(FirePostArgs a) {
  __shared__ double sh[4 * FIRE_WAVES];
  __shared__ double sc[3];
  const int s = blockIdx.x;
  if (a.active && a.active[s] == 0) return;
  const int i0 = a.sys_start[s], i1 = a.sys_start[s + 1];
  const tsx_fire_params& p = a.p;
  const bool cellf = p.cell_filter != TSX_CELL_NONE;
  if (threadIdx.x == 0 && a.flags && a.flags[s]) {  // the list of s was rebuilt at this cell
    for (int k = 0; k < 9; ++k) a.cell_build[9 * s + k] = a.cell[9 * s + k];
    a.flags[s] = 0;
  }
  double fmax_atom = 0.0;
  if (p.flavor == TSX_FIRE_VV && !a.init) {
    const double dt = a.dt[s];
    double acc[3] = {0.0, 0.0, 0.0};  // power, |v|^2, |F|^2
    for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS) {
      const double hm = 0.5 * dt / a.masses[i];
      double f2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double f = a.forces[3 * (int64_t)i + k];
        const double v = a.vel[3 * (int64_t)i + k] + hm * f;
        a.vel[3 * (int64_t)i + k] = v;
        acc[0] += f * v;
        acc[1] += v * v;
        acc[2] += f * f;
        f2 += f * f;
      }
      fmax_atom = fmax(fmax_atom, sqrt(f2));
    }
    block_sum_n<3>(acc, sh);
    if (threadIdx.x == 0) {
      double power = acc[0], vv = acc[1], ff = acc[2];
      double cv[9], cf[9];
      if (cellf) {
        const double hm = 0.5 * dt / a.cell_masses[s];
        for (int k = 0; k < 9; ++k) {
          cf[k] = a.cell_forces[9 * s + k];
          cv[k] = a.cell_vel[9 * s + k] + hm * cf[k];
          power += cf[k] * cv[k];
        }
      }
      const bool pos = power > 0.0;
      double alpha = a.alpha[s], dtn = dt;
      int npos = a.n_pos[s];
      if (pos) {
        npos += 1;
        if (npos > p.n_min) {
          dtn = fmin(dtn * p.f_inc, p.dt_max);
          alpha *= p.f_alpha;
        }
      } else {
        dtn *= p.f_dec;
        alpha = p.alpha_start;
        npos = 0;
      }
      a.dt[s] = dtn;
      a.alpha[s] = alpha;
      a.n_pos[s] = npos;
      if (cellf)
        for (int k = 0; k < 9; ++k) {
          vv += cv[k] * cv[k];
          ff += cf[k] * cf[k];
        }
      const double vs = sqrt(vv), fsn = sqrt(ff);
      if (cellf)
        for (int k = 0; k < 9; ++k)
          a.cell_vel[9 * s + k] =
              pos ? (1.0 - alpha) * cv[k] + alpha * (cf[k] / (fsn + p.eps) * vs) : 0.0;
      sc[0] = pos ? 1.0 - alpha : 0.0;
      sc[1] = pos ? alpha * (vs / (fsn + p.eps)) : 0.0;
    }
    __syncthreads();
    const double ck = sc[0], cm = sc[1];
    for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS)
      for (int k = 0; k < 3; ++k) {
        const double f = a.forces[3 * (int64_t)i + k];
        a.vel[3 * (int64_t)i + k] = ck * a.vel[3 * (int64_t)i + k] + f * cm;
      }
  } else {
    for (int i = i0 + threadIdx.x; i < i1; i += FIRE_THREADS) {
      double f2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double f = a.forces[3 * (int64_t)i + k];
        f2 += f * f;
      }
      fmax_atom = fmax(fmax_atom, sqrt(f2));
    }
  }
  if (a.fmax) {
    // quantities.py:327-342 (max |F_i|) and the cell-force row norm of
    // runners.py:527-534
    fmax_atom = wave_max(fmax_atom);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = fmax_atom;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0.0;
      for (int w = 0; w < FIRE_WAVES; ++w) m = fmax(m, sh[w]);
      a.fmax[2 * s] = m;
      double mc = 0.0;
      if (cellf)
        for (int r = 0; r < 3; ++r) {
          const double* c = a.cell_forces + 9 * (int64_t)s + 3 * r;
          mc = fmax(mc, sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
        }
      a.fmax[2 * s + 1] = mc;
    }
  }
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int64_t tsx_fire_ws_bytes(int32_t n_systems) {
  return (int64_t)sizeof(double) * SYSW * (n_systems > 0 ? n_systems : 0);
}

int tsx_fire_pre(const tsx_fire_params* prm, int32_t n_systems, const int32_t* sys_start,
                 const int32_t* atom_sys, int64_t n_atoms, double* pos, int32_t pos_stride,
                 double* vel, const double* forces, const double* masses, double* dt,
                 double* alpha, int32_t* n_pos, double* cell_col, const double* ref_cell,
                 const double* cell_factor, double* cell_pos, double* cell_vel,
                 const double* cell_forces, const int32_t* active, const double* cell_build,
                 const double* xref, double r_list, double cutoff, int32_t* flags, void* ws,
                 void* stream) {
  TSX_REQUIRE(prm != nullptr, "params required");
  TSX_REQUIRE(n_systems >= 0 && n_atoms >= 0, "sizes must be >= 0");
  TSX_REQUIRE(pos_stride == 3 || pos_stride == 4, "pos_stride must be 3 or 4");
  TSX_REQUIRE(prm->flavor == TSX_FIRE_ASE || prm->flavor == TSX_FIRE_VV, "unknown FIRE flavour");
  TSX_REQUIRE(prm->cell_filter >= TSX_CELL_NONE && prm->cell_filter <= TSX_CELL_FRECHET,
              "unknown cell filter");
  TSX_REQUIRE(ws != nullptr, "workspace of tsx_fire_ws_bytes(n_systems) required");
  const bool cellf = prm->cell_filter != TSX_CELL_NONE;
  TSX_REQUIRE(!cellf || (ref_cell && cell_factor && cell_pos && cell_vel && cell_forces),
              "cell-filter arrays required");
  TSX_REQUIRE(!flags || (cell_build && xref && r_list > cutoff && cutoff > 0.0),
              "list check needs cell_build, xref and r_list > cutoff > 0");
  if (n_systems == 0) return TSX_OK;
  hipStream_t st = (hipStream_t)stream;
  FireArgs a{*prm,        n_systems,   n_atoms,  sys_start,   atom_sys, pos,
             pos_stride,  vel,         forces,   masses,      dt,       alpha,
             n_pos,       cell_col,    ref_cell, cell_factor, cell_pos, cell_vel,
             cell_forces, active,      cell_build, xref,      r_list,   cutoff,
             flags,       (double*)ws};
  k_fire_atoms1<<<n_systems, FIRE_THREADS, 0, st>>>(a);
  TSX_CHECK_LAUNCH();
  const bool sys_work = prm->flavor == TSX_FIRE_ASE || flags != nullptr;
  if (sys_work) {
    k_fire_sys<<<blocks_for(n_systems, 64), 64, 0, st>>>(a);
    TSX_CHECK_LAUNCH();
  }
  if (n_atoms > 0 && (prm->flavor == TSX_FIRE_ASE || flags != nullptr)) {
    k_fire_atoms2<<<blocks_for(n_atoms, 256), 256, 0, st>>>(a);
    TSX_CHECK_LAUNCH();
  }
  return TSX_OK;
}

int tsx_fire_post(const tsx_fire_params* prm, int32_t n_systems, const int32_t* sys_start,
                  double* vel, const double* forces, const double* masses, double* dt,
                  double* alpha, int32_t* n_pos, const double* cell_col, const double* ref_cell,
                  const double* cell_factor, const double* pressure, const double* stress,
                  double* cell_vel, double* cell_forces, const double* cell_masses,
                  const int32_t* active, double* fmax, double* cell_build, int32_t* flags,
                  int32_t init, void* stream) {
  TSX_REQUIRE(prm != nullptr, "params required");
  TSX_REQUIRE(n_systems >= 0, "n_systems must be >= 0");
  TSX_REQUIRE(prm->flavor == TSX_FIRE_ASE || prm->flavor == TSX_FIRE_VV, "unknown FIRE flavour");
  TSX_REQUIRE(prm->cell_filter >= TSX_CELL_NONE && prm->cell_filter <= TSX_CELL_FRECHET,
              "unknown cell filter");
  const bool cellf = prm->cell_filter != TSX_CELL_NONE;
  TSX_REQUIRE(!cellf || (ref_cell && cell_factor && pressure && stress && cell_forces),
              "cell-filter arrays required");
  TSX_REQUIRE(!(cellf && prm->flavor == TSX_FIRE_VV && !init) || (cell_vel && cell_masses),
              "VV flavour with a cell filter needs cell_vel and cell_masses");
  TSX_REQUIRE(!flags || cell_build, "flags need cell_build");
  if (n_systems == 0) return TSX_OK;
  hipStream_t st = (hipStream_t)stream;
  FirePostArgs a{*prm,        n_systems,   sys_start, vel,        forces,      masses,
                 dt,          alpha,       n_pos,     cell_col,   ref_cell,    cell_factor,
                 pressure,    stress,      cell_vel,  cell_forces, cell_masses, active,
                 fmax,        cell_build,  flags,     init};
  if (cellf) {
    k_fire_cell_forces<<<blocks_for(n_systems, 64), 64, 0, st>>>(a);
    TSX_CHECK_LAUNCH();
  }
  k_fire_post<<<n_systems, FIRE_THREADS, 0, st>>>(a);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
