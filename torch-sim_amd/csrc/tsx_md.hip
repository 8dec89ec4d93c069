// Molecular-dynamics kernels (gfx950): per-step kernels that mirror the
// reference step functions one to one, and the fused graph-capturable step.
//
// Per-step kernels (used by nve_step / nvt_langevin_step):
//   tsx_md_kick        p += F dt                        (md.py:169-187)
//   tsx_md_drift       x += p/m dt                      (md.py:190-208)
//   tsx_md_vv_half     B(dt/2) A(dt)                    (nve.py:121-123)
//   tsx_md_baoa        B(dt/2) A(dt/2) O(dt) A(dt/2)    (nvt.py:229-233, 27-86)
// Fused step (tsx_md_fused_step): one launch evaluates the pair forces on the
// current positions, completes the reference step with B(dt/2), and -- unless
// it is the last step -- runs the first half of the next step (B A O A for
// Langevin, B A(dt) for NVE) into the other position buffer, checking the
// displacement since the last list build against half the Verlet skin.
#include "tsx_pair.h"

namespace tsx {

constexpr int MAX_SPECIES_MD = 8;

__global__ void k_kick(double* __restrict__ p, const double* __restrict__ f, double dt,
                       int64_t n3) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) p[k] = p[k] + f[k] * dt;
}

__global__ void k_drift(double* __restrict__ x, const double* __restrict__ p,
                        const double* __restrict__ m, double dt, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 3 * n) x[k] = x[k] + (p[k] / m[k / 3]) * dt;
}

__global__ void k_vv_half(double* __restrict__ x, double* __restrict__ p,
                          const double* __restrict__ f, const double* __restrict__ m, double dt,
                          int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * n) return;
  const double pk = p[k] + f[k] * (0.5 * dt);
  p[k] = pk;
  x[k] = x[k] + (pk / m[k / 3]) * dt;
}

// noise: either a caller tensor [N,3] or Philox(seed, step, atom).
__device__ __forceinline__ void langevin_noise(const double* __restrict__ noise, int64_t i,
                                               uint64_t seed, uint64_t step, double* z) {
  if (noise) {
    z[0] = noise[3 * i];
    z[1] = noise[3 * i + 1];
    z[2] = noise[3 * i + 2];
  } else {
    normal3(seed, step, (uint32_t)i, z);
  }
}

__global__ void k_baoa(double* __restrict__ x, double* __restrict__ p,
                       const double* __restrict__ f, const double* __restrict__ m,
                       const int32_t* __restrict__ atom_sys, const double* __restrict__ kT,
                       const double* __restrict__ gamma, double dt,
                       const double* __restrict__ noise, uint64_t seed, uint64_t step,
                       int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = atom_sys[i];
  const double hdt = 0.5 * dt;
  const double mi = m[i];
  const double c1 = exp(-gamma[s] * dt);
  const double c2 = sqrt(kT[s] * (1.0 - c1 * c1));
  const double sm = sqrt(mi);
  double z[3];
  langevin_noise(noise, i, seed, step, z);
  for (int c = 0; c < 3; ++c) {
    double pc = p[3 * i + c] + f[3 * i + c] * hdt;
    double xc = x[3 * i + c] + (pc / mi) * hdt;
    pc = c1 * pc + c2 * sm * z[c];
    xc = xc + (pc / mi) * hdt;
    p[3 * i + c] = pc;
    x[3 * i + c] = xc;
  }
}

// ------------------------------------------------------------ fused step
struct FusedArgs {
  const double* params;
  int ns;
  const double* pos_in;  // [N,4] (x, y, z, species bits)
  double* pos_out;       // [N,4]
  const double* xref;    // [N,4] positions at the last list build
  const double* cell_col;
  const int32_t* sys_start;
  const int32_t* atom_sys;
  const uint32_t* nbr;
  const int32_t* count;
  const int32_t* order;
  int64_t n;
  int jbits;
  int64_t ld;
  double rc2;
  double* p;       // [N,3]
  const double* m; // [N]
  double* f;       // [N,3]
  double* e_atom;  // [N]
  double* w_atom;  // [N,6] or null
  double dt;
  const double* kT;     // [S]
  const double* gamma;  // [S]
  const double* noise;  // [N,3] or null
  uint64_t seed;
  const int64_t* step;  // device counter: reference step whose first half runs here
  int end_kick;
  int do_next;
  double half_skin2;
  int32_t* flag_set;  // [S] rebuild request of each system for the next step
};

template <int KIND, int TPA, bool LANGEVIN, bool VIRIAL>
__global__ __launch_bounds__(256) void k_md_fused(FusedArgs A) {
  __shared__ PairP sp[MAX_SPECIES_MD * MAX_SPECIES_MD];
  for (int k = threadIdx.x; k < A.ns * A.ns; k += blockDim.x)
    sp[k] = make_pair_params(KIND, A.params[3 * k], A.params[3 * k + 1], A.params[3 * k + 2]);
  __syncthreads();
  const int64_t gid = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t t = gid / TPA;
  const int lane = (int)(gid % TPA);
  const bool valid = t < A.n;
  const int64_t tt = valid ? t : 0;
  const int i = A.order[tt];
  const int s = A.atom_sys[i];
  const double4 xi = reinterpret_cast<const double4*>(A.pos_in)[i];
  const int ti = A.ns > 1 ? (int)__double_as_longlong(xi.w) : 0;
  const RowAcc a = accumulate_row<KIND, TPA, VIRIAL, 4>(
      sp, A.ns, nullptr, A.pos_in, A.cell_col + 9 * (int64_t)s, A.sys_start[s], xi.x, xi.y, xi.z,
      ti, A.nbr, tt, valid ? A.count[tt] : 0, lane, A.jbits, A.ld, A.rc2);
  if (!valid || lane != 0) return;
  const double hdt = 0.5 * A.dt;
  const double mi = A.m[i];
  double p0 = A.p[3 * (int64_t)i], p1 = A.p[3 * (int64_t)i + 1], p2 = A.p[3 * (int64_t)i + 2];
  if (A.end_kick) {  // B(dt/2) closing the current reference step
    p0 = p0 + a.fx * hdt;
    p1 = p1 + a.fy * hdt;
    p2 = p2 + a.fz * hdt;
  }
  double4 xo = xi;
  if (A.do_next) {
    p0 = p0 + a.fx * hdt;  // B(dt/2) opening the next step
    p1 = p1 + a.fy * hdt;
    p2 = p2 + a.fz * hdt;
    if (LANGEVIN) {
      xo.x = xo.x + (p0 / mi) * hdt;
      xo.y = xo.y + (p1 / mi) * hdt;
      xo.z = xo.z + (p2 / mi) * hdt;
      const double c1 = exp(-A.gamma[s] * A.dt);
      const double c2 = sqrt(A.kT[s] * (1.0 - c1 * c1));
      const double sm = sqrt(mi);
      double z[3];
      langevin_noise(A.noise, i, A.seed, (uint64_t)A.step[0], z);
      p0 = c1 * p0 + c2 * sm * z[0];
      p1 = c1 * p1 + c2 * sm * z[1];
      p2 = c1 * p2 + c2 * sm * z[2];
      xo.x = xo.x + (p0 / mi) * hdt;
      xo.y = xo.y + (p1 / mi) * hdt;
      xo.z = xo.z + (p2 / mi) * hdt;
    } else {
      xo.x = xo.x + (p0 / mi) * A.dt;
      xo.y = xo.y + (p1 / mi) * A.dt;
      xo.z = xo.z + (p2 / mi) * A.dt;
    }
    if (A.xref && A.flag_set) {
      const double4 r = reinterpret_cast<const double4*>(A.xref)[i];
      const double ex = xo.x - r.x, ey = xo.y - r.y, ez = xo.z - r.z;
      if (ex * ex + ey * ey + ez * ez > A.half_skin2) A.flag_set[s] = 1;
    }
  }
  reinterpret_cast<double4*>(A.pos_out)[i] = xo;
  A.p[3 * (int64_t)i] = p0;
  A.p[3 * (int64_t)i + 1] = p1;
  A.p[3 * (int64_t)i + 2] = p2;
  A.f[3 * (int64_t)i] = a.fx;
  A.f[3 * (int64_t)i + 1] = a.fy;
  A.f[3 * (int64_t)i + 2] = a.fz;
  A.e_atom[i] = a.e;
  if (VIRIAL)
    for (int c = 0; c < 6; ++c) A.w_atom[6 * (int64_t)i + c] = a.w[c];
}

// per-system energy sums (fixed order), clear this step's rebuild gates and
// advance the device step counter
// gated copy of the positions into the displacement reference
// (counts the rebuilds that actually ran in ``counter`` when given)
__global__ void k_copy_ref(const double* __restrict__ pos, double* __restrict__ xref, int64_t n4,
                           const int32_t* __restrict__ gate, int32_t* __restrict__ counter) {
  if (gate && *gate == 0) return;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n4) xref[k] = pos[k];
  if (k == 0 && counter) counter[0] += 1;
}

template <int KIND, int TPA, bool LANGEVIN>
static void launch_fused_v(bool vir, hipStream_t st, int nblk, const FusedArgs& a) {
  if (vir)
    k_md_fused<KIND, TPA, LANGEVIN, true><<<nblk, 256, 0, st>>>(a);
  else
    k_md_fused<KIND, TPA, LANGEVIN, false><<<nblk, 256, 0, st>>>(a);
}

template <int KIND, int TPA>
static void launch_fused_l(bool lang, bool vir, hipStream_t st, int nblk, const FusedArgs& a) {
  if (lang)
    launch_fused_v<KIND, TPA, true>(vir, st, nblk, a);
  else
    launch_fused_v<KIND, TPA, false>(vir, st, nblk, a);
}

template <int KIND>
static int launch_fused_t(int tpa, bool lang, bool vir, hipStream_t st, int nblk,
                          const FusedArgs& a) {
  switch (tpa) {
    case 1: launch_fused_l<KIND, 1>(lang, vir, st, nblk, a); return 0;
    case 2: launch_fused_l<KIND, 2>(lang, vir, st, nblk, a); return 0;
    case 4: launch_fused_l<KIND, 4>(lang, vir, st, nblk, a); return 0;
    case 8: launch_fused_l<KIND, 8>(lang, vir, st, nblk, a); return 0;
    default: return -1;
  }
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int tsx_md_kick(double* p, const double* f, double dt, int64_t n_atoms, void* stream) {
  if (n_atoms == 0) return TSX_OK;
  k_kick<<<blocks_for(3 * n_atoms, 256), 256, 0, (hipStream_t)stream>>>(p, f, dt, 3 * n_atoms);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_md_drift(double* x, const double* p, const double* m, double dt, int64_t n_atoms,
                 void* stream) {
  if (n_atoms == 0) return TSX_OK;
  k_drift<<<blocks_for(3 * n_atoms, 256), 256, 0, (hipStream_t)stream>>>(x, p, m, dt, n_atoms);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_md_vv_half(double* x, double* p, const double* f, const double* m, double dt,
                   int64_t n_atoms, void* stream) {
  if (n_atoms == 0) return TSX_OK;
  k_vv_half<<<blocks_for(3 * n_atoms, 256), 256, 0, (hipStream_t)stream>>>(x, p, f, m, dt,
                                                                            n_atoms);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_md_baoa(double* x, double* p, const double* f, const double* m, const int32_t* atom_sys,
                const double* kT, const double* gamma, double dt, const double* noise,
                uint64_t seed, uint64_t step, int64_t n_atoms, void* stream) {
  TSX_REQUIRE(kT && gamma && atom_sys, "kT, gamma and atom_sys are required");
  if (n_atoms == 0) return TSX_OK;
  k_baoa<<<blocks_for(n_atoms, 256), 256, 0, (hipStream_t)stream>>>(
      x, p, f, m, atom_sys, kT, gamma, dt, noise, seed, step, n_atoms);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_md_fused_step(int32_t kind, const double* params, int32_t n_species,
                      const double* pos_in, double* pos_out, const double* xref,
                      const double* cell_col, const int32_t* sys_start,
                      const int32_t* atom_sys, const uint32_t* nbr, const int32_t* count,
                      const int32_t* order, int64_t n_atoms, int32_t jbits, int32_t tpa,
                      int64_t ld, double cutoff, double* p, const double* m, double* f,
                      double* e_atom, double* w_atom, double dt, int32_t langevin,
                      const double* kT, const double* gamma, const double* noise,
                      uint64_t seed, const int64_t* step, int32_t end_kick, int32_t do_next,
                      double half_skin, int32_t* flag_set, void* stream) {
  TSX_REQUIRE(kind >= TSX_LJ && kind <= TSX_SOFT_SPHERE, "unknown pair kind %d", kind);
  TSX_REQUIRE(n_species >= 1 && n_species <= MAX_SPECIES_MD, "n_species out of range");
  TSX_REQUIRE(pos_in != pos_out, "pos_in and pos_out must be different buffers");
  TSX_REQUIRE(!langevin || (kT && gamma && step), "Langevin needs kT, gamma and step");
  TSX_REQUIRE(jbits >= 1 && jbits <= 26, "jbits must be in [1, 26]");
  if (n_atoms == 0) return TSX_OK;
  FusedArgs a;
  a.params = params;
  a.ns = n_species;
  a.pos_in = pos_in;
  a.pos_out = pos_out;
  a.xref = xref;
  a.cell_col = cell_col;
  a.sys_start = sys_start;
  a.atom_sys = atom_sys;
  a.nbr = nbr;
  a.count = count;
  a.order = order;
  a.n = n_atoms;
  a.jbits = jbits;
  a.ld = ld;
  a.rc2 = cutoff * cutoff;
  a.p = p;
  a.m = m;
  a.f = f;
  a.e_atom = e_atom;
  a.w_atom = w_atom;
  a.dt = dt;
  a.kT = kT;
  a.gamma = gamma;
  a.noise = noise;
  a.seed = seed;
  a.step = step;
  a.end_kick = end_kick;
  a.do_next = do_next;
  a.half_skin2 = half_skin * half_skin;
  a.flag_set = flag_set;
  hipStream_t st = (hipStream_t)stream;
  const int nblk = blocks_for(n_atoms * tpa, 256);
  const bool vir = w_atom != nullptr;
  int rc;
  if (kind == TSX_LJ)
    rc = launch_fused_t<TSX_LJ>(tpa, langevin != 0, vir, st, nblk, a);
  else if (kind == TSX_MORSE)
    rc = launch_fused_t<TSX_MORSE>(tpa, langevin != 0, vir, st, nblk, a);
  else
    rc = launch_fused_t<TSX_SOFT_SPHERE>(tpa, langevin != 0, vir, st, nblk, a);
  TSX_REQUIRE(rc == 0, "tpa must be 1, 2, 4 or 8");
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_copy_ref(const double* pos, double* xref, int64_t n_doubles, const int32_t* gate,
                 int32_t* counter, void* stream) {
  if (n_doubles == 0) return TSX_OK;
  k_copy_ref<<<blocks_for(n_doubles, 256), 256, 0, (hipStream_t)stream>>>(pos, xref, n_doubles,
                                                                          gate, counter);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
