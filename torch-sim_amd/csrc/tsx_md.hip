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
#include <type_traits>

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
  const double inv_m = 1.0 / mi;  // the fused step's drift arithmetic
  double z[3];
  langevin_noise(noise, i, seed, step, z);
  for (int c = 0; c < 3; ++c) {
    double pc = p[3 * i + c] + f[3 * i + c] * hdt;
    double xc = x[3 * i + c] + (pc * inv_m) * hdt;
    pc = c1 * pc + c2 * sm * z[c];
    xc = xc + (pc * inv_m) * hdt;
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
  const int32_t* sys_start;
  const int32_t* atom_sys;
  const uint32_t* nbr;   // stencil-word list
  const int32_t* count;
  const int32_t* order;
  int64_t n;
  int jbits;
  int64_t ld;
  double rc;
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
  double half_skin;
  int32_t* flag_set;    // [S] rebuild request of each system for the next step
  const float* rel_in;  // [N,4] r - A of pos_in
  float* rel_out;       // [N,4] r - A of pos_out
  const double* anchor; // [N,4] bin anchors A
  const double* ustep;  // [S,12] a_k / nb_k (rows), sum_k |a_k / nb_k|
  int ortho;            // every cell orthorhombic (ustep off-diagonals zero)
  int stream_words;     // load the list words non-temporally
};

// The FP64 side of the cutoff decision for entries within the rounding
// margin: (r_j - A_j) - (r_i - A_i) + d . ustep.
struct ExactCtx {
  const double4* pos_s;  // this system's packed positions
  const double4* anc_s;  // this system's anchors
  int il;                // this row's atom within the system (r_i - A_i when needed)
  const double* u;       // ustep row of the system
  double rc2;
};

// Walk this lane's share of row t with stencil words: one 16-byte gather of
// the neighbour's bin-relative position per entry and the FP32 displacement
// d = (rel_j - rel_i) + d_x u_0 + d_y u_1 + d_z u_2.  Entries with
// d^2 >= hi2 are out, d^2 < lo2 in; between the two (the proven FP32
// rounding margin around rc) the FP64 displacement decides, so the pair set
// is the FP64 one (pair_potential.py:181-194).  The pair function is FP32 on
// d, accumulation FP64 (pair_potential.py:419-430).  UNR entries at a time:
// all words, then all gathers in flight.
// entries per lane with their gathers in flight together
constexpr int UNR_REL = 8;

template <bool VIRIAL>
__device__ __forceinline__ void add_pair(RowAcc& a, float e, float fs, float dx, float dy,
                                         float dz) {
  const double f = (double)fs;
  const double ddx = (double)dx, ddy = (double)dy, ddz = (double)dz;
  a.fx = fma(-f, ddx, a.fx);
  a.fy = fma(-f, ddy, a.fy);
  a.fz = fma(-f, ddz, a.fz);
  a.e = fma(0.5, (double)e, a.e);
  if (VIRIAL) {
    const double hf = 0.5 * f;
    a.w[0] += hf * ddx * ddx;
    a.w[1] += hf * ddy * ddy;
    a.w[2] += hf * ddz * ddz;
    a.w[3] += hf * ddy * ddz;
    a.w[4] += hf * ddx * ddz;
    a.w[5] += hf * ddx * ddy;
  }
}

// d = b + s_x u_0 + s_y u_1 + s_z u_2 for a stencil offset s and b = rel_j -
// rel_i, per component, as an FMA chain onto b.  With ORTHO (every cell
// orthorhombic: ustep rows have zero off-diagonal entries) one FMA per
// component: an FMA with a zero factor returns its addend, so the general
// chain gives the same bits on such cells.
template <bool ORTHO>
__device__ __forceinline__ void stencil_fold(float fx, float fy, float fz, const float* uf,
                                             float bx, float by, float bz, float& dx, float& dy,
                                             float& dz) {
  if (ORTHO) {
    dx = fmaf(fx, uf[0], bx);
    dy = fmaf(fy, uf[4], by);
    dz = fmaf(fz, uf[8], bz);
  } else {
    dx = fmaf(fz, uf[6], fmaf(fy, uf[3], fmaf(fx, uf[0], bx)));
    dy = fmaf(fz, uf[7], fmaf(fy, uf[4], fmaf(fx, uf[1], by)));
    dz = fmaf(fz, uf[8], fmaf(fy, uf[5], fmaf(fx, uf[2], bz)));
  }
}

// FP32 displacement and squared distance of one list entry from its word and
// the neighbour's relative position
template <bool ORTHO>
__device__ __forceinline__ void rel_disp(uint32_t w, const float4& q, int jbits, int sbits,
                                         const float* uf, float xri, float yri, float zri,
                                         float& dx, float& dy, float& dz, float& d2) {
  const float fx = (float)nl_shift(w, jbits, sbits, 0);
  const float fy = (float)nl_shift(w, jbits, sbits, 1);
  const float fz = (float)nl_shift(w, jbits, sbits, 2);
  stencil_fold<ORTHO>(fx, fy, fz, uf, q.x - xri, q.y - yri, q.z - zri, dx, dy, dz);
  d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Raw buffer loads: a uniform 128-bit resource, a per-lane 32-bit byte
// offset and a scalar byte offset, so a batch's word addresses cost no VALU
// work (the eight entries of a lane differ by scalar multiples of the row
// stride).  No range clamp (num_records at its maximum): every offset below
// addresses an element of the array.  aux 2 sets the non-temporal bit.
constexpr int BUF_FLAGS = 0x00020000;  // gfx9 dword 3: 32-bit data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xffffffff,
                                           BUF_FLAGS);
}

template <bool NT>
__device__ __forceinline__ uint32_t buf_word(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                             uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, NT ? 2 : 0);
}

// the neighbour's relative position: x, y, z (and the species bits with MULTI)
template <bool MULTI>
__device__ __forceinline__ float4 buf_rel(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  if (MULTI) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
  } else {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)voff, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       0.f);
  }
}

// NTW: load the list words non-temporally (a compile-time choice: as a
// runtime branch the compiler merges the two loads and drops the hint).
// Byte offsets are 32-bit: the caller guarantees 4 * (rows * TPA) and
// 4 * 7 * ld * TPA below 2^32 and 16 * N below 2^32.
template <int KIND, int TPA, bool VIRIAL, bool MULTI, bool ORTHO, bool NTW>
__device__ __forceinline__ RowAcc accumulate_rel(const PairP* sp, const PairP& p0, int ns, int ti,
                                                 const float* __restrict__ rel_all,
                                                 uint32_t rel_base, float xri, float yri,
                                                 float zri, const float* uf,
                                                 const uint32_t* __restrict__ nbr, int64_t t,
                                                 int cnt, int lane, int jbits, int64_t ld,
                                                 float lo2, float hi2, const ExactCtx& ex) {
  RowAcc a;
  a.fx = a.fy = a.fz = a.e = 0.0;
  for (int c = 0; c < 6; ++c) a.w[c] = 0.0;
  const int sbits = (32 - jbits) / 3;
  const uint32_t jmask = (1u << jbits) - 1u;
  const int64_t stride = ld * TPA;
  const int iters = cnt > lane ? (cnt - lane + TPA - 1) / TPA : 0;
  const __amdgpu_buffer_rsrc_t rrel = buf_rsrc(rel_all);
  const uint32_t woff = (uint32_t)(t * TPA + lane) * 4u;  // entry 0 of this lane's share
  const uint32_t sstep = (uint32_t)stride * 4u;           // bytes between a lane's entries
  // Branch-free batches: across a wave nearly every entry of a batch is in
  // range on some lane, so the pair arithmetic runs for all of them anyway;
  // per-entry branches would only add exec-mask traffic.  Only a lane's last
  // batch can run past its row's end: there the lane re-reads its last entry
  // (the load needs no guard) and the entries past the end add exact zeros,
  // so the sums keep the bits of an entry-by-entry walk.
  auto batch = [&](int k0, auto tail_tag) {
    constexpr bool TAIL = decltype(tail_tag)::value;
    uint32_t w[UNR_REL];
    if (!TAIL) {
      // the resource's base steps by whole batches (scalar), lanes keep woff
      const __amdgpu_buffer_rsrc_t rw = buf_rsrc(nbr + (int64_t)k0 * stride);
#pragma unroll
      for (int u = 0; u < UNR_REL; ++u) w[u] = buf_word<NTW>(rw, woff, (uint32_t)u * sstep);
    } else {
      const __amdgpu_buffer_rsrc_t rw = buf_rsrc(nbr + (int64_t)k0 * stride);
#pragma unroll
      for (int u = 0; u < UNR_REL; ++u) {
        const int k = k0 + u < iters ? u : iters - 1 - k0;  // past the end: the last entry
        w[u] = buf_word<NTW>(rw, woff + (uint32_t)k * sstep, 0u);
      }
    }
    float4 q[UNR_REL];
#pragma unroll
    for (int u = 0; u < UNR_REL; ++u) q[u] = buf_rel<MULTI>(rrel, ((w[u] & jmask) + rel_base) << 4);
    uint32_t amb = 0;  // entries within the rounding margin of rc
#pragma unroll
    for (int u = 0; u < UNR_REL; ++u) {
      const bool valid = !TAIL || k0 + u < iters;
      float dx, dy, dz, d2;
      rel_disp<ORTHO>(w[u], q[u], jbits, sbits, uf, xri, yri, zri, dx, dy, dz, d2);
      const bool in = valid && d2 < lo2;
      amb |= (valid && !(d2 < lo2) && d2 < hi2) ? (1u << u) : 0u;
      float e, fs;
      // a discarded entry may have d2 = 0 (inf / NaN here): the select drops it
      pair_eval_pos<KIND>(d2, MULTI ? sp[ti * ns + __float_as_int(q[u].w)] : p0, e, fs);
      add_pair<VIRIAL>(a, in ? e : 0.f, in ? fs : 0.f, dx, dy, dz);
    }
    while (amb) {  // rare: the FP64 displacement decides
      const double4 xi4 = ex.pos_s[ex.il];
      const double4 ai4 = ex.anc_s[ex.il];
      const double xi = xi4.x - ai4.x, yi = xi4.y - ai4.y, zi = xi4.z - ai4.z;
      const int u = __builtin_ctz(amb);
      amb &= amb - 1u;
      uint32_t wu = w[0];
#pragma unroll
      for (int v = 1; v < UNR_REL; ++v) wu = u == v ? w[v] : wu;
      const uint32_t jl = wu & jmask;
      const double ix = (double)nl_shift(wu, jbits, sbits, 0);
      const double iy = (double)nl_shift(wu, jbits, sbits, 1);
      const double iz = (double)nl_shift(wu, jbits, sbits, 2);
      const double4 pj = ex.pos_s[jl];
      const double4 aj = ex.anc_s[jl];
      const double gx = ((pj.x - aj.x) - xi) + fma(iz, ex.u[6], fma(iy, ex.u[3], ix * ex.u[0]));
      const double gy = ((pj.y - aj.y) - yi) + fma(iz, ex.u[7], fma(iy, ex.u[4], ix * ex.u[1]));
      const double gz = ((pj.z - aj.z) - zi) + fma(iz, ex.u[8], fma(iy, ex.u[5], ix * ex.u[2]));
      const double d2e = gx * gx + gy * gy + gz * gz;
      if (!(d2e < ex.rc2)) continue;
      float e, fs;
      pair_eval<KIND>((float)d2e,
                      MULTI ? sp[ti * ns + (int)__double_as_longlong(pj.w)] : p0, e, fs);
      add_pair<VIRIAL>(a, e, fs, (float)gx, (float)gy, (float)gz);
    }
  };
  const int full = iters - iters % UNR_REL;  // entries in whole batches
  int k0 = 0;
  for (; k0 < full; k0 += UNR_REL) batch(k0, std::false_type{});
  if (k0 < iters) batch(k0, std::true_type{});
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

// LJ and Morse without the virial: four waves per SIMD (107-111 VGPRs with
// one species pair, whose constants stay in registers).  The gather-bound
// walk gains from the fourth wave's loads in flight; a fifth (96 VGPRs, a
// few spills outside the walk) loses more to the scheduler's tighter
// budget than it gains.  The virial and soft-sphere variants keep the
// default budget: capped, they would spill.
template <int KIND, int TPA, bool LANGEVIN, bool VIRIAL, bool MULTI, bool ORTHO, bool NTW = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(VIRIAL || KIND == TSX_SOFT_SPHERE ? 1 : 4)))
void k_md_fused(FusedArgs A) {
  // Prologue chain before the first neighbour gather: row -> (atom, system)
  // -> (system start, frame, own relative position).  Rows of a system cover
  // its own index range [sys_start[s], sys_start[s + 1]), so atom_sys[t] is
  // the row's system without waiting for order[t]; rel_in[i] holds
  // (float)(pos_in[i] - anchor[i]) exactly (written by the build or the
  // previous step from the same FP64 difference), so the FP64 records are
  // read only by the tail and the rare FP64 cutoff path.
  const int64_t gid = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t t = gid / TPA;
  const int lane = (int)(gid % TPA);
  const bool valid = t < A.n;
  const int64_t tt = valid ? t : 0;
  const int i = A.order[tt];
  const int s = A.atom_sys[tt];
  const int base = A.sys_start[s];
  const float4 ri = reinterpret_cast<const float4*>(A.rel_in)[i];
  const int ti = MULTI ? __float_as_int(ri.w) : 0;
  // one species pair: its constants in registers; several: an LDS table
  __shared__ PairP sp[MAX_SPECIES_MD * MAX_SPECIES_MD];
  PairP pp{};
  if (MULTI) {
    for (int k = threadIdx.x; k < A.ns * A.ns; k += blockDim.x)
      sp[k] = make_pair_params(KIND, A.params[3 * k], A.params[3 * k + 1], A.params[3 * k + 2]);
    __syncthreads();
  } else {
    pp = make_pair_params(KIND, A.params[0], A.params[1], A.params[2]);
  }
  ExactCtx ex;
  ex.il = i - base;
  // the system's bin steps and the FP32 rounding margin around rc: every
  // term of the FP32 displacement is bounded by M = rc + 2 skin + 3 sum|u_k|
  // and the displacement carries fewer than 32 roundings of 2^-24 M
  const double* us = A.ustep + 12 * (int64_t)s;
  float uf[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) uf[k] = (float)us[k];
  const double big = A.rc + 4.0 * A.half_skin + 3.0 * us[9];
  const double tol = 64.0 * 5.9604644775390625e-8 * big;  // 64 * 2^-24 * M
  const float lo2 = (float)((A.rc - tol) * (A.rc - tol) * (1.0 - 1.0 / 1048576.0));
  const float hi2 = (float)((A.rc + tol) * (A.rc + tol) * (1.0 + 1.0 / 1048576.0));
  ex.pos_s = reinterpret_cast<const double4*>(A.pos_in) + base;
  ex.anc_s = reinterpret_cast<const double4*>(A.anchor) + base;
  ex.u = us;
  ex.rc2 = A.rc * A.rc;
  const int cnt = valid ? A.count[tt] : 0;
  const RowAcc a = accumulate_rel<KIND, TPA, VIRIAL, MULTI, ORTHO, NTW>(
      sp, pp, A.ns, ti, A.rel_in, (uint32_t)base, ri.x, ri.y, ri.z, uf, A.nbr, tt, cnt, lane,
      A.jbits, A.ld, lo2, hi2, ex);
  // Philox normals of (seed, step, atom) for the O step: with two or more
  // lanes per atom, lanes 0 and 1 each draw one of normal3's two blocks (the
  // same operations on each block, so the same values), in one instruction
  // stream instead of two on lane 0
  double zn[3] = {0.0, 0.0, 0.0};
  if (LANGEVIN && TPA >= 2 && A.do_next && !A.noise) {
    const uint64_t step = (uint64_t)A.step[0];
    const uint32_t k0 = (uint32_t)A.seed, k1 = (uint32_t)(A.seed >> 32);
    const u4 r = philox4x32_10(
        u4{(uint32_t)i, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)(lane & 1)}, k0, k1);
    const double amp = sqrt(-2.0 * log(u01_53(r.x, r.y)));
    double sn, cs;
    sincospi(2.0 * u01_53(r.z, r.w), &sn, &cs);
    zn[0] = amp * cs;
    zn[1] = amp * sn;
    zn[2] = __shfl_xor(zn[0], 1, 64);  // block 1's first normal, from lane 1
  }
  if (!valid || lane != 0) return;
  // own position and anchor again (cached): not kept live across the row walk
  const double4 xi = reinterpret_cast<const double4*>(A.pos_in)[i];
  const double4 ai = reinterpret_cast<const double4*>(A.anchor)[i];
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
    const double inv_m = 1.0 / mi;  // one division for the drifts
    if (LANGEVIN) {
      xo.x = xo.x + (p0 * inv_m) * hdt;
      xo.y = xo.y + (p1 * inv_m) * hdt;
      xo.z = xo.z + (p2 * inv_m) * hdt;
      const double c1 = exp(-A.gamma[s] * A.dt);
      const double c2 = sqrt(A.kT[s] * (1.0 - c1 * c1));
      const double sm = sqrt(mi);
      double z[3] = {zn[0], zn[1], zn[2]};
      if (A.noise || TPA == 1) langevin_noise(A.noise, i, A.seed, (uint64_t)A.step[0], z);
      p0 = c1 * p0 + c2 * sm * z[0];
      p1 = c1 * p1 + c2 * sm * z[1];
      p2 = c1 * p2 + c2 * sm * z[2];
      xo.x = xo.x + (p0 * inv_m) * hdt;
      xo.y = xo.y + (p1 * inv_m) * hdt;
      xo.z = xo.z + (p2 * inv_m) * hdt;
    } else {
      xo.x = xo.x + (p0 * inv_m) * A.dt;
      xo.y = xo.y + (p1 * inv_m) * A.dt;
      xo.z = xo.z + (p2 * inv_m) * A.dt;
    }
    if (A.xref && A.flag_set) {
      const double4 r = reinterpret_cast<const double4*>(A.xref)[i];
      const double gx = xo.x - r.x, gy = xo.y - r.y, gz = xo.z - r.z;
      if (gx * gx + gy * gy + gz * gz > A.half_skin * A.half_skin) A.flag_set[s] = 1;
    }
  }
  reinterpret_cast<double4*>(A.pos_out)[i] = xo;
  float4 ro;
  ro.x = (float)(xo.x - ai.x);
  ro.y = (float)(xo.y - ai.y);
  ro.z = (float)(xo.z - ai.z);
  ro.w = __int_as_float((int)__double_as_longlong(xo.w));
  reinterpret_cast<float4*>(A.rel_out)[i] = ro;
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

// the orthorhombic variants only without the virial (the engine's kernels)
template <int KIND, int TPA, bool LANGEVIN>
static void launch_fused_v(bool vir, hipStream_t st, int nblk, const FusedArgs& a) {
  const bool multi = a.ns > 1;
  if (vir && multi)
    k_md_fused<KIND, TPA, LANGEVIN, true, true, false><<<nblk, 256, 0, st>>>(a);
  else if (vir)
    k_md_fused<KIND, TPA, LANGEVIN, true, false, false><<<nblk, 256, 0, st>>>(a);
  else if (multi && a.ortho)
    k_md_fused<KIND, TPA, LANGEVIN, false, true, true><<<nblk, 256, 0, st>>>(a);
  else if (multi)
    k_md_fused<KIND, TPA, LANGEVIN, false, true, false><<<nblk, 256, 0, st>>>(a);
  else if (a.ortho && a.stream_words)  // one large box: words streamed past L2
    k_md_fused<KIND, TPA, LANGEVIN, false, false, true, true><<<nblk, 256, 0, st>>>(a);
  else if (a.stream_words)
    k_md_fused<KIND, TPA, LANGEVIN, false, false, false, true><<<nblk, 256, 0, st>>>(a);
  else if (a.ortho)
    k_md_fused<KIND, TPA, LANGEVIN, false, false, true><<<nblk, 256, 0, st>>>(a);
  else
    k_md_fused<KIND, TPA, LANGEVIN, false, false, false><<<nblk, 256, 0, st>>>(a);
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
                      double half_skin, int32_t* flag_set, const float* rel_in, float* rel_out,
                      const double* anchor, const double* ustep, int32_t orthorhombic,
                      int32_t stream_words,
                      void* stream) {
  (void)cell_col;  // the frame (anchor, ustep) carries the geometry
  TSX_REQUIRE(kind >= TSX_LJ && kind <= TSX_SOFT_SPHERE, "unknown pair kind %d", kind);
  TSX_REQUIRE(n_species >= 1 && n_species <= MAX_SPECIES_MD, "n_species out of range");
  TSX_REQUIRE(pos_in != pos_out, "pos_in and pos_out must be different buffers");
  TSX_REQUIRE(!langevin || (kT && gamma && step), "Langevin needs kT, gamma and step");
  TSX_REQUIRE(jbits >= 1 && jbits <= 22, "jbits must be in [1, 22]");
  TSX_REQUIRE(rel_in && rel_out && rel_in != rel_out && anchor && ustep,
              "the fused step needs the list frame: rel_in, rel_out, anchor, ustep");
  TSX_REQUIRE(cutoff > 0.0 && half_skin >= 0.0, "need cutoff > 0 and half_skin >= 0");
  TSX_REQUIRE(ld >= n_atoms, "ld must be at least n_atoms");
  // 32-bit byte offsets of the buffer loads: a batch spans 8 row strides of
  // words, relative positions 16 bytes per atom
  TSX_REQUIRE(ld * (int64_t)tpa * 4 * 8 < ((int64_t)1 << 32) && n_atoms * 16 < ((int64_t)1 << 32),
              "list too large for one fused launch (ld * tpa = %lld): split the batch",
              (long long)(ld * tpa));
  if (n_atoms == 0) return TSX_OK;
  FusedArgs a;
  a.params = params;
  a.ns = n_species;
  a.pos_in = pos_in;
  a.pos_out = pos_out;
  a.xref = xref;
  a.sys_start = sys_start;
  a.atom_sys = atom_sys;
  a.nbr = nbr;
  a.count = count;
  a.order = order;
  a.n = n_atoms;
  a.jbits = jbits;
  a.ld = ld;
  a.rc = cutoff;
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
  a.half_skin = half_skin;
  a.flag_set = flag_set;
  a.rel_in = rel_in;
  a.rel_out = rel_out;
  a.anchor = anchor;
  a.ustep = ustep;
  a.ortho = orthorhombic != 0;
  a.stream_words = stream_words != 0;
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
