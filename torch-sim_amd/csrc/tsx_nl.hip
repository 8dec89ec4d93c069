// Cell-list neighbor list over concatenated periodic systems (gfx950).
//
// Replaces the reference's padded, host-synchronising linked-cell search
// (torch_sim/transforms.py:1009-1373, torch_sim/neighbors/torch_nl.py:156-215)
// with per-system fractional grids.  Every stage after ``setup`` can be gated
// per system, so a Verlet-skin engine rebuilds only the systems whose atoms
// moved more than half the skin:
//   setup     : one workgroup computes each system's inverse cell, plane
//               spacings, grid size nb_k, stencil half-width m_k and bin
//               offsets; a grid kernel maps every bin to its system.
//   bin       : one thread per atom, fractional coordinate -> bin, atomic count.
//   scan      : one workgroup per system scans its bin counts.
//   scatter   : atoms -> bin order (one-bin systems keep atom order).
//   sort+wrap : one wave per bin ranks its atoms by index in LDS
//               (deterministic order) and writes the wrapped FP64 coordinates
//               (transforms.py:160-210) in bin order.
//   fill      : one workgroup per bin stages the stencil's candidate atoms in
//               LDS; each thread owns one atom of the bin and keeps the images
//               with FP64 |wr_i - wr_j - S.L|^2 < r_list^2 (torch_nl.py:87),
//               unwrapped-frame shift (torch_nl.py:214), ELL write with
//               overflow detection.
#include "tsx_common.h"
#include "tsx_mat33.h"

#include <cstdarg>
#include <mutex>

namespace tsx {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Per-system geometry record (doubles):
//   [0..8]  row cell L (row k = lattice vector k)
//   [9..17] Linv with f = r . Linv  (f_k = sum_m r_m Linv[m][k])
//   [18..20] plane spacings w_k,  [21] |det|,  [22] active (1.0/0.0)
constexpr int GEOM = 24;
// Per-system grid record (int32): nb[3], m[3], bin_off, flags (bit k = pbc_k)
constexpr int GRID = 8;
constexpr int META = 16;  // [0] total bins

struct NLWs {
  double* geom;
  int32_t* grid;
  int32_t* meta;
  int32_t* bin_count;  // bin_cap + 1
  int32_t* bin_start;  // bin_cap + 1
  int32_t* atom_bin;   // N
  int32_t* slot;       // N
  double* wr;          // N x 4 (wrapped xyz, packed wrap shift)
  int32_t* scan_tmp;   // scan block sums
  int32_t* unsorted;   // N: atoms scattered to their bins, before the in-bin sort
  int32_t* act;        // [S + 1] exclusive prefix of the bins of gated-on systems
  int2* witem;         // [bin_cap] (system, bin) of each active-bin work item
  uint32_t* cw;        // N, bin order: partial list word (local index, negated wrap)
  int32_t* wmax;       // S: largest wrap-shift magnitude of the system's atoms
};

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t ws_layout(int64_t n, int32_t s, int32_t bin_cap, NLWs* w, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  const int64_t nb = (int64_t)bin_cap + 1;
  void* p_geom = take(sizeof(double) * GEOM * (size_t)(s > 0 ? s : 1));
  void* p_grid = take(sizeof(int32_t) * GRID * (size_t)(s > 0 ? s : 1));
  void* p_meta = take(sizeof(int32_t) * META);
  void* p_bc = take(sizeof(int32_t) * nb);
  void* p_bs = take(sizeof(int32_t) * nb);
  void* p_ab = take(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  void* p_sl = take(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  void* p_wr = take(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
  void* p_st = take(sizeof(int32_t) * 4096);
  void* p_un = take(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  void* p_act = take(sizeof(int32_t) * ((size_t)(s > 0 ? s : 1) + 1));
  void* p_wi = take(sizeof(int2) * (size_t)(bin_cap > 0 ? bin_cap : 1));
  void* p_cw = take(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  void* p_wm = take(sizeof(int32_t) * (size_t)(s > 0 ? s : 1));
  if (w) {
    w->geom = (double*)p_geom;
    w->grid = (int32_t*)p_grid;
    w->meta = (int32_t*)p_meta;
    w->bin_count = (int32_t*)p_bc;
    w->bin_start = (int32_t*)p_bs;
    w->atom_bin = (int32_t*)p_ab;
    w->slot = (int32_t*)p_sl;
    w->wr = (double*)p_wr;
    w->scan_tmp = (int32_t*)p_st;
    w->unsorted = (int32_t*)p_un;
    w->act = (int32_t*)p_act;
    w->witem = (int2*)p_wi;
    w->cw = (uint32_t*)p_cw;
    w->wmax = (int32_t*)p_wm;
  }
  return off;
}

// Pack a wrap shift (three ints in [-2^20, 2^20)) into 64 bits.
__device__ __forceinline__ double pack_w(int a, int b, int c) {
  const uint64_t m = (1ull << 21) - 1;
  const uint64_t v = ((uint64_t)(a + (1 << 20)) & m) | (((uint64_t)(b + (1 << 20)) & m) << 21) |
                     (((uint64_t)(c + (1 << 20)) & m) << 42);
  return __longlong_as_double((long long)v);
}
__device__ __forceinline__ void unpack_w(double d, int* w) {
  const uint64_t v = (uint64_t)__double_as_longlong(d);
  const uint64_t m = (1ull << 21) - 1;
  w[0] = (int)(v & m) - (1 << 20);
  w[1] = (int)((v >> 21) & m) - (1 << 20);
  w[2] = (int)((v >> 42) & m) - (1 << 20);
}

// Three two's-complement sbits-wide fields (x, y, z), low field first.
__device__ __forceinline__ uint32_t pack_fields(int x, int y, int z, int sbits) {
  const uint32_t m = (1u << sbits) - 1u;
  return ((uint32_t)x & m) | (((uint32_t)y & m) << sbits) | (((uint32_t)z & m) << (2 * sbits));
}

// Field-wise sum of two packed words, modulo 2^sbits per field: the low bits
// of every field are added with the top bits (H) cleared, so no carry leaves
// a field, and the top bits are added by xor.
__device__ __forceinline__ uint32_t field_add(uint32_t a, uint32_t b, uint32_t H) {
  return ((a & ~H) + (b & ~H)) ^ ((a ^ b) & H);
}

__device__ __forceinline__ bool pbc_of(const uint8_t* pbc, int per_sys, int s, int k) {
  return pbc[per_sys ? 3 * s + k : k] != 0;
}

// ------------------------------------------------------------------ setup
__global__ void k_setup(const double* __restrict__ cell_col, const uint8_t* __restrict__ pbc,
                        int pbc_per_sys, int n_sys, double r_list, int q, int bin_cap,
                        int zero_bins,
                        double* __restrict__ geom, int32_t* __restrict__ grid,
                        int32_t* __restrict__ meta, int32_t* __restrict__ status,
                        int32_t* __restrict__ bin_count, const int32_t* __restrict__ gate,
                        double* __restrict__ ustep) {
  if (gate && *gate == 0) return;
  // every block zeroes part of the bin counters [0, bin_cap]
  if (zero_bins)
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= bin_cap;
         b += (int64_t)gridDim.x * blockDim.x)
      bin_count[b] = 0;
  if (blockIdx.x != 0) return;
  __shared__ int32_t sh_nb[1024];
  __shared__ int32_t sh_carry;
  if (threadIdx.x == 0) sh_carry = 0;
  __syncthreads();
  for (int base = 0; base < n_sys; base += blockDim.x) {
    const int s = base + threadIdx.x;
    int nbins = 0;
    if (s < n_sys) {
      double L[9], inv[9];
      for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 3; ++m) L[3 * k + m] = cell_col[9 * s + 3 * m + k];
      const double det = det33(L);
      const bool any_pbc =
          pbc_of(pbc, pbc_per_sys, s, 0) || pbc_of(pbc, pbc_per_sys, s, 1) ||
          pbc_of(pbc, pbc_per_sys, s, 2);
      const bool invertible = isfinite(det) && fabs(det) > 1e-12;
      const bool active = any_pbc && invertible;
      if (any_pbc && !invertible) atomicOr(status, TSX_STATUS_SINGULAR_CELL);
      // f = r . Linv  with Linv = inverse of L (row convention)
      if (invertible) {
        inv33(L, inv);
      } else {
        for (int k = 0; k < 9; ++k) inv[k] = 0.0;
      }
      double* g = geom + (size_t)GEOM * s;
      for (int k = 0; k < 9; ++k) g[k] = L[k];
      for (int k = 0; k < 9; ++k) g[9 + k] = inv[k];
      int nb[3], mm[3], flags = 0;
      for (int k = 0; k < 3; ++k) {
        // plane spacing along k = 1 / |column k of Linv|
        const double c2 = inv[k] * inv[k] + inv[3 + k] * inv[3 + k] + inv[6 + k] * inv[6 + k];
        const double w = invertible ? 1.0 / sqrt(c2) : 0.0;
        g[18 + k] = w;
        const bool p = active && pbc_of(pbc, pbc_per_sys, s, k);
        if (p) {
          flags |= 1 << k;
          double nbf = floor((double)q * w / r_list);
          if (nbf < 1.0) nbf = 1.0;
          if (nbf > 1024.0) nbf = 1024.0;
          nb[k] = (int)nbf;
          mm[k] = (int)ceil(r_list * nb[k] / w);
        } else {
          nb[k] = 1;
          mm[k] = 0;
        }
      }
      g[21] = fabs(det);
      g[22] = active ? 1.0 : 0.0;
      g[23] = 0.0;
      int32_t* gr = grid + (size_t)GRID * s;
      for (int k = 0; k < 3; ++k) {
        gr[k] = nb[k];
        gr[3 + k] = mm[k];
      }
      gr[7] = flags;
      nbins = nb[0] * nb[1] * nb[2];
      if (ustep) {  // bin step along each lattice vector a_k / nb_k, and sum_k |a_k / nb_k|
        double* u = ustep + 12 * (size_t)s;
        double su = 0.0;
        for (int k = 0; k < 3; ++k) {
          double n2 = 0.0;
          for (int c = 0; c < 3; ++c) {
            u[3 * k + c] = L[3 * k + c] / nb[k];
            n2 += u[3 * k + c] * u[3 * k + c];
          }
          su += sqrt(n2);
        }
        u[9] = su;
        u[10] = u[11] = 0.0;
      }
    }
    // block exclusive scan of nbins
    sh_nb[threadIdx.x] = nbins;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int v = threadIdx.x >= (unsigned)off ? sh_nb[threadIdx.x - off] : 0;
      __syncthreads();
      sh_nb[threadIdx.x] += v;
      __syncthreads();
    }
    const int incl = sh_nb[threadIdx.x];
    if (s < n_sys) grid[(size_t)GRID * s + 6] = sh_carry + incl - nbins;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) sh_carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    meta[0] = sh_carry;
    if (zero_bins && sh_carry > bin_cap) atomicOr(status, TSX_STATUS_BIN_OVERFLOW);
  }
}

// fractional coordinates, wrap shift and wrapped position of one atom
struct WrapOut {
  double f[3];  // wrapped fractional coordinates
  int w[3];     // integer wrap shift
  double wr[3]; // wrapped cartesian position
};

__device__ __forceinline__ WrapOut wrap_atom(const double* __restrict__ r, const double* g,
                                             int flags) {
  WrapOut o;
  const bool active = g[22] != 0.0;
  if (!active) {
    for (int k = 0; k < 3; ++k) {
      o.f[k] = 0.0;
      o.w[k] = 0;
      o.wr[k] = r[k];
    }
    return o;
  }
  double f[3];
  for (int k = 0; k < 3; ++k) f[k] = r[0] * g[9 + k] + r[1] * g[12 + k] + r[2] * g[15 + k];
  for (int k = 0; k < 3; ++k) {
    double fl = 0.0;
    if (flags & (1 << k)) fl = floor(f[k]);
    o.w[k] = (int)fl;
    o.f[k] = f[k] - fl;
  }
  for (int m = 0; m < 3; ++m)
    o.wr[m] = g[m] * o.f[0] + g[3 + m] * o.f[1] + g[6 + m] * o.f[2];
  return o;
}

__device__ __forceinline__ int bin_of(const WrapOut& o, const int32_t* gr) {
  int b[3];
  for (int k = 0; k < 3; ++k) {
    if (gr[7] & (1 << k)) {
      int v = (int)floor(o.f[k] * gr[k]);
      v = v < 0 ? 0 : (v >= gr[k] ? gr[k] - 1 : v);
      b[k] = v;
    } else {
      b[k] = 0;
    }
  }
  return gr[6] + (b[0] * gr[1] + b[1]) * gr[2] + b[2];
}

template <typename Tin, typename T>
__global__ void k_scan_1block(const Tin* __restrict__ in, int64_t n, T* __restrict__ out, const int32_t* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ T sh[1024];
  const int64_t chunk = (n + blockDim.x - 1) / blockDim.x;
  const int64_t lo = (int64_t)threadIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  T s = 0;
  for (int64_t k = lo; k < hi; ++k) s += (T)in[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    T v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : (T)0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  T run = sh[threadIdx.x] - s;
  for (int64_t k = lo; k < hi; ++k) {
    out[k] = run;
    run += (T)in[k];
  }
  if (threadIdx.x == blockDim.x - 1) out[n] = sh[threadIdx.x];
}

// Multi-block exclusive scan for large n: per-block sums, scan of the sums,
// then per-block scan with the carried offset.
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_TILE = SCAN_ITEMS * SCAN_THREADS;

__global__ void k_scan_partials(const int32_t* __restrict__ in, int64_t n,
                                int64_t* __restrict__ part) {
  __shared__ double sh[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  const double t = block_sum((double)s, sh);  // exact for counts < 2^53
  if (threadIdx.x == 0) part[blockIdx.x] = (int64_t)t;
}

__global__ void k_scan_final(const int32_t* __restrict__ in, int64_t n,
                             const int64_t* __restrict__ part_scan, int64_t* __restrict__ out) {
  __shared__ int64_t sh[SCAN_THREADS];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  const int64_t lo = base + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t v[SCAN_ITEMS];
  int64_t s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = lo + k;
    v[k] = idx < n ? in[idx] : 0;
    s += v[k];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    int64_t t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = part_scan[blockIdx.x] + sh[threadIdx.x] - s;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = lo + k;
    if (idx < n) out[idx] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) out[n] = run;
}


__device__ __forceinline__ bool gated_off(const int32_t* __restrict__ gate, int s) {
  return gate != nullptr && gate[s] == 0;
}

// The per-system kernels below run on a grid of n_sys * B blocks: block
// (s, part) handles system s, striding its atoms / bins by B, so systems whose
// gate is off cost one load per block.

// ----------------------------------------------------------- active bins

// One workgroup: act[s] = number of bins of the gated-on systems before s,
// act[n_sys] = their total.  The per-bin fill is a persistent grid that
// walks this compact list of active bins.
__device__ __forceinline__ void active_scan(const int32_t* __restrict__ grid,
                                            const int32_t* __restrict__ meta, int bin_cap,
                                            int n_sys, const int32_t* __restrict__ gate,
                                            int32_t* __restrict__ act, int32_t* sh,
                                            int32_t& carry) {
  // 1024 threads: wave scans, then the 16 wave totals (sh[0..15])
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const bool ok = meta[0] <= bin_cap;
  for (int c0 = 0; c0 < n_sys; c0 += 1024) {
    const int s = c0 + threadIdx.x;
    int v = 0;
    if (ok && s < n_sys && !gated_off(gate, s)) {
      const int32_t* gr = grid + (size_t)GRID * s;
      v = gr[0] * gr[1] * gr[2];
    }
    const int incl = wave_incl_scan(v);
    if (lane == 63) sh[wid] = incl;
    __syncthreads();
    int excl = carry + incl - v;
    for (int q = 0; q < wid; ++q) excl += sh[q];
    if (s < n_sys) act[s] = excl;
    __syncthreads();  // every thread has read carry and sh
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) act[n_sys] = carry;
}

// the scan above as a launch of its own (one 1024-thread workgroup); the
// one-workgroup-per-system build runs it in its first workgroup instead
__global__ __launch_bounds__(1024) void k_active(const int32_t* __restrict__ grid,
                                                 const int32_t* __restrict__ meta, int bin_cap,
                                                 int n_sys, const int32_t* __restrict__ gate,
                                                 int32_t* __restrict__ act) {
  __shared__ int32_t sh[1024];
  __shared__ int32_t carry;
  active_scan(grid, meta, bin_cap, n_sys, gate, act, sh, carry);
}

// system owning active work item w: last s with act[s] <= w (act is
// non-decreasing; systems with no active bins are skipped automatically)
__device__ __forceinline__ int active_system(const int32_t* __restrict__ act, int n_sys, int w) {
  int lo = 0, hi = n_sys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (act[mid] <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// -------------------------------------------------------------------- bin
// Bin by wrapped fractional coordinate, atomic slot, and (engine) the
// displacement reference copy of the position.
__global__ __launch_bounds__(256) void k_bin(
    const double* __restrict__ pos, int ps, const int32_t* __restrict__ sys_start, int n_sys,
    int B, const double* __restrict__ geom, const int32_t* __restrict__ grid,
    const int32_t* __restrict__ meta, int bin_cap, int32_t* __restrict__ bin_count,
    int32_t* __restrict__ atom_bin, int32_t* __restrict__ slot,
    const int32_t* __restrict__ gate, double* __restrict__ xref) {
  const int s = blockIdx.x / B, part = blockIdx.x % B;
  if (s >= n_sys || meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const double* g = geom + (size_t)GEOM * s;
  const int e = sys_start[s + 1];
  for (int i = sys_start[s] + part * blockDim.x + threadIdx.x; i < e; i += B * blockDim.x) {
    const double* r = pos + (int64_t)ps * i;
    const WrapOut o = wrap_atom(r, g, gr[7]);
    const int b = bin_of(o, gr);
    atom_bin[i] = b;
    slot[i] = atomicAdd(&bin_count[b], 1);
    if (xref) {
      double4 v;
      v.x = r[0];
      v.y = r[1];
      v.z = r[2];
      v.w = ps == 4 ? r[3] : 0.0;
      reinterpret_cast<double4*>(xref)[i] = v;
    }
  }
}

// ---------------------------------------------------------- per-system scan
// Exclusive scan of one system's bin counts into bin_start (one workgroup
// per system).  Tiles of 4096 bins: each thread sums 16 consecutive counts,
// a wave shuffle scan and four wave totals give the offsets, so a tile costs
// two barriers (a 1M-atom box has ~4.4e5 bins).
constexpr int SCAN_PER = 16;

__global__ __launch_bounds__(256) void k_scan_sys(
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    const int32_t* __restrict__ sys_start, const int32_t* __restrict__ bin_count,
    int32_t* __restrict__ bin_start, const int32_t* __restrict__ gate,
    int32_t* __restrict__ counter, const double* __restrict__ cell_col,
    double* __restrict__ cell_build, int32_t* __restrict__ wmax) {
  __shared__ int32_t wsum[4];
  __shared__ int32_t carry;
  const int s = blockIdx.x;
  if (meta[0] > bin_cap || gated_off(gate, s)) return;
  if (threadIdx.x == 0) wmax[s] = 0;  // raised by the sort of this build
  const int32_t* gr = grid + (size_t)GRID * s;
  const int off = gr[6], nb = gr[0] * gr[1] * gr[2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = sys_start[s];
  __syncthreads();
  constexpr int TILE = 256 * SCAN_PER;
  for (int c0 = 0; c0 < nb; c0 += TILE) {
    const int base = c0 + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int tsum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      v[k] = base + k < nb ? bin_count[off + base + k] : 0;
      tsum += v[k];
    }
    const int incl = wave_incl_scan(tsum);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int run = carry + incl - tsum;
    for (int q = 0; q < wid; ++q) run += wsum[q];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      if (base + k < nb) bin_start[off + base + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bin_start[off + nb] = sys_start[s + 1];
    if (counter) atomicAdd(counter, 1);
  }
  if (cell_build && threadIdx.x < 9) cell_build[9 * s + threadIdx.x] = cell_col[9 * s + threadIdx.x];
}

// The same per-system scan in three launches for batches of a few large
// systems (one workgroup would walk every bin of a 10^5-bin system): chunk
// scans, then a scan of the chunk totals per system, then the offsets.
// Bit-identical to k_scan_sys (an integer prefix).
constexpr int SCAN_CHUNKS = 256;  // chunks per system

__global__ __launch_bounds__(256) void k_scan_chunk(
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    const int32_t* __restrict__ bin_count, int32_t* __restrict__ bin_start,
    const int32_t* __restrict__ gate, int32_t* __restrict__ part) {
  __shared__ int32_t wsum[4];
  __shared__ int32_t carry;
  const int s = blockIdx.x / SCAN_CHUNKS, j = blockIdx.x % SCAN_CHUNKS;
  if (meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const int off = gr[6], nb = gr[0] * gr[1] * gr[2];
  const int chunk = (nb + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
  const int lo = min(nb, j * chunk), hi = min(nb, lo + chunk);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  constexpr int TILE = 256 * SCAN_PER;
  for (int c0 = lo; c0 < hi; c0 += TILE) {
    const int base = c0 + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int tsum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      v[k] = base + k < hi ? bin_count[off + base + k] : 0;
      tsum += v[k];
    }
    const int incl = wave_incl_scan(tsum);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int run = carry + incl - tsum;
    for (int q = 0; q < wid; ++q) run += wsum[q];
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      if (base + k < hi) bin_start[off + base + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = carry;
}

__global__ __launch_bounds__(SCAN_CHUNKS) void k_scan_parts(
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    const int32_t* __restrict__ sys_start, int32_t* __restrict__ bin_start,
    const int32_t* __restrict__ gate, int32_t* __restrict__ counter,
    const double* __restrict__ cell_col, double* __restrict__ cell_build,
    int32_t* __restrict__ wmax, int32_t* __restrict__ part) {
  __shared__ int32_t wsum[SCAN_CHUNKS / 64];
  const int s = blockIdx.x;
  if (meta[0] > bin_cap || gated_off(gate, s)) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int v = part[s * SCAN_CHUNKS + threadIdx.x];
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int run = sys_start[s] + incl - v;
  for (int q = 0; q < wid; ++q) run += wsum[q];
  part[s * SCAN_CHUNKS + threadIdx.x] = run;
  if (threadIdx.x == 0) {
    // Systems' bins are contiguous, so the end of system s is the first bin
    // of system s + 1, which k_scan_add completes from its local prefix; a
    // write here would be added to.  Only the last system's end is free.
    const int32_t* gr = grid + (size_t)GRID * s;
    if (s == (int)gridDim.x - 1) bin_start[gr[6] + gr[0] * gr[1] * gr[2]] = sys_start[s + 1];
    wmax[s] = 0;  // raised by the sort of this build
    if (counter) atomicAdd(counter, 1);
  }
  if (cell_build && threadIdx.x < 9) cell_build[9 * s + threadIdx.x] = cell_col[9 * s + threadIdx.x];
}

__global__ __launch_bounds__(256) void k_scan_add(
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    int32_t* __restrict__ bin_start, const int32_t* __restrict__ gate,
    const int32_t* __restrict__ part) {
  const int s = blockIdx.x / SCAN_CHUNKS, j = blockIdx.x % SCAN_CHUNKS;
  if (meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const int off = gr[6], nb = gr[0] * gr[1] * gr[2];
  const int chunk = (nb + SCAN_CHUNKS - 1) / SCAN_CHUNKS;
  const int lo = min(nb, j * chunk), hi = min(nb, lo + chunk);
  const int add = part[blockIdx.x];
  for (int b = lo + (int)threadIdx.x; b < hi; b += 256) bin_start[off + b] += add;
}

__global__ __launch_bounds__(256) void k_scatter(
    const int32_t* __restrict__ sys_start, int n_sys, int B, const int32_t* __restrict__ grid,
    const int32_t* __restrict__ meta, int bin_cap, const int32_t* __restrict__ atom_bin,
    const int32_t* __restrict__ slot, const int32_t* __restrict__ bin_start,
    int32_t* __restrict__ unsorted, const int32_t* __restrict__ gate) {
  const int s = blockIdx.x / B, part = blockIdx.x % B;
  if (s >= n_sys || meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const bool single = gr[0] * gr[1] * gr[2] == 1;
  const int e = sys_start[s + 1];
  for (int i = sys_start[s] + part * blockDim.x + threadIdx.x; i < e; i += B * blockDim.x) {
    // one-bin systems keep atom order; otherwise the atomic slot, sorted later
    const int dst = single ? i : bin_start[atom_bin[i]] + slot[i];
    unsorted[dst] = i;
  }
}

// ---------------------------------------------------------- sort + wrap
// One thread per atom of a rebuilt system: its rank among the atoms of its
// bin (by atom index, so the order is unique and does not depend on the
// scatter's atomic slots), its place order[bin_start + rank], its wrapped
// coordinates at that place, its partial list word cw (system-local index in
// the low jbits and above it, for image-shift words, the negated wrap shift
// in the three fields -- a neighbour's word is this plus the image shift and
// the centre's wrap shift, field by field -- or, for stencil words, the
// atom's bin z index), the largest wrap magnitude of the system, the
// bin-relative frame (anchor and rel) when asked for, and the bin counter
// reset for the next build.  Single-bin systems were scattered in atom order
// already.
// Everything a build writes for atom i of system s at bin-order place dst:
// order, wrapped coordinates with the packed wrap shift, the partial list
// word, and the bin-relative frame (anchor, rel) when asked for.  Returns the
// atom's largest wrap magnitude.
// rloc: the atom's position record (x, y, z, w) already in registers, or null
// to read it from pos
__device__ __forceinline__ int emit_atom(int i, int dst, int s, int lb, const WrapOut& o,
                                         const double* __restrict__ pos, int ps,
                                         const double* rloc,
                                         const int32_t* __restrict__ sys_start, const double* g,
                                         const int32_t* gr, int jbits, int sbits, int stencil,
                                         int32_t* __restrict__ order, double* __restrict__ wr,
                                         uint32_t* __restrict__ cw, double* __restrict__ anchor,
                                         float* __restrict__ rel) {
  order[dst] = i;
  double4 v;
  v.x = o.wr[0];
  v.y = o.wr[1];
  v.z = o.wr[2];
  v.w = pack_w(o.w[0], o.w[1], o.w[2]);
  reinterpret_cast<double4*>(wr)[dst] = v;
  const int bz = lb % gr[2], by = (lb / gr[2]) % gr[1], bx = lb / (gr[1] * gr[2]);
  cw[dst] = (uint32_t)(i - sys_start[s]) |
            ((stencil ? (uint32_t)bz : pack_fields(-o.w[0], -o.w[1], -o.w[2], sbits)) << jbits);
  if (anchor || rel) {
    const double* r = rloc ? rloc : pos + (int64_t)ps * i;
    double A[3];
    if (g[22] == 0.0) {  // no lattice: anchor at the system's first atom
      const double* r0 = pos + (int64_t)ps * sys_start[s];
      for (int m = 0; m < 3; ++m) A[m] = r0[m];
    } else {
      const int bb[3] = {bx, by, bz};
      double c[3];
      for (int k = 0; k < 3; ++k) {
        if (gr[7] & (1 << k)) {
          c[k] = (double)o.w[k] + (double)bb[k] / (double)gr[k];
        } else {  // non-periodic direction: the first atom's coordinate
          const double* r0 = pos + (int64_t)ps * sys_start[s];
          c[k] = r0[0] * g[9 + k] + r0[1] * g[12 + k] + r0[2] * g[15 + k];
        }
      }
      for (int m = 0; m < 3; ++m) A[m] = c[0] * g[m] + c[1] * g[3 + m] + c[2] * g[6 + m];
    }
    if (anchor) {
      double4 av;
      av.x = A[0];
      av.y = A[1];
      av.z = A[2];
      av.w = 0.0;
      reinterpret_cast<double4*>(anchor)[i] = av;
    }
    if (rel) {
      float4 rv;
      rv.x = (float)(r[0] - A[0]);
      rv.y = (float)(r[1] - A[1]);
      rv.z = (float)(r[2] - A[2]);
      rv.w = ps == 4 ? __int_as_float((int)__double_as_longlong(r[3])) : 0.f;
      reinterpret_cast<float4*>(rel)[i] = rv;
    }
  }
  return max(abs(o.w[0]), max(abs(o.w[1]), abs(o.w[2])));
}

__global__ __launch_bounds__(256) void k_sort_atom(
    const double* __restrict__ pos, int ps, const int32_t* __restrict__ sys_start, int n_sys,
    int B, const double* __restrict__ geom, const int32_t* __restrict__ grid,
    const int32_t* __restrict__ meta, int bin_cap, const int32_t* __restrict__ atom_bin,
    const int32_t* __restrict__ bin_start, const int32_t* __restrict__ unsorted,
    int32_t* __restrict__ bin_count, int32_t* __restrict__ order, double* __restrict__ wr,
    const int32_t* __restrict__ gate, int jbits, uint32_t* __restrict__ cw,
    int32_t* __restrict__ wmax, int stencil, double* __restrict__ anchor,
    float* __restrict__ rel) {
  const int s = blockIdx.x / B, part = blockIdx.x % B;
  if (s >= n_sys || meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const double* g = geom + (size_t)GEOM * s;
  const bool single = gr[0] * gr[1] * gr[2] == 1;
  const int e = sys_start[s + 1];
  const int sbits = (32 - jbits) / 3;
  int wm = 0;
  for (int i = sys_start[s] + part * blockDim.x + threadIdx.x; i < e; i += B * blockDim.x) {
    const int b = atom_bin[i];
    const int lo = bin_start[b], hi = bin_start[b + 1];
    int dst;
    if (single) {
      dst = i;
    } else {
      int rank = 0;
      for (int k = lo; k < hi; ++k) rank += unsorted[k] < i;
      dst = lo + rank;
    }
    const WrapOut o = wrap_atom(pos + (int64_t)ps * i, g, gr[7]);
    wm = max(wm, emit_atom(i, dst, s, b - gr[6], o, pos, ps, nullptr, sys_start, g, gr, jbits, sbits,
                           stencil, order, wr, cw, anchor, rel));
    bin_count[b] = 0;
  }
  for (int off = 32; off > 0; off >>= 1) wm = max(wm, __shfl_xor(wm, off, 64));
  if ((threadIdx.x & 63) == 0 && wm > 0) atomicMax(&wmax[s], wm);
}

// ------------------------------------------------------ build, one workgroup
// Bin, scan, scatter and in-bin sort of one system by one workgroup, for
// batches whose systems have at most SB_MAXN atoms: per-atom bins and slots
// and the scattered atoms stay in LDS, bin counters too when the system has
// at most SB_MAXB bins (else the global counters, left cleared).  It writes
// what bin + scan_sys + scatter + sort_atom write, in the same order, with
// one launch of one block per system instead of four launches of
// per-system block groups.
constexpr int SB_MAXN = 2048;
constexpr int SB_MAXB = 2048;
constexpr int SB_THREADS = 1024;

// Exclusive scan of v[0..n) in place by the whole block (v[n] = total when
// write_total).
__device__ void block_excl_scan(int32_t* v, int n, int32_t* wsum, int32_t* carry,
                                bool write_total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int W = SB_THREADS / 64;
  constexpr int PER = 4;
  if (threadIdx.x == 0) *carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += SB_THREADS * PER) {
    const int b0 = c0 + threadIdx.x * PER;
    int x[PER];
    int t = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      x[k] = b0 + k < n ? v[b0 + k] : 0;
      t += x[k];
    }
    const int incl = wave_incl_scan(t);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int run = *carry + incl - t;
    for (int q = 0; q < wid; ++q) run += wsum[q];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      if (b0 + k < n) v[b0 + k] = run;
      run += x[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int q = 0; q < W; ++q) tot += wsum[q];
      *carry += tot;
    }
    __syncthreads();
  }
  if (write_total && threadIdx.x == 0) v[n] = *carry;
  __syncthreads();
}

__global__ __launch_bounds__(SB_THREADS) void k_build_small(
    const double* __restrict__ pos, int ps, const int32_t* __restrict__ sys_start, int n_sys,
    const double* __restrict__ geom, const int32_t* __restrict__ grid,
    const int32_t* __restrict__ meta, int bin_cap, int32_t* __restrict__ bin_count,
    int32_t* __restrict__ atom_bin, int32_t* __restrict__ bin_start,
    int32_t* __restrict__ order, double* __restrict__ wr, const int32_t* __restrict__ gate,
    double* __restrict__ xref, int32_t* __restrict__ counter,
    const double* __restrict__ cell_col, double* __restrict__ cell_build, int jbits,
    uint32_t* __restrict__ cw, int32_t* __restrict__ wmax, int stencil,
    double* __restrict__ anchor, float* __restrict__ rel, int32_t* __restrict__ status,
    int32_t* __restrict__ act, int2* __restrict__ witem, const double* __restrict__ red_e,
    double* __restrict__ red_energy, int32_t* __restrict__ red_gate, int64_t* __restrict__ red_step) {
  static_assert(SB_THREADS == 1024, "the active-bin scan needs 1024 threads");
  __shared__ int32_t lcnt[SB_MAXB + 1];  // bin counts, then exclusive starts
  __shared__ int32_t abin[SB_MAXN];      // local bin of each atom
  __shared__ int32_t aslot[SB_MAXN];     // slot of each atom in its bin
  __shared__ int32_t uns[SB_MAXN];       // atoms scattered to their bins
  __shared__ int32_t wsum[SB_THREADS / 64];
  __shared__ int32_t carry;
  __shared__ int32_t s_wm;
  __shared__ double red_part[SB_THREADS / 256][4];
  const int s = blockIdx.x;
  // the per-bin fill's active-bin prefix (gates and grids are final already)
  if (act && s == 0) active_scan(grid, meta, bin_cap, n_sys, gate, act, lcnt, carry);
  // Workgroups past the systems reduce the previous MD step's energies, four
  // systems each, 256 threads per system in tsx_md_reduce's order: a thread
  // sums its strided atoms of the system's first 2048, each wave sums its
  // lanes by the xor tree, the group's first thread adds its four waves.
  if (s >= n_sys) {
    if (!red_e) return;
    constexpr int RT = 256;
    const int g = threadIdx.x / RT, tg = threadIdx.x % RT;
    const int sr = (s - n_sys) * (SB_THREADS / RT) + g;
    if (sr < n_sys) {
      const int64_t lo = sys_start[sr], end = sys_start[sr + 1];
      const int64_t hi = lo + 2048 < end ? lo + 2048 : end;
      double acc = 0.0;
      if (lo < hi) {
        // the (at most eight) loads of a thread are issued together; the
        // adds keep the strided order
        double v[2048 / RT];
#pragma unroll
        for (int u = 0; u < 2048 / RT; ++u) {
          const int64_t k = lo + tg + (int64_t)u * RT;
          v[u] = red_e[k < hi ? k : lo];
        }
#pragma unroll
        for (int u = 0; u < 2048 / RT; ++u)
          if (lo + tg + (int64_t)u * RT < hi) acc += v[u];
      }
      const double t = wave_sum(acc);
      if ((tg & 63) == 0) red_part[g][tg >> 6] = t;
    }
    __syncthreads();
    if (sr < n_sys && tg == 0) {
      double t = 0.0;
      for (int w = 0; w < RT / 64; ++w) t += red_part[g][w];
      red_energy[sr] = t;
      red_gate[sr] = 0;
      if (sr == 0) red_step[0] += 1;
    }
    return;
  }
  if (s >= n_sys || meta[0] > bin_cap || gated_off(gate, s)) return;
  const int32_t* gr = grid + (size_t)GRID * s;
  const double* g = geom + (size_t)GEOM * s;
  const int base = sys_start[s], n_s = sys_start[s + 1] - base;
  // This thread's atoms (a = threadIdx.x + u SB_THREADS, at most two): their
  // position records are loaded once, here, so that this round trip overlaps
  // the work-item prefix's, and the rank pass below reuses them from registers.
  static_assert(SB_MAXN <= 2 * SB_THREADS, "at most two atoms per thread");
  double rr[2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int a = threadIdx.x + u * SB_THREADS;
    for (int k = 0; k < 4; ++k) rr[u][k] = 0.0;
    if (a < n_s && n_s <= SB_MAXN) {  // an empty system reads nothing
      const double* r = pos + (int64_t)ps * (base + a);
      for (int k = 0; k < 3; ++k) rr[u][k] = r[k];
      if (ps == 4) rr[u][3] = r[3];
    }
  }
  if (witem) {
    // this system's work items of the per-bin fill, written before any
    // early exit so that [0, act[n_sys]) has no holes: its active-bin prefix
    // (the bins of the gated-on systems before it), then (s, bin) per bin
    int v = 0;
    for (int k = threadIdx.x; k < s; k += SB_THREADS)
      if (!gated_off(gate, k))
        v += grid[(size_t)GRID * k] * grid[(size_t)GRID * k + 1] * grid[(size_t)GRID * k + 2];
    v = wave_sum_int(v);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    int first = 0;
    for (int q = 0; q < SB_THREADS / 64; ++q) first += wsum[q];
    // a system this path refuses (too large, below) gets bin -1: the fill
    // skips its items instead of reading bin ranges this build never wrote
    const int nb0 = gr[0] * gr[1] * gr[2];
    const bool refused = n_s > SB_MAXN;
    for (int b = threadIdx.x; b < nb0; b += SB_THREADS)
      witem[first + b] = make_int2(s, refused ? -1 : gr[6] + b);
    __syncthreads();  // wsum is reused below
  }
  if (n_s > SB_MAXN) {  // max_system_atoms was wrong: refuse, never overrun LDS
    if (threadIdx.x == 0) atomicOr(status, TSX_STATUS_BAD_SIZE);
    return;
  }
  const int off = gr[6], nb = gr[0] * gr[1] * gr[2];
  const bool lds = nb <= SB_MAXB;  // else the (cleared) global counters
  int32_t* cnt = lds ? lcnt : bin_count + off;
  const int sbits = (32 - jbits) / 3;
  if (lds)
    for (int k = threadIdx.x; k <= nb; k += SB_THREADS) lcnt[k] = 0;
  if (threadIdx.x == 0) s_wm = 0;
  __syncthreads();
  WrapOut ow[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int a = threadIdx.x + u * SB_THREADS;
    if (a >= n_s) break;
    const int i = base + a;
    ow[u] = wrap_atom(rr[u], g, gr[7]);
    const int b = bin_of(ow[u], gr);
    abin[a] = b - off;
    aslot[a] = atomicAdd(&cnt[b - off], 1);
    atom_bin[i] = b;
    if (xref) {
      double4 v;
      v.x = rr[u][0];
      v.y = rr[u][1];
      v.z = rr[u][2];
      v.w = rr[u][3];
      reinterpret_cast<double4*>(xref)[i] = v;
    }
  }
  __syncthreads();
  // exclusive bin starts in place; the global counters of this system are
  // entries [0, nb) only (entry nb is the next system's first bin)
  block_excl_scan(cnt, nb, wsum, &carry, lds);
  for (int k = threadIdx.x; k < nb; k += SB_THREADS) bin_start[off + k] = base + cnt[k];
  if (threadIdx.x == 0) bin_start[off + nb] = base + n_s;
  for (int a = threadIdx.x; a < n_s; a += SB_THREADS) uns[cnt[abin[a]] + aslot[a]] = a;
  __syncthreads();
  int wm = 0;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int a = threadIdx.x + u * SB_THREADS;
    if (a >= n_s) break;
    const int lb = abin[a];
    const int lo = cnt[lb], hi = lb + 1 < nb ? cnt[lb + 1] : n_s;
    int rank = 0;
    for (int k = lo; k < hi; ++k) rank += uns[k] < a;
    const int i = base + a;
    wm = max(wm, emit_atom(i, base + lo + rank, s, lb, ow[u], pos, ps, rr[u], sys_start, g, gr,
                           jbits, sbits, stencil, order, wr, cw, anchor, rel));
  }
  for (int o2 = 32; o2 > 0; o2 >>= 1) wm = max(wm, __shfl_xor(wm, o2, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&s_wm, wm);
  __syncthreads();
  if (!lds)
    for (int k = threadIdx.x; k < nb; k += SB_THREADS) cnt[k] = 0;
  if (threadIdx.x == 0) {
    wmax[s] = s_wm;
    if (counter) atomicAdd(counter, 1);
  }
  if (cell_build && threadIdx.x < 9) cell_build[9 * s + threadIdx.x] = cell_col[9 * s + threadIdx.x];
}

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

// squared distance |a - b - c|^2 in the reference operation order, no FMA
#pragma clang fp contract(off)
__device__ __forceinline__ double dist2_ref(double ax, double ay, double az, double bx,
                                            double by, double bz, double cx, double cy,
                                            double cz) {
  const double dx = (ax - bx) - cx;
  const double dy = (ay - by) - cy;
  const double dz = (az - bz) - cz;
  return (dx * dx + dy * dy) + dz * dz;
}
#pragma clang fp contract(on)

__device__ __forceinline__ int64_t ell_index(int64_t t, int n, int tpa, int64_t ld) {
  return (int64_t)(n / tpa) * (ld * tpa) + t * tpa + (n % tpa);
}


// Component m of the image shift S . L (row cell g) as one FMA chain, the
// same in every fill kernel so that they all write identical rows.
__device__ __forceinline__ double shift_comp(int sx, int sy, int sz, const double* g, int m) {
  return fma((double)sz, g[6 + m], fma((double)sy, g[3 + m], (double)sx * g[m]));
}

// Largest |S_k| over the stencil images of any bin of a system.
__device__ __forceinline__ int stencil_smax(const int32_t* gr) {
  int r = 0;
  for (int k = 0; k < 3; ++k) {
    if (gr[3 + k] == 0) continue;
    const int lo = -floordiv(-gr[3 + k], gr[k]);
    const int hi = floordiv(gr[k] - 1 + gr[3 + k], gr[k]);
    r = max(r, max(lo, hi));
  }
  return r;
}

// Stencil offsets |d_k| <= m_k fit the word fields.
__device__ __forceinline__ bool stencil_fits(const int32_t* gr, int bias) {
  return gr[3] < bias && gr[4] < bias && gr[5] < bias;
}

// Per-system constants of the list word.  A system is "fast" when its image
// shifts and wrap shifts cannot push any field of S' = S + w_i - w_j out of
// range; its words are then field_add(cw_j, S + w_i) with no checks.
struct WordCtx {
  int jbits, sbits, bias;
  uint32_t jmask, H;
};

__device__ __forceinline__ WordCtx word_ctx(int jbits) {
  WordCtx c;
  c.jbits = jbits;
  c.sbits = (32 - jbits) / 3;
  c.bias = 1 << (c.sbits - 1);
  c.jmask = (1u << jbits) - 1u;
  c.H = pack_fields(c.bias, c.bias, c.bias, c.sbits) << jbits;
  return c;
}

// The checked word of one candidate (any wrap magnitude): returns false and
// sets ovf when a field of S' does not fit.
__device__ __forceinline__ bool checked_word(const WordCtx& c, int jl, int sx, int sy, int sz,
                                             const int* w_i, double wj_w, bool half, int il,
                                             bool& match, uint32_t& word, bool& ovf) {
  int w_j[3];
  unpack_w(wj_w, w_j);
  const int ux = sx + w_i[0] - w_j[0];
  const int uy = sy + w_i[1] - w_j[1];
  const int uz = sz + w_i[2] - w_j[2];
  if (match && half) {
    if (il != jl) {
      match = il < jl;
    } else {
      const int first = ux != 0 ? ux : (uy != 0 ? uy : uz);
      match = first > 0;
    }
  }
  if (match && (ux < -c.bias || ux >= c.bias || uy < -c.bias || uy >= c.bias || uz < -c.bias ||
                uz >= c.bias)) {
    ovf = true;
    match = false;
  }
  word = nl_word(jl, ux, uy, uz, c.jbits, c.sbits);
  return match;
}

// The half-list rule on a fast word (self images: first nonzero shift > 0).
__device__ __forceinline__ bool half_keep(const WordCtx& c, uint32_t word, int il, int jl) {
  if (il != jl) return il < jl;
  const int ux = nl_shift(word, c.jbits, c.sbits, 0);
  const int uy = nl_shift(word, c.jbits, c.sbits, 1);
  const int uz = nl_shift(word, c.jbits, c.sbits, 2);
  const int first = ux != 0 ? ux : (uy != 0 ? uy : uz);
  return first > 0;
}

// ------------------------------------------------------------------- fill
// Persistent grid of 256-thread workgroups walking the active-bin list, one
// bin per workgroup at a time.  For one bin the workgroup enumerates the
// stencil images (bin + lattice shift S, m_k bins each way, one image per
// thread), scans their atom counts, writes each candidate's image into a
// table, and stages the candidates of all images in LDS at once (FILL_CH at
// a time; the next work item is loaded a bin ahead); a group of L lanes per
// centre atom (L * centre-atoms-per-round = 256, L <= 64) then tests
// candidates sub, sub + L, ...; a ballot over the group orders the matches,
// so every row
// lists its neighbours in candidate order -- image order (dx, dy, dz
// nested), then atom index -- for any L (deterministic).
constexpr int FILL_T = 256;    // threads per block
// 512 staged candidates (a dense bin stages about 500) keep one workgroup's
// LDS near 25 KB, so six workgroups share a CU: the fill is bound by the
// latency of its per-bin phases, and resident workgroups hide it.
constexpr int FILL_CH = 512;   // staged candidates per chunk
constexpr int FILL_IMG = 128;  // images per batch (one per thread)

// Inclusive prefix sum over the wave of non-negative v, one ballot per bit
// of the largest v: exact, and no lane-address registers.
__device__ __forceinline__ int wave_incl_scan_bits(int v) {
  const uint64_t le = (2ull << (threadIdx.x & 63)) - 1ull;  // lanes <= this one
  int r = 0;
  for (int b = 0; b < 31; ++b) {
    if (__ballot((v >> b) != 0) == 0) break;
    r += __popcll(__ballot((v >> b) & 1) & le) << b;
  }
  return r;
}

template <bool HALF, bool STENCIL>
__global__ __launch_bounds__(FILL_T) void k_fill(
    const int32_t* __restrict__ sys_start, int n_sys, const double* __restrict__ geom,
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    const int32_t* __restrict__ act, const int32_t* __restrict__ bin_start,
    const int32_t* __restrict__ order, const double* __restrict__ wr,
    const uint32_t* __restrict__ cw, const int32_t* __restrict__ wmax, double rl2, int self_int,
    int jbits, int lt, int kmax, int64_t ld, uint32_t* __restrict__ nbr,
    int32_t* __restrict__ count, int32_t* __restrict__ status,
    const int2* __restrict__ witem) {
  constexpr bool stencil = STENCIL;  // list word format (TSX_WORDS_STENCIL)
  __shared__ double4 cpos[FILL_CH];
  __shared__ uint32_t ccw[FILL_CH];
  __shared__ int16_t cimg[FILL_CH];
  __shared__ double ics[FILL_IMG][3];
  __shared__ int32_t isv[FILL_IMG][3];
  __shared__ uint32_t sw[FILL_IMG];
  __shared__ int32_t ibeg[FILL_IMG];
  __shared__ int32_t ipre[FILL_IMG + 1];
  __shared__ int32_t wsum[FILL_T / 64];
  if (meta[0] > bin_cap) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const WordCtx wc = word_ctx(jbits);
  const int64_t stride = ld << lt;
  const int tmask = (1 << lt) - 1;
  const int total = act[n_sys];
  // Work item wi: (system, bin) from the build's table (bin -1: a system the
  // build refused), or found by a search of the active-bin prefix.
  auto item_of = [&](int wi) -> int2 {
    if (witem) return witem[wi];
    const int si = active_system(act, n_sys, wi);
    return make_int2(si, grid[(size_t)GRID * si + 6] + (wi - act[si]));
  };
  // What a bin needs before its image ranges: the system's grid record, the
  // bin's atom range, the system's first atom (and wrap bound), loaded
  // together; the next item's entry is loaded one item ahead, so a bin waits
  // on one round trip before its image ranges instead of a chain of them.
  struct Rec {
    int s, b, c0, c1, base, wm;
    int gr[7];
  };
  auto load_rec = [&](int2 it) {
    Rec r{};
    r.s = it.x;
    r.b = it.y;
    if (it.y >= 0) {
      const int32_t* gp = grid + (size_t)GRID * it.x;
      for (int k = 0; k < 7; ++k) r.gr[k] = gp[k];
      r.c0 = bin_start[it.y];
      r.c1 = bin_start[it.y + 1];
      r.base = sys_start[it.x];
      if (!STENCIL) r.wm = wmax[it.x];
    }
    return r;
  };
  const int first = xcd_block(), step = gridDim.x;
  int2 nxt = make_int2(0, -1);
  if (first < total) nxt = item_of(first);
  for (int wi_ = first; wi_ < total; wi_ += step) {
    const int2 it = nxt;
    if (wi_ + step < total) nxt = item_of(wi_ + step);
    const Rec cur = load_rec(it);
    const int s = cur.s, b = cur.b;
    if (b < 0) continue;  // a system the build refused (uniform over the block)
    const double* g = geom + (size_t)GEOM * s;
    const int* gr = cur.gr;
    const int base = cur.base;
    const int off = gr[6];
    const int wx = 2 * gr[3] + 1, wy = 2 * gr[4] + 1, wz = 2 * gr[5] + 1;
    const int n_img = wx * wy * wz;
    const int c0 = cur.c0, c1 = cur.c1;
    const int nc = c1 - c0;
    if (nc == 0) continue;  // uniform over the block
    const bool fast = stencil ? stencil_fits(gr, wc.bias)
                              : stencil_smax(gr) + 2 * cur.wm < wc.bias;
    if (stencil && !fast) {  // stencil offsets wider than the word fields
      if (tid == 0) atomicOr(status, TSX_STATUS_SHIFT_OVERFLOW);
      continue;
    }
    const int local = b - off;
    const int bz0 = local % gr[2], by0 = (local / gr[2]) % gr[1], bx0 = local / (gr[1] * gr[2]);
    int per = 4;  // centre atoms per round (power of two, L = FILL_T / per <= 64)
    while (per < nc && per < FILL_T) per <<= 1;
    const int L = FILL_T / per;
    const int slot_a = tid / L, sub = tid % L;
    const int gbase = lane & ~(L - 1);
    const uint64_t gmask = (L == 64) ? ~0ull : (((1ull << L) - 1ull) << gbase);
    const uint64_t below = ((1ull << sub) - 1ull) << gbase;
    for (int r0 = c0; r0 < c1; r0 += per) {  // rounds of centre atoms
      const int t = r0 + slot_a;
      const bool own = t < c1;
      // centre loads: first used by the tests, so their latency overlaps
      // the image ranges' loads, the scan and the staging
      double4 wi = make_double4(0.0, 0.0, 0.0, 0.0);
      int i = -1;
      if (own) {
        wi = reinterpret_cast<const double4*>(wr)[t];
        i = order[t];
      }
      const int64_t rowbase = (int64_t)t << lt;
      int cnt = 0;
      bool shift_ovf = false;
      for (int ib0 = 0; ib0 < n_img; ib0 += FILL_IMG) {
        const int nib = min(FILL_IMG, n_img - ib0);
        // image tid of this batch: its bin range is loaded before the barrier
        int n_q = 0, kb = 0, dx = 0, dy = 0, dz = 0, sx = 0, sy = 0, sz = 0;
        if (tid < nib) {
          const int im = ib0 + tid;
          dz = im % wz - gr[5];
          dy = (im / wz) % wy - gr[4];
          dx = im / (wy * wz) - gr[3];
          const int cx = bx0 + dx, cy = by0 + dy, cz = bz0 + dz;
          sx = floordiv(cx, gr[0]);
          sy = floordiv(cy, gr[1]);
          sz = floordiv(cz, gr[2]);
          const int nbin = off + ((cx - sx * gr[0]) * gr[1] + (cy - sy * gr[1])) * gr[2] +
                           (cz - sz * gr[2]);
          kb = bin_start[nbin];
          n_q = bin_start[nbin + 1] - kb;
        }
        __syncthreads();  // previous users of the image and candidate tables are done
        if (tid < nib) {
          // S . L (transforms.py:563-565)
          ics[tid][0] = shift_comp(sx, sy, sz, g, 0);
          ics[tid][1] = shift_comp(sx, sy, sz, g, 1);
          ics[tid][2] = shift_comp(sx, sy, sz, g, 2);
          isv[tid][0] = sx;
          isv[tid][1] = sy;
          isv[tid][2] = sz;
          sw[tid] = (stencil ? pack_fields(dx, dy, dz, wc.sbits) : pack_fields(sx, sy, sz, wc.sbits))
                    << jbits;
          ibeg[tid] = kb;
        }
        // exclusive prefix of the image sizes over the block
        const int incl = wave_incl_scan_bits(n_q);
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int excl = incl - n_q;
        for (int q = 0; q < wid; ++q) excl += wsum[q];
        if (tid < nib) ipre[tid] = excl;
        if (tid == FILL_T - 1) ipre[nib] = excl + n_q;
        // image of each candidate of a chunk, written by the image threads
        // (the first chunk's under the same barrier as the prefix)
        auto image_table = [&](int k0, int nk) {
          if (tid < nib) {
            const int e = min(excl + n_q, k0 + nk);
            for (int c = max(excl, k0); c < e; ++c) cimg[c - k0] = (int16_t)tid;
          }
        };
        image_table(0, FILL_CH);
        __syncthreads();
        const int ncand = ipre[nib];
        for (int k0 = 0; k0 < ncand; k0 += FILL_CH) {
          const int nk = min(FILL_CH, ncand - k0);
          if (k0 > 0) {
            __syncthreads();  // the previous chunk's tests are done
            image_table(k0, nk);
            __syncthreads();
          }
          for (int c = tid; c < nk; c += FILL_T) {
            const int q = cimg[c];
            const int k = ibeg[q] + (k0 + c - ipre[q]);
            cpos[c] = reinterpret_cast<const double4*>(wr)[k];
            ccw[c] = cw[k];
          }
          __syncthreads();
          // the centre's word keys (its loads were issued with the round)
          int w_i[3] = {0, 0, 0};
          if (own) unpack_w(wi.w, w_i);
          const int il = i - base;
          const uint32_t wiw = pack_fields(w_i[0], w_i[1], w_i[2], wc.sbits) << jbits;
          // candidate c of the chunk against this lane's centre atom; the
          // distance of a clamped index is computed unconditionally so that
          // the LDS reads of several candidates can be in flight together
          struct Cand {
            double4 wj;
            int q;
            double d2;
          };
          auto dist = [&](int c) -> Cand {
            Cand r;
            const int cc = min(c, nk - 1);
            r.wj = cpos[cc];
            r.q = cimg[cc];
            r.d2 = dist2_ref(wi.x, wi.y, wi.z, r.wj.x, r.wj.y, r.wj.z, ics[r.q][0], ics[r.q][1],
                             ics[r.q][2]);
            return r;
          };
          auto test = [&](int c, const Cand& k, uint32_t& word) -> bool {
            bool match = false;
            if (own && c < nk && k.d2 < rl2) {
              const int q = k.q;
              const uint32_t cwj = ccw[c];
              const int jl = (int)(cwj & wc.jmask);
              if (fast) {
                const uint32_t swq = sw[q];
                word = stencil ? (uint32_t)jl | swq
                               : (uint32_t)jl |
                                     field_add(cwj & ~wc.jmask, field_add(swq, wiw, wc.H), wc.H);
                match = self_int || !(swq == 0 && jl == il);
                if (HALF && match) match = half_keep(wc, word, il, jl);
              } else {
                const int sx = isv[q][0], sy = isv[q][1], sz = isv[q][2];
                match = self_int || !(jl == il && (sx | sy | sz) == 0);
                checked_word(wc, jl, sx, sy, sz, w_i, k.wj.w, HALF, il, match, word, shift_ovf);
              }
            }
            return match;
          };
          // matches ordered by a ballot over the group, in candidate order
          auto emit = [&](bool match, uint32_t word) {
            const uint64_t bal = __ballot(match);
            if (match) {
              const int pos_n = cnt + __popcll(bal & below);
              if (pos_n < kmax) nbr[rowbase + (int64_t)(pos_n >> lt) * stride + (pos_n & tmask)] = word;
            }
            cnt += __popcll(bal & gmask);
          };
          // uniform trip counts over the block
          if (STENCIL) {
            // two candidates per lane and trip, so the LDS reads of the second
            // overlap the first's arithmetic (68 VGPRs; the other word formats
            // would pass the 80 that six workgroups per CU allow)
            for (int cb = 0; cb < nk; cb += 2 * L) {
              const Cand k0c = dist(cb + sub), k1c = dist(cb + L + sub);
              uint32_t w0 = 0, w1 = 0;
              const bool m0 = test(cb + sub, k0c, w0);
              const bool m1 = test(cb + L + sub, k1c, w1);
              emit(m0, w0);
              if (cb + L < nk) emit(m1, w1);
            }
          } else {
            for (int cb = 0; cb < nk; cb += L) {
              uint32_t w0 = 0;
              const bool m0 = test(cb + sub, dist(cb + sub), w0);
              emit(m0, w0);
            }
          }
        }
      }
      // every lane of a group holds the same count; lane sub == 0 reports
      const uint64_t ovf = __ballot(shift_ovf);
      if (own && sub == 0) {
        if (cnt > kmax) {
          atomicOr(status, TSX_STATUS_NL_OVERFLOW);
          cnt = kmax;
        }
        if (ovf & gmask) atomicOr(status, TSX_STATUS_SHIFT_OVERFLOW);
        count[t] = cnt;
      }
    }
    // the next bin's first image batch opens with a barrier before it
    // rewrites the tables, so none is needed here
  }
}

// ---------------------------------------------------------- fill, per atom
// A group of L lanes per list row t (atom order[t], rows in bin order).  The
// group walks the stencil images of the atom's bin in order; within a
// candidate range lane `sub` tests candidates kb + sub, kb + sub + L, ...; a
// ballot over the group compacts the matches, so the row is written in
// candidate order (the order a single thread would meet them).  Candidates
// are read from L2: rows of one bin share their candidates, so the reads
// coalesce.  No LDS and no staging, so sparse bins cost only their
// candidates, and partial (gated) rebuilds still fill the chip with L lanes
// per rebuilt atom.  Four lanes per atom (dense rows, full builds of large
// systems) run five waves per SIMD (96 VGPRs, 2-3 spilled); two lanes keep
// the default four.
template <int L, bool HALF, bool STENCIL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L == 4 ? 5 : 1)))
void k_fill_atom(
    const int32_t* __restrict__ sys_start, const int32_t* __restrict__ atom_sys,
    int64_t n_atoms, const double* __restrict__ geom, const int32_t* __restrict__ grid,
    const int32_t* __restrict__ meta, int bin_cap, const int32_t* __restrict__ atom_bin,
    const int32_t* __restrict__ bin_start, const int32_t* __restrict__ order,
    const double* __restrict__ wr, const uint32_t* __restrict__ cw,
    const int32_t* __restrict__ wmax, double rl2, int self_int, int jbits, int lt, int kmax,
    int64_t ld, uint32_t* __restrict__ nbr, int32_t* __restrict__ count,
    int32_t* __restrict__ status, const int32_t* __restrict__ gate) {
  constexpr bool stencil = STENCIL;  // list word format (TSX_WORDS_STENCIL)
  const int64_t gid = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t t = gid / L;
  const int sub = (int)(gid % L);
  if (t >= n_atoms || meta[0] > bin_cap) return;
  const int i = order[t];
  const int s = atom_sys[i];
  if (gated_off(gate, s)) return;
  const int lane = threadIdx.x & 63;
  const int gbase = lane & ~(L - 1);
  const uint64_t gmask = (L == 64 ? ~0ull : ((1ull << L) - 1ull)) << gbase;
  const uint64_t below = ((1ull << sub) - 1ull) << gbase;
  const double* g = geom + (size_t)GEOM * s;
  const int32_t* gr = grid + (size_t)GRID * s;
  const int base = sys_start[s];
  const int il = i - base;
  const int off = gr[6];
  const int nbx = gr[0], nby = gr[1], nbz = gr[2];
  const int mx = gr[3], my = gr[4], mz = gr[5];
  const int local = atom_bin[i] - off;
  const int bz0 = local % nbz, by0 = (local / nbz) % nby, bx0 = local / (nby * nbz);
  const WordCtx wc = word_ctx(jbits);
  const bool fast =
      stencil ? stencil_fits(gr, wc.bias) : stencil_smax(gr) + 2 * wmax[s] < wc.bias;
  if (stencil && !fast) {  // stencil offsets wider than the word fields
    if (sub == 0) atomicOr(status, TSX_STATUS_SHIFT_OVERFLOW);
    return;
  }
  const double4 wi = reinterpret_cast<const double4*>(wr)[t];
  int w_i[3];
  unpack_w(wi.w, w_i);
  const uint32_t wiw = pack_fields(w_i[0], w_i[1], w_i[2], wc.sbits) << jbits;
  const int64_t rowbase = t << lt;
  const int64_t stride = ld << lt;
  const int tmask = (1 << lt) - 1;
  int cnt = 0;
  bool shift_ovf = false;
  // Stencil columns (dx, dy); along z the bins of a column are contiguous in
  // memory, so each run of z-bins with one lattice shift sz is one contiguous
  // candidate range (image order dx, dy, dz and bin order are kept).
  for (int dx = -mx; dx <= mx; ++dx) {
    const int cx = bx0 + dx;
    const int sx = floordiv(cx, nbx);
    const int ix = cx - sx * nbx;
    for (int dy = -my; dy <= my; ++dy) {
      const int cy = by0 + dy;
      const int sy = floordiv(cy, nby);
      const int iy = cy - sy * nby;
      const int colbase = off + (ix * nby + iy) * nbz;
      const int cz_end = bz0 + mz;
      for (int cz = bz0 - mz; cz <= cz_end;) {
        const int sz = floordiv(cz, nbz);
        const int iz = cz - sz * nbz;
        const int run = min(nbz - iz, cz_end - cz + 1);
        cz += run;
        const int kb = bin_start[colbase + iz], ke = bin_start[colbase + iz + run];
        if (kb == ke) continue;
        const double c0 = shift_comp(sx, sy, sz, g, 0);
        const double c1 = shift_comp(sx, sy, sz, g, 1);
        const double c2 = shift_comp(sx, sy, sz, g, 2);
        const uint32_t iw = field_add(pack_fields(sx, sy, sz, wc.sbits) << jbits, wiw, wc.H);
        const bool zimg = (sx | sy | sz) == 0;
        // stencil words: dz of a candidate is its bin's z index plus zoff
        const int zoff = sz * nbz - bz0;
        // two candidates per lane in flight; out-of-range lanes re-read the
        // run's first candidate, so every load is unconditional
        for (int k0 = kb; k0 < ke; k0 += 2 * L) {
          const int ka = k0 + sub, kc = k0 + L + sub;
          const bool va = ka < ke, vc = kc < ke;
          const int kar = va ? ka : kb, kcr = vc ? kc : kb;
          const double4 wa = reinterpret_cast<const double4*>(wr)[kar];
          const double4 wb = reinterpret_cast<const double4*>(wr)[kcr];
          const uint32_t ca = cw[kar], cb = cw[kcr];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            if (k0 + u * L >= ke) break;  // uniform over the group
            const double4 wj = u == 0 ? wa : wb;
            const uint32_t cwj = u == 0 ? ca : cb;
            bool match = false;
            uint32_t word = 0;
            const double d2 = dist2_ref(wi.x, wi.y, wi.z, wj.x, wj.y, wj.z, c0, c1, c2);
            if ((u == 0 ? va : vc) && d2 < rl2) {
              const int jl = (int)(cwj & wc.jmask);
              match = self_int || !(zimg && jl == il);
              if (stencil) {
                word = (uint32_t)jl |
                       (pack_fields(dx, dy, (int)(cwj >> jbits) + zoff, wc.sbits) << jbits);
              } else if (fast) {
                word = (uint32_t)jl | field_add(cwj & ~wc.jmask, iw, wc.H);
                if (HALF && match) match = half_keep(wc, word, il, jl);
              } else {
                checked_word(wc, jl, sx, sy, sz, w_i, wj.w, HALF, il, match, word, shift_ovf);
              }
            }
            const uint64_t bal = __ballot(match) & gmask;
            if (match) {
              const int pos = cnt + __popcll(bal & below);
              if (pos < kmax) nbr[rowbase + (int64_t)(pos >> lt) * stride + (pos & tmask)] = word;
            }
            cnt += __popcll(bal);
          }
        }
      }
    }
  }
  const uint64_t ovf = __ballot(shift_ovf) & gmask;
  if (sub != 0) return;
  if (cnt > kmax) {
    atomicOr(status, TSX_STATUS_NL_OVERFLOW);
    cnt = kmax;
  }
  if (ovf) atomicOr(status, TSX_STATUS_SHIFT_OVERFLOW);
  count[t] = cnt;
}

// --------------------------------------------------------- fill, per wave
// Per-bin fill for stencil words (the MD engine's lists) with one wave per
// active bin and every centre atom of the bin on every lane: a lane takes a
// candidate of the bin's stencil (image order, then bin order, 64 per trip),
// reads it once from L2 and tests it against up to NC centre atoms; one
// ballot per centre orders the matches.  The candidate fetch, image lookup
// and word assembly are paid once per candidate instead of once per
// (candidate, centre), and the four waves of a workgroup run independent
// bins with no workgroup barrier.  Rows equal k_fill's entry for entry.
// floor(a / b) for |a| < 2^20 and 1 <= b < 2^20 through one FP32 product,
// with rb the hardware reciprocal of b (within 1 ulp, relative error at most
// 2^-23): x = (a + 1/2) / b = q + (r + 1/2) / b lies at least 1 / (2b) inside
// [q, q + 1), and the product misses x by at most x (2^-23 + 2^-24), which is
// below 1 / (2b) whenever |a| + 1/2 < 2^22 / 3, so it floors to q exactly
// (no integer divide)
__device__ __forceinline__ float rcp_small(int b) { return __builtin_amdgcn_rcpf((float)b); }

__device__ __forceinline__ int floordiv_small(int a, float rb) {
  return (int)floorf(((float)a + 0.5f) * rb);
}

// a wave-uniform double, moved to scalar registers
__device__ __forceinline__ double uniform_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

constexpr int FW_T = 256;     // threads per workgroup
constexpr int FW_W = FW_T / 64;
constexpr int FW_IMG = 128;   // stencil images per batch (two per lane)
constexpr int FW_CH = 1024;   // candidates per image-table chunk
// Centre atoms per pass over the candidates (NC): a bin with more atoms walks
// its stencil's candidates again.  4 suits sparse bins; 8 suits dense ones,
// where fcc atoms on bin faces and corners make a third of the bins hold 5
// or more (one unit cell per bin: 4 atoms on average, up to 10): 1.003
// passes per bin against 1.36 with 4 (and 1.05 with 6).
constexpr int FW_C_SPARSE = 4;
constexpr int FW_C_DENSE = 8;
constexpr double FW_DENSE_OCC = 3.0;  // mean atoms per bin from which NC is 8

// NARROW: row slots addressed by a 24-bit multiply from a per-centre row
// pointer (ld tpa below 2^24 and every slot index of a row below 2^31);
// otherwise by 64-bit arithmetic
// NC = 4 (sparse bins, one large box) at seven waves per SIMD (72 VGPRs, the
// LDS limit); NC = 8 keeps the compiler's allocation (86 VGPRs with two
// candidates per trip: five waves), where forcing six or seven waves cost
// the dense Cu batch 2-4 %.
template <bool NARROW, int NC>
__global__ __launch_bounds__(FW_T)
__attribute__((amdgpu_waves_per_eu(NC == FW_C_SPARSE ? 7 : 1)))
void k_fill_wave(
    const int32_t* __restrict__ sys_start, int n_sys, const double* __restrict__ geom,
    const int32_t* __restrict__ grid, const int32_t* __restrict__ meta, int bin_cap,
    const int32_t* __restrict__ act, const int2* __restrict__ witem,
    const int32_t* __restrict__ bin_start, const double* __restrict__ wr,
    const uint32_t* __restrict__ cw, double rl2, int self_int, int jbits, int lt, int kmax,
    int64_t ld, uint32_t* __restrict__ nbr, int32_t* __restrict__ count,
    int32_t* __restrict__ status) {
  __shared__ int32_t s_ibeg[FW_W][FW_IMG];
  __shared__ int32_t s_ipre[FW_W][FW_IMG];
  __shared__ double s_ics[FW_W][FW_IMG][3];
  __shared__ uint32_t s_sw[FW_W][FW_IMG];
  __shared__ uint8_t s_cimg[FW_W][FW_CH];
  if (meta[0] > bin_cap) return;
  const int lane = threadIdx.x & 63;
  // the wave's index, wave-uniform by construction: stated so that the work
  // items, bin records and centre atoms derived from it load through the
  // scalar cache into SGPRs
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const WordCtx wc = word_ctx(jbits);
  const int64_t stride = ld << lt;
  const int tmask = (1 << lt) - 1;
  const int total_items = act[n_sys];
  const int gstride = gridDim.x * FW_W;
  auto item_of = [&](int wi) -> int2 {
    if (witem) return witem[wi];
    // the grid-wide build writes no items: search the active-bin prefix
    const int si = active_system(act, n_sys, wi);
    return make_int2(si, grid[(size_t)GRID * si + 6] + (wi - act[si]));
  };
  // What a bin needs before its images: the system's grid record and first
  // atom, the bin's atom range.  The next item's record is loaded while the
  // current bin runs, so a bin starts with its image ranges' round trip.
  struct Rec {
    int s, b, c0, c1, base;
    int gr[7];
  };
  auto load_rec = [&](int wi) {
    Rec r{};
    r.b = -1;
    if (wi < total_items) {
      const int2 it = item_of(wi);
      r.s = it.x;
      r.b = it.y;
      if (it.y >= 0) {
        const int32_t* gp = grid + (size_t)GRID * it.x;
        for (int k = 0; k < 7; ++k) r.gr[k] = gp[k];
        r.c0 = bin_start[it.y];
        r.c1 = bin_start[it.y + 1];
        r.base = sys_start[it.x];
      }
    }
    return r;
  };
  // a candidate of the current image chunk: its record and its image's
  // shift and word fields
  struct Cand {
    double4 wj;
    uint32_t cwj;
    int q;
  };
  auto cand_load = [&](int c, int nk, int k0) {
    Cand r;
    const int cc = c < nk ? c : nk - 1;
    r.q = s_cimg[w][cc];
    const int k = s_ibeg[w][r.q] + (k0 + cc - s_ipre[w][r.q]);
    r.wj = reinterpret_cast<const double4*>(wr)[k];
    r.cwj = cw[k];
    return r;
  };
  int wi = xcd_block() * FW_W + w;
  Rec nxt = load_rec(wi);
  for (; wi < total_items; wi += gstride) {
    const Rec cur = nxt;
    nxt = load_rec(wi + gstride);
    if (cur.b < 0) continue;  // a system the build refused (uniform over the wave)
    const int c0 = cur.c0, c1 = cur.c1;
    if (c0 == c1) continue;
    const int* gr = cur.gr;
    if (!stencil_fits(gr, wc.bias)) {  // stencil offsets wider than the word fields
      if (lane == 0) atomicOr(status, TSX_STATUS_SHIFT_OVERFLOW);
      continue;
    }
    const double* g = geom + (size_t)GEOM * cur.s;
    const int off = gr[6];
    const int wx = 2 * gr[3] + 1, wy = 2 * gr[4] + 1, wz = 2 * gr[5] + 1;
    const int n_img = wx * wy * wz;
    const int local = cur.b - off;
    // reciprocals for the bin and image decodes and the lattice shifts
    const float r_wz = rcp_small(wz), r_wyz = rcp_small(wy * wz);
    const float r_nx = rcp_small(gr[0]), r_ny = rcp_small(gr[1]), r_nz = rcp_small(gr[2]);
    int bx0, by0, bz0;
    if (gr[0] * gr[1] * gr[2] <= (1 << 20)) {  // wave-uniform
      const int lxy = floordiv_small(local, r_nz);
      bz0 = local - lxy * gr[2];
      bx0 = floordiv_small(lxy, r_ny);
      by0 = lxy - bx0 * gr[1];
    } else {
      bz0 = local % gr[2];
      by0 = (local / gr[2]) % gr[1];
      bx0 = local / (gr[1] * gr[2]);
    }
    for (int a0 = c0; a0 < c1; a0 += NC) {  // passes of up to NC centre atoms
      const int na = min(NC, c1 - a0);
      double cx[NC], cy[NC], cz[NC];
      int il[NC], cnt[NC];
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        const int t = a0 + (a < na ? a : 0);
        const double4 v = reinterpret_cast<const double4*>(wr)[t];
        cx[a] = uniform_f64(v.x);
        cy[a] = uniform_f64(v.y);
        cz[a] = uniform_f64(v.z);
        il[a] = __builtin_amdgcn_readfirstlane((int)(cw[t] & wc.jmask));
        cnt[a] = 0;
      }
      for (int ib0 = 0; ib0 < n_img; ib0 += FW_IMG) {
        // this lane's two images of the batch: bin range, shift, word fields
        int n_q[2], kb[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int im = ib0 + lane + 64 * h;
          n_q[h] = 0;
          kb[h] = 0;
          if (im < n_img) {
            // unsigned 24-bit products of the non-negative image and bin
            // indices (all far below 2^24)
            const int ix = floordiv_small(im, r_wyz);
            const int iyz = im - (int)__umul24((uint32_t)ix, (uint32_t)(wy * wz));
            const int iy = floordiv_small(iyz, r_wz);
            const int dz = iyz - (int)__umul24((uint32_t)iy, (uint32_t)wz) - gr[5];
            const int dy = iy - gr[4];
            const int dx = ix - gr[3];
            const int qx = bx0 + dx, qy = by0 + dy, qz = bz0 + dz;
            const int shx = floordiv_small(qx, r_nx), shy = floordiv_small(qy, r_ny),
                      shz = floordiv_small(qz, r_nz);
            const uint32_t wx_ = (uint32_t)(qx - shx * gr[0]), wy_ = (uint32_t)(qy - shy * gr[1]),
                           wz_ = (uint32_t)(qz - shz * gr[2]);
            const int nbin = off + (int)(__umul24(__umul24(wx_, (uint32_t)gr[1]) + wy_,
                                                  (uint32_t)gr[2]) +
                                         wz_);
            kb[h] = bin_start[nbin];
            n_q[h] = bin_start[nbin + 1] - kb[h];
            const int q = lane + 64 * h;
            s_ics[w][q][0] = shift_comp(shx, shy, shz, g, 0);
            s_ics[w][q][1] = shift_comp(shx, shy, shz, g, 1);
            s_ics[w][q][2] = shift_comp(shx, shy, shz, g, 2);
            s_sw[w][q] = pack_fields(dx, dy, dz, wc.sbits) << jbits;
          }
        }
        // images in order: lanes' first images (0..63), then their second
        const int inc0 = wave_incl_scan(n_q[0]);
        const int inc1 = wave_incl_scan(n_q[1]);
        const int tot0 = __shfl(inc0, 63, 64);
        const int pre0 = inc0 - n_q[0], pre1 = tot0 + inc1 - n_q[1];
        const int total = tot0 + __shfl(inc1, 63, 64);
        s_ibeg[w][lane] = kb[0];
        s_ibeg[w][lane + 64] = kb[1];
        s_ipre[w][lane] = pre0;
        s_ipre[w][lane + 64] = pre1;
        for (int k0 = 0; k0 < total; k0 += FW_CH) {
          const int nk = min(FW_CH, total - k0);
          // candidate -> image table of this chunk, written by the image lanes
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int p0 = h ? pre1 : pre0;
            const int e = min(p0 + n_q[h], k0 + nk);
            for (int c = max(p0, k0); c < e; ++c) s_cimg[w][c - k0] = (uint8_t)(lane + 64 * h);
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (NC == FW_C_DENSE) {
            // uniform trips over the wave, two candidates per lane: both
            // records in flight together, tested in candidate order (86
            // VGPRs, five waves per SIMD: +0.8 % on the dense Cu batch)
            for (int cb = 0; cb < nk; cb += 128) {
              Cand cds[2];
              cds[0] = cand_load(cb + lane, nk, k0);
              cds[1] = cand_load(cb + 64 + lane, nk, k0);
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                if (h == 1 && cb + 64 >= nk) break;  // uniform over the wave
                const Cand& cd = cds[h];
                const bool valid = cb + 64 * h + lane < nk;
                const double sx = s_ics[w][cd.q][0], sy = s_ics[w][cd.q][1], sz = s_ics[w][cd.q][2];
                const uint32_t swq = s_sw[w][cd.q];
                const int jl = (int)(cd.cwj & wc.jmask);
                const uint32_t word = (uint32_t)jl | swq;
#pragma unroll
                for (int a = 0; a < NC; ++a) {
                  if (a >= na) break;  // uniform over the wave
                  const double d2 = dist2_ref(cx[a], cy[a], cz[a], cd.wj.x, cd.wj.y, cd.wj.z, sx, sy, sz);
                  const bool match = valid && d2 < rl2 && (self_int || !(swq == 0 && jl == il[a]));
                  const uint64_t bal = __ballot(match);
                  if (match) {
                    const int pos = cnt[a] + (int)__builtin_amdgcn_mbcnt_hi(
                                                 (uint32_t)(bal >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    if (pos < kmax) {
                      if (NARROW) {
                        uint32_t* rowp = nbr + ((int64_t)(a0 + a) << lt);  // wave-uniform
                        rowp[__umul24((uint32_t)(pos >> lt), (uint32_t)stride) +
                             (uint32_t)(pos & tmask)] = word;
                      } else {
                        nbr[((int64_t)(a0 + a) << lt) + (int64_t)(pos >> lt) * stride +
                            (pos & tmask)] = word;
                      }
                    }
                  }
                  cnt[a] += __popcll(bal);
                }
              }
            }
          } else {
            for (int cb = 0; cb < nk; cb += 64) {  // uniform trips over the wave
              const Cand cd = cand_load(cb + lane, nk, k0);
              const bool valid = cb + lane < nk;
              const double sx = s_ics[w][cd.q][0], sy = s_ics[w][cd.q][1], sz = s_ics[w][cd.q][2];
              const uint32_t swq = s_sw[w][cd.q];
              const int jl = (int)(cd.cwj & wc.jmask);
              const uint32_t word = (uint32_t)jl | swq;
#pragma unroll
              for (int a = 0; a < NC; ++a) {
                if (a >= na) break;  // uniform over the wave
                const double d2 = dist2_ref(cx[a], cy[a], cz[a], cd.wj.x, cd.wj.y, cd.wj.z, sx, sy, sz);
                const bool match = valid && d2 < rl2 && (self_int || !(swq == 0 && jl == il[a]));
                const uint64_t bal = __ballot(match);
                if (match) {
                  const int pos = cnt[a] + (int)__builtin_amdgcn_mbcnt_hi(
                                               (uint32_t)(bal >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                  if (pos < kmax) {
                    if (NARROW) {
                      uint32_t* rowp = nbr + ((int64_t)(a0 + a) << lt);  // wave-uniform
                      rowp[__umul24((uint32_t)(pos >> lt), (uint32_t)stride) +
                           (uint32_t)(pos & tmask)] = word;
                    } else {
                      nbr[((int64_t)(a0 + a) << lt) + (int64_t)(pos >> lt) * stride +
                          (pos & tmask)] = word;
                    }
                  }
                }
                cnt[a] += __popcll(bal);
              }
            }
          }
          // every lane's reads of this chunk's table are done before it is rewritten
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      }
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        if (a >= na) break;
        if (lane == 0) {
          int cn = cnt[a];
          if (cn > kmax) {
            atomicOr(status, TSX_STATUS_NL_OVERFLOW);
            cn = kmax;
          }
          count[a0 + a] = cn;
        }
      }
    }
  }
}

template <int NC>
static int fill_wave_resident_blocks() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  if (cache[dev] > 0) return cache[dev];
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_fill_wave<true, NC>, FW_T, 0) != hipSuccess ||
      cus <= 0 || per <= 0)
    return 1024;
  cache[dev] = cus * per;
  return cache[dev];
}

// Workgroups of the per-bin fill that fit on the device at once: its grid is
// persistent, so a larger grid only adds a second, serial round of
// workgroups.  Queried once per device and HALF variant.
template <bool HALF, bool STENCIL>
static int fill_resident_blocks() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  if (cache[dev] > 0) return cache[dev];
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_fill<HALF, STENCIL>, FILL_T, 0) !=
          hipSuccess ||
      cus <= 0 || per <= 0)
    return 1024;
  cache[dev] = cus * per;
  return cache[dev];
}
template <int L, bool HALF>
static void launch_fill_atom(hipStream_t st, const int32_t* sys_start, const int32_t* atom_sys,
                             int64_t n, const NLWs& w, int bin_cap, const int32_t* order,
                             double rl2, int self_int, int jbits, int lt, int kmax, int64_t ld,
                             uint32_t* nbr, int32_t* count, int32_t* status,
                             const int32_t* gate, int stencil) {
  // stencil words come with full lists only (checked by the caller)
  if (stencil)
    k_fill_atom<L, false, true><<<blocks_for(n * L, 256), 256, 0, st>>>(
        sys_start, atom_sys, n, w.geom, w.grid, w.meta, bin_cap, w.atom_bin, w.bin_start, order,
        w.wr, w.cw, w.wmax, rl2, self_int, jbits, lt, kmax, ld, nbr, count, status, gate);
  else
    k_fill_atom<L, HALF, false><<<blocks_for(n * L, 256), 256, 0, st>>>(
        sys_start, atom_sys, n, w.geom, w.grid, w.meta, bin_cap, w.atom_bin, w.bin_start, order,
        w.wr, w.cw, w.wmax, rl2, self_int, jbits, lt, kmax, ld, nbr, count, status, gate);
}

__global__ void k_emit(const uint32_t* __restrict__ nbr, const int32_t* __restrict__ count,
                       const int32_t* __restrict__ order, const int64_t* __restrict__ row_ptr,
                       const int32_t* __restrict__ sys_start, const int32_t* __restrict__ atom_sys,
                       int64_t n_atoms, int jbits, int tpa, int64_t ld, int64_t* __restrict__ map,
                       int64_t n_pairs, int64_t* __restrict__ sysmap, double* __restrict__ shifts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_atoms) return;
  const int i = order[t];
  const int s = atom_sys[i];
  const int base = sys_start[s];
  const int sbits = (32 - jbits) / 3;
  const uint32_t jmask = (jbits >= 32) ? 0xffffffffu : ((1u << jbits) - 1u);
  int64_t p = row_ptr[t];
  const int c = count[t];
  for (int n = 0; n < c; ++n, ++p) {
    const uint32_t w = nbr[ell_index(t, n, tpa, ld)];
    map[p] = i;
    map[n_pairs + p] = base + (int)(w & jmask);
    sysmap[p] = s;
    shifts[3 * p + 0] = (double)nl_shift(w, jbits, sbits, 0);
    shifts[3 * p + 1] = (double)nl_shift(w, jbits, sbits, 1);
    shifts[3 * p + 2] = (double)nl_shift(w, jbits, sbits, 2);
  }
}

__global__ void k_wrap_out(const double* __restrict__ pos, const double* __restrict__ geom,
                           const int32_t* __restrict__ grid, const int32_t* __restrict__ atom_sys,
                           int64_t n, double* __restrict__ wrapped, double* __restrict__ shifts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = atom_sys[i];
  const WrapOut o =
      wrap_atom(pos + 3 * i, geom + (size_t)GEOM * s, grid[(size_t)GRID * s + 7]);
  for (int k = 0; k < 3; ++k) {
    wrapped[3 * i + k] = o.wr[k];
    shifts[3 * i + k] = (double)o.w[k];
  }
}

// ------------------------------------------------------- Verlet skin check
// A list built at cell Lb with radius r_list stays valid for cutoff rc while
// every atom, pulled back to the build cell (Lb L^-1 r), is within thr of its
// position at the build: pairs outside the list were >= r_list apart then,
// so they are >= sigma (r_list - 2 thr) >= rc apart now, where
//   thr = (r_list - rc / sigma) / 2,   sigma = 1 - ||L Lb^-1 - I||_F
// bounds the smallest singular value of the cell map from below.
constexpr int CHECK_W = 10;  // per-system workspace: pull-back [9], thr^2

__global__ void k_check_sys(const double* __restrict__ cell_col,
                            const double* __restrict__ cell_build, int n_sys, double r_list,
                            double cutoff, double* __restrict__ ws, int32_t* __restrict__ flags) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const M33 cur = m33_load(cell_col + 9 * (int64_t)s);
  const M33 bld = m33_load(cell_build + 9 * (int64_t)s);
  const M33 a = m33_mul(cur, m33_inv(bld));
  const double sig = 1.0 - m33_frob(m33_sub(a, m33_eye()));
  double* w = ws + (int64_t)CHECK_W * s;
  m33_store(m33_mul(bld, m33_inv(cur)), w);
  int flag = 0;
  if (!(sig > 0.0) || !(sig * r_list > cutoff)) {
    w[9] = -1.0;
    flag = 1;
  } else {
    const double thr = 0.5 * (r_list - cutoff / sig);
    w[9] = thr * thr;
  }
  flags[s] = flag;
}

__global__ void k_check_atoms(const double* __restrict__ pos, int ps,
                              const double* __restrict__ xref,
                              const int32_t* __restrict__ atom_sys, int64_t n,
                              const double* __restrict__ ws, int32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = atom_sys[i];
  const double* w = ws + (int64_t)CHECK_W * s;
  const double* r = pos + (int64_t)ps * i;
  double rt[3];
  m33_apply(m33_load(w), r, rt);
  const double4 x0 = reinterpret_cast<const double4*>(xref)[i];
  const double dx = rt[0] - x0.x, dy = rt[1] - x0.y, dz = rt[2] - x0.z;
  if (!(dx * dx + dy * dy + dz * dz <= w[9])) flags[s] = 1;
}

// ------------------------------------------------- no-allocation build
// Workspace of tsx_nl_build_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  In host code ``scratch`` names a host
// stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, looked up once per device (with that device current)
// and kept in g_scratch_ptr; only that address is handed to kernels.
constexpr size_t NOALLOC_WS_BYTES = 36ull << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_WS_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of ``scratch`` on the current device (nullptr and an error
// message on failure).
static void* scratch_on_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("tsx_nl_build_noalloc: no usable current device (%s, device %d)",
              hipGetErrorString(e), dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("tsx_nl_build_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace tsx

using namespace tsx;

extern "C" {

int64_t tsx_nl_check_ws_bytes(int32_t n_systems) {
  return (int64_t)sizeof(double) * CHECK_W * (n_systems > 0 ? n_systems : 1);
}

int tsx_nl_check(const double* pos, int32_t pos_stride, const double* cell_col,
                 const double* cell_build, const double* xref, const int32_t* atom_sys,
                 int64_t n_atoms, int32_t n_systems, double r_list, double cutoff,
                 int32_t* flags, void* ws, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(pos_stride == 3 || pos_stride == 4, "pos_stride must be 3 or 4");
  TSX_REQUIRE(r_list > cutoff && cutoff > 0.0, "need r_list > cutoff > 0");
  TSX_REQUIRE(ws && flags && cell_build && xref, "ws, flags, cell_build and xref required");
  hipStream_t st = (hipStream_t)stream;
  k_check_sys<<<blocks_for(n_systems, 64), 64, 0, st>>>(cell_col, cell_build, n_systems, r_list,
                                                        cutoff, (double*)ws, flags);
  TSX_CHECK_LAUNCH();
  if (n_atoms > 0) {
    k_check_atoms<<<blocks_for(n_atoms, 256), 256, 0, st>>>(pos, pos_stride, xref, atom_sys,
                                                            n_atoms, (const double*)ws, flags);
    TSX_CHECK_LAUNCH();
  }
  return TSX_OK;
}

int tsx_nl_build_noalloc(const double* pos, int32_t pos_stride, const double* cell_col,
                         const uint8_t* pbc, int32_t pbc_per_system, const int32_t* sys_start,
                         const int32_t* atom_sys, int64_t n_atoms, int32_t n_systems,
                         double r_list, int32_t bins_per_rlist, int32_t self_interaction,
                         int32_t half, int32_t jbits, int32_t tpa, int32_t kmax, int64_t ld,
                         int32_t bin_cap, uint32_t* nbr, int32_t* count, int32_t* order,
                         double* xref, int32_t* status, const int32_t* gate, int32_t* counter,
                         double* cell_build, int32_t fill_mode, int32_t word_mode,
                         const tsx_nl_frame* frame, int32_t max_system_atoms, void* stream) {
  TSX_REQUIRE(n_atoms >= 0 && n_atoms <= TSX_NOALLOC_MAX_ATOMS,
              "tsx_nl_build_noalloc: n_atoms %lld outside [0, %d]", (long long)n_atoms,
              TSX_NOALLOC_MAX_ATOMS);
  TSX_REQUIRE(n_systems > 0 && n_systems <= TSX_NOALLOC_MAX_SYSTEMS,
              "tsx_nl_build_noalloc: n_systems %d outside [1, %d]", n_systems,
              TSX_NOALLOC_MAX_SYSTEMS);
  TSX_REQUIRE(bin_cap >= 1 && bin_cap <= TSX_NOALLOC_MAX_BINS,
              "tsx_nl_build_noalloc: bin_cap %d outside [1, %d]", bin_cap, TSX_NOALLOC_MAX_BINS);
  TSX_REQUIRE(ws_layout(n_atoms, n_systems, bin_cap, nullptr, nullptr) <= NOALLOC_WS_BYTES,
              "tsx_nl_build_noalloc: workspace exceeds the static scratch");
  void* ws = scratch_on_current_device();
  if (!ws) return TSX_ERR_HIP;
  // the scratch holds whatever the last call on this device left: rebuild the
  // per-system geometry and clear the bin counters every time (setup = 1)
  return tsx_nl_build(pos, pos_stride, cell_col, pbc, pbc_per_system, sys_start, atom_sys,
                      n_atoms, n_systems, r_list, bins_per_rlist, self_interaction, half, jbits,
                      tpa, kmax, ld, bin_cap, 1, ws, nbr, count, order, xref, status, gate,
                      counter, cell_build, fill_mode, word_mode, frame, max_system_atoms, stream);
}

int tsx_version(void) { return 100; }

const char* tsx_last_error(void) { return tsx::g_err; }

int64_t tsx_nl_ws_bytes(int64_t n_atoms, int32_t n_systems, int32_t bin_cap) {
  return (int64_t)ws_layout(n_atoms, n_systems, bin_cap, nullptr, nullptr);
}

int tsx_nl_build(const double* pos, int32_t pos_stride, const double* cell_col,
                 const uint8_t* pbc, int32_t pbc_per_system, const int32_t* sys_start,
                 const int32_t* atom_sys, int64_t n_atoms, int32_t n_systems, double r_list,
                 int32_t bins_per_rlist, int32_t self_interaction, int32_t half, int32_t jbits,
                 int32_t tpa, int32_t kmax, int64_t ld, int32_t bin_cap, int32_t setup, void* ws,
                 uint32_t* nbr, int32_t* count, int32_t* order, double* xref, int32_t* status,
                 const int32_t* gate, int32_t* counter, double* cell_build, int32_t fill_mode,
                 int32_t word_mode, const tsx_nl_frame* frame, int32_t max_system_atoms,
                 void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(n_atoms >= 0 && n_atoms < (1ll << 31), "n_atoms out of range");
  TSX_REQUIRE(pos_stride == 3 || pos_stride == 4, "pos_stride must be 3 or 4");
  TSX_REQUIRE(jbits >= 1 && jbits <= 26, "jbits must be in [1, 26]");
  TSX_REQUIRE(tpa == 1 || tpa == 2 || tpa == 4 || tpa == 8 || tpa == 16,
              "tpa must be a power of two <= 16");
  TSX_REQUIRE(ld >= n_atoms, "ld must be >= n_atoms");
  TSX_REQUIRE(kmax >= 0 && kmax % tpa == 0, "kmax must be a non-negative multiple of tpa");
  TSX_REQUIRE(r_list > 0.0, "r_list must be positive");
  TSX_REQUIRE(bins_per_rlist >= 1 && bins_per_rlist <= 4, "bins_per_rlist in [1,4]");
  TSX_REQUIRE(bin_cap >= 1, "bin_cap must be >= 1");
  TSX_REQUIRE(fill_mode >= TSX_FILL_AUTO && fill_mode <= TSX_FILL_WAVE_DENSE, "unknown fill_mode");
  TSX_REQUIRE(word_mode == TSX_WORDS_SHIFT || word_mode == TSX_WORDS_STENCIL,
              "unknown word_mode");
  TSX_REQUIRE(word_mode == TSX_WORDS_SHIFT || (!half && jbits <= 22),
              "stencil words need a full list and jbits <= 22");
  hipStream_t st = (hipStream_t)stream;
  NLWs w;
  ws_layout(n_atoms, n_systems, bin_cap, &w, (char*)ws);
  if (setup) {
    const int zb = blocks_for((int64_t)bin_cap + 1, 1024);
    k_setup<<<zb < 1024 ? zb : 1024, 1024, 0, st>>>(
        cell_col, pbc, pbc_per_system, n_systems, r_list, bins_per_rlist, bin_cap, 1, w.geom,
        w.grid, w.meta, status, w.bin_count, nullptr, frame ? frame->ustep : nullptr);
    TSX_CHECK_LAUNCH();
  }
  TSX_REQUIRE(!(frame && frame->e_atom) || n_atoms > 0,
              "the build's per-system reduction needs atoms");
  if (n_atoms == 0) return TSX_OK;
  // Fill kernel: per-bin LDS staging pays off for dense lists of many
  // systems, where gated rebuilds touch a few percent of the atoms and one
  // workgroup per bin shares the staged candidates among the bin's atoms; the
  // per-atom kernel wins for sparse bins and for single large systems.  With
  // stencil words the one-wave-per-bin fill takes every dense list, one large
  // system included (the 1M-atom LJ box: 2.58e9 against 2.46e9 atom-steps/s
  // with the per-atom fill).
  bool per_bin = fill_mode == TSX_FILL_BIN;
  const bool dense = bins_per_rlist >= 2 && kmax >= 64;
  if (fill_mode == TSX_FILL_AUTO) per_bin = dense && n_systems > 1;
  const bool small = max_system_atoms > 0 && max_system_atoms <= SB_MAXN;
  const bool per_wave = fill_mode >= TSX_FILL_WAVE ||
                        (fill_mode == TSX_FILL_AUTO && dense && word_mode == TSX_WORDS_STENCIL);
  TSX_REQUIRE(!per_wave || word_mode == TSX_WORDS_STENCIL, "TSX_FILL_WAVE needs stencil words");
  if (per_wave) per_bin = true;  // the per-bin work items
  const bool red = frame && frame->e_atom;
  TSX_REQUIRE(!red || small, "the build's per-system reduction needs max_system_atoms in (0, 2048]");
  TSX_REQUIRE(!red || (frame->energy && frame->gate_clear && frame->step),
              "the build's per-system reduction needs energy, gate_clear and step");
  if (small) {
    // every system fits one workgroup: bin, scan, scatter and sort in one launch
    // past the systems' workgroups, a quarter as many reduce energies
    const int n_red = red ? (n_systems + SB_THREADS / 256 - 1) / (SB_THREADS / 256) : 0;
    k_build_small<<<n_systems + n_red, SB_THREADS, 0, st>>>(
        pos, pos_stride, sys_start, n_systems, w.geom, w.grid, w.meta, bin_cap, w.bin_count,
        w.atom_bin, w.bin_start, order, w.wr, gate, xref, counter, cell_col, cell_build, jbits,
        w.cw, w.wmax, word_mode, frame ? frame->anchor : nullptr, frame ? frame->rel : nullptr,
        status, per_bin ? w.act : nullptr, per_bin ? w.witem : nullptr,
        red ? frame->e_atom : nullptr, red ? frame->energy : nullptr,
        red ? frame->gate_clear : nullptr, red ? frame->step : nullptr);
    TSX_CHECK_LAUNCH();
  } else {
    // blocks per system: whole-grid parallelism for one large system, a few
    // blocks each for hundreds of small ones
    const int b_atoms = n_systems >= 2048 ? 1 : 2048 / n_systems;
    k_bin<<<n_systems * b_atoms, 256, 0, st>>>(pos, pos_stride, sys_start, n_systems, b_atoms,
                                               w.geom, w.grid, w.meta, bin_cap, w.bin_count,
                                               w.atom_bin, w.slot, gate, xref);
    TSX_CHECK_LAUNCH();
    if (n_systems <= 4096 / SCAN_CHUNKS) {  // a few large systems: chunked scan
      k_scan_chunk<<<n_systems * SCAN_CHUNKS, 256, 0, st>>>(w.grid, w.meta, bin_cap, w.bin_count,
                                                             w.bin_start, gate, w.scan_tmp);
      TSX_CHECK_LAUNCH();
      k_scan_parts<<<n_systems, SCAN_CHUNKS, 0, st>>>(w.grid, w.meta, bin_cap, sys_start,
                                                      w.bin_start, gate, counter, cell_col,
                                                      cell_build, w.wmax, w.scan_tmp);
      TSX_CHECK_LAUNCH();
      k_scan_add<<<n_systems * SCAN_CHUNKS, 256, 0, st>>>(w.grid, w.meta, bin_cap, w.bin_start,
                                                           gate, w.scan_tmp);
      TSX_CHECK_LAUNCH();
    } else {
      k_scan_sys<<<n_systems, 256, 0, st>>>(w.grid, w.meta, bin_cap, sys_start, w.bin_count,
                                            w.bin_start, gate, counter, cell_col, cell_build,
                                            w.wmax);
      TSX_CHECK_LAUNCH();
    }
    k_scatter<<<n_systems * b_atoms, 256, 0, st>>>(sys_start, n_systems, b_atoms, w.grid, w.meta,
                                                   bin_cap, w.atom_bin, w.slot, w.bin_start,
                                                   w.unsorted, gate);
    TSX_CHECK_LAUNCH();
    k_sort_atom<<<n_systems * b_atoms, 256, 0, st>>>(pos, pos_stride, sys_start, n_systems, b_atoms,
                                                     w.geom, w.grid, w.meta, bin_cap, w.atom_bin,
                                                     w.bin_start, w.unsorted, w.bin_count, order,
                                                     w.wr, gate, jbits, w.cw, w.wmax, word_mode,
                                                     frame ? frame->anchor : nullptr,
                                                     frame ? frame->rel : nullptr);
    TSX_CHECK_LAUNCH();
  }
  const double rl2 = r_list * r_list;
  int lt = 0;
  while ((1 << lt) < tpa) ++lt;
  if (per_bin) {
    if (!small) {  // the one-workgroup-per-system build wrote act already
      k_active<<<1, 1024, 0, st>>>(w.grid, w.meta, bin_cap, n_systems, gate, w.act);
      TSX_CHECK_LAUNCH();
    }
    // stencil words come with full lists only (checked above)
    const bool sten = word_mode == TSX_WORDS_STENCIL;
    if (per_wave) {
      // centres per pass from the mean bin occupancy (bin_cap is the caller's
      // bin count plus a margin) unless the mode names it
      const bool dense_bins = fill_mode == TSX_FILL_WAVE_DENSE ||
                              (fill_mode != TSX_FILL_WAVE_SPARSE &&
                               (double)n_atoms >= FW_DENSE_OCC * (double)bin_cap);
      const int g_res_w = dense_bins ? fill_wave_resident_blocks<FW_C_DENSE>()
                                     : fill_wave_resident_blocks<FW_C_SPARSE>();
      const int g_w = (bin_cap + FW_W - 1) / FW_W < g_res_w ? (bin_cap + FW_W - 1) / FW_W : g_res_w;
      const int64_t stride_w = ld * tpa;
      const bool narrow = stride_w < (1 << 24) && (int64_t)(kmax / tpa + 1) * stride_w < (1ll << 31);
      auto launch = [&](auto kern) {
        kern<<<g_w, FW_T, 0, st>>>(sys_start, n_systems, w.geom, w.grid, w.meta, bin_cap, w.act,
                                   small ? w.witem : nullptr, w.bin_start, w.wr, w.cw, rl2,
                                   self_interaction, jbits, lt, kmax, ld, nbr, count, status);
      };
      if (narrow && dense_bins)
        launch(k_fill_wave<true, FW_C_DENSE>);
      else if (narrow)
        launch(k_fill_wave<true, FW_C_SPARSE>);
      else if (dense_bins)
        launch(k_fill_wave<false, FW_C_DENSE>);
      else
        launch(k_fill_wave<false, FW_C_SPARSE>);
      TSX_CHECK_LAUNCH();
      return TSX_OK;
    }
    const int g_res = half   ? fill_resident_blocks<true, false>()
                      : sten ? fill_resident_blocks<false, true>()
                             : fill_resident_blocks<false, false>();
    const int g_fill = bin_cap < g_res ? bin_cap : g_res;
#define TSX_FILL_BIN_LAUNCH(HH, SS)                                                               \
  k_fill<HH, SS><<<g_fill, FILL_T, 0, st>>>(sys_start, n_systems, w.geom, w.grid, w.meta, bin_cap, \
                                            w.act, w.bin_start, order, w.wr, w.cw, w.wmax, rl2,    \
                                            self_interaction, jbits, lt, kmax, ld, nbr, count,     \
                                            status, small ? w.witem : nullptr)
    if (half)
      TSX_FILL_BIN_LAUNCH(true, false);
    else if (sten)
      TSX_FILL_BIN_LAUNCH(false, true);
    else
      TSX_FILL_BIN_LAUNCH(false, false);
#undef TSX_FILL_BIN_LAUNCH
  } else {
    // lanes per atom from the row capacity (dense lists: more candidates per
    // column); fewer lanes leave more atoms per wave for the column arithmetic
#define TSX_FILL_ATOM_LAUNCH(LL, HH)                                                       \
  launch_fill_atom<LL, HH>(st, sys_start, atom_sys, n_atoms, w, bin_cap, order, rl2,       \
                           self_interaction, jbits, lt, kmax, ld, nbr, count, status, gate,   \
                           word_mode)
    if (kmax >= 64) {
      if (half) TSX_FILL_ATOM_LAUNCH(4, true);
      else TSX_FILL_ATOM_LAUNCH(4, false);
    } else {
      if (half) TSX_FILL_ATOM_LAUNCH(2, true);
      else TSX_FILL_ATOM_LAUNCH(2, false);
    }
#undef TSX_FILL_ATOM_LAUNCH
  }
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int64_t tsx_scan_ws_bytes(int64_t n) {
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  return (int64_t)sizeof(int64_t) * (2 * tiles + 4);
}

int tsx_nl_count_pairs(const int32_t* count, int64_t n, int64_t* row_ptr, void* ws,
                       void* stream) {
  TSX_REQUIRE(n >= 0, "n must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles <= 1) {
    k_scan_1block<int32_t, int64_t><<<1, 1024, 0, st>>>(count, n, row_ptr, nullptr);
    TSX_CHECK_LAUNCH();
    return TSX_OK;
  }
  TSX_REQUIRE(ws != nullptr, "workspace required for n > %d", SCAN_TILE);
  int64_t* part = (int64_t*)ws;
  int64_t* part_scan = part + tiles + 1;
  k_scan_partials<<<(int)tiles, SCAN_THREADS, 0, st>>>(count, n, part);
  TSX_CHECK_LAUNCH();
  k_scan_1block<int64_t, int64_t><<<1, 1024, 0, st>>>(part, tiles, part_scan, nullptr);
  TSX_CHECK_LAUNCH();
  k_scan_final<<<(int)tiles, SCAN_THREADS, 0, st>>>(count, n, part_scan, row_ptr);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int tsx_nl_emit_pairs(const uint32_t* nbr, const int32_t* count, const int32_t* order,
                      const int64_t* row_ptr, const int32_t* sys_start, const int32_t* atom_sys,
                      int64_t n_atoms, int32_t jbits, int32_t tpa, int64_t ld, int64_t* mapping,
                      int64_t n_pairs, int64_t* system_mapping, double* shifts, void* stream) {
  if (n_atoms == 0 || n_pairs == 0) return TSX_OK;
  hipStream_t st = (hipStream_t)stream;
  k_emit<<<blocks_for(n_atoms, 256), 256, 0, st>>>(nbr, count, order, row_ptr, sys_start,
                                                   atom_sys, n_atoms, jbits, tpa, ld, mapping,
                                                   n_pairs, system_mapping, shifts);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

int64_t tsx_wrap_ws_bytes(int32_t n_systems) {
  return (int64_t)ws_layout(0, n_systems, 1, nullptr, nullptr);
}

int tsx_wrap_positions(const double* pos, const double* cell_col, const uint8_t* pbc,
                       int32_t pbc_per_system, const int32_t* atom_sys, int64_t n_atoms,
                       int32_t n_systems, void* ws, int32_t* status, double* wrapped,
                       double* shifts, void* stream) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  hipStream_t st = (hipStream_t)stream;
  NLWs w;
  ws_layout(0, n_systems, 1, &w, (char*)ws);
  // geometry only: no bin counters are touched (the workspace holds 2)
  k_setup<<<1, 1024, 0, st>>>(cell_col, pbc, pbc_per_system, n_systems, 1.0, 1, 1, 0, w.geom,
                              w.grid, w.meta, status, w.bin_count, nullptr, nullptr);
  TSX_CHECK_LAUNCH();
  if (n_atoms == 0) return TSX_OK;
  k_wrap_out<<<blocks_for(n_atoms, 256), 256, 0, st>>>(pos, w.geom, w.grid, atom_sys, n_atoms,
                                                       wrapped, shifts);
  TSX_CHECK_LAUNCH();
  return TSX_OK;
}

}  // extern "C"
