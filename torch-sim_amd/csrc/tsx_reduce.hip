// Per-system reductions over concatenated systems of any size.
//
// Each system is cut into chunks of RED_CH atoms counted from its own first
// atom, one workgroup per (system, chunk): a single 1M-atom system and 512
// small ones both fill the chip, and a system's summation order depends only
// on its own size (a system sums bitwise alike alone or inside any batch).
// When every system fits one chunk, the workgroup owning a system writes its
// results directly; otherwise a second kernel adds each system's chunk
// partials in chunk order (the kernel boundary orders the partial writes; a
// device-scope fence per workgroup would write back the L2).
//
// Reference: the per-system index_add sums of pair_potential.py:351-464
// (energy, stress) and md.py / quantities.py segment sums.
#include "tsx_common.h"

namespace tsx {

constexpr int RED_CH = 2048;
constexpr int RED_THREADS = 256;
constexpr int RED_WAVES = RED_THREADS / 64;
constexpr int RED_MAXV = 7;

// per-atom values: e (V = 1) or e and the 6 virial components (V = 7)
template <int V>
__device__ __forceinline__ void load_vals(const double* e_atom, const double* w_atom, int64_t k,
                                          double* v) {
  v[0] = e_atom[k];
  if (V == 7)
    for (int c = 0; c < 6; ++c) v[1 + c] = w_atom[6 * k + c];
}

struct SegOut {
  double* energy;
  double* stress;
  const double* cell_col;
  int32_t* gate_clear;
  int64_t* step;
  int advance;
};

template <int V>
__device__ void finish_system(const SegOut& o, int s, const double* v) {
  o.energy[s] = v[0];
  if (V == 7 && o.stress) {
    const double* C = o.cell_col + 9 * (int64_t)s;
    const double vol = fabs(det33(C));
    double* st = o.stress + 9 * (int64_t)s;
    st[0] = -v[1] / vol;
    st[4] = -v[2] / vol;
    st[8] = -v[3] / vol;
    st[5] = st[7] = -v[4] / vol;
    st[2] = st[6] = -v[5] / vol;
    st[1] = st[3] = -v[6] / vol;
  }
  if (o.gate_clear) o.gate_clear[s] = 0;
}

template <int V>
__global__ __launch_bounds__(RED_THREADS) void k_seg_reduce(
    const double* __restrict__ e_atom, const double* __restrict__ w_atom,
    const int32_t* __restrict__ sys_start, int chunks, int n_sys, double* __restrict__ part,
    SegOut o, double* __restrict__ stresses) {
  __shared__ double sh[RED_WAVES][V];
  const int s = blockIdx.x / chunks, c = blockIdx.x % chunks;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t lo = (int64_t)sys_start[s] + (int64_t)c * RED_CH;
  const int64_t end = sys_start[s + 1];
  const int64_t hi = lo + RED_CH < end ? lo + RED_CH : end;
  double acc[V];
  for (int q = 0; q < V; ++q) acc[q] = 0.0;
  if (V == 1 && lo < hi) {
    // the (at most RED_CH / RED_THREADS) loads of a thread are issued
    // together; the adds keep the strided order of the loop below
    constexpr int PT = RED_CH / RED_THREADS;
    double v[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int64_t k = lo + threadIdx.x + (int64_t)u * RED_THREADS;
      v[u] = e_atom[k < hi ? k : lo];
    }
#pragma unroll
    for (int u = 0; u < PT; ++u)
      if (lo + threadIdx.x + (int64_t)u * RED_THREADS < hi) acc[0] += v[u];
  } else {
    for (int64_t k = lo + threadIdx.x; k < hi; k += RED_THREADS) {
      double v[V];
      load_vals<V>(e_atom, w_atom, k, v);
      for (int q = 0; q < V; ++q) acc[q] += v[q];
    }
  }
  for (int q = 0; q < V; ++q) {
    const double t = wave_sum(acc[q]);
    if (lane == 0) sh[wid][q] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[V];
    for (int q = 0; q < V; ++q) {
      double t = 0.0;
      for (int w = 0; w < RED_WAVES; ++w) t += sh[w][q];
      v[q] = t;
    }
    if (chunks == 1) {
      finish_system<V>(o, s, v);
      if (o.advance && o.step && s == 0) o.step[0] += 1;
    } else {
      for (int q = 0; q < V; ++q) part[((int64_t)s * chunks + c) * V + q] = v[q];
    }
  }
  if (V == 7 && stresses && lo < hi) {
    const double iv = 1.0 / fabs(det33(o.cell_col + 9 * (int64_t)s));
    for (int64_t k = lo + threadIdx.x; k < hi; k += RED_THREADS) {
      const double* wk = w_atom + 6 * k;
      double* st = stresses + 9 * k;
      st[0] = -wk[0] * iv;
      st[4] = -wk[1] * iv;
      st[8] = -wk[2] * iv;
      st[5] = st[7] = -wk[3] * iv;
      st[2] = st[6] = -wk[4] * iv;
      st[1] = st[3] = -wk[5] * iv;
    }
  }
}

// systems larger than one chunk: one workgroup per system adds its chunk
// partials (strided per thread, then the fixed-order block tree)
template <int V>
__global__ __launch_bounds__(RED_THREADS) void k_seg_finish(const int32_t* __restrict__ sys_start,
                                                            int chunks, int n_sys,
                                                            const double* __restrict__ part,
                                                            SegOut o) {
  __shared__ double sh[RED_WAVES][V];
  const int s = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n_s = (int64_t)sys_start[s + 1] - sys_start[s];
  int nc = (int)((n_s + RED_CH - 1) / RED_CH);
  nc = nc < chunks ? nc : chunks;  // a too small max_system_atoms truncates, never overreads
  double acc[V];
  for (int k = 0; k < V; ++k) acc[k] = 0.0;
  for (int c = threadIdx.x; c < nc; c += RED_THREADS)
    for (int k = 0; k < V; ++k) acc[k] += part[((int64_t)s * chunks + c) * V + k];
  for (int k = 0; k < V; ++k) {
    const double t = wave_sum(acc[k]);
    if (lane == 0) sh[wid][k] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[V];
    for (int k = 0; k < V; ++k) {
      double t = 0.0;
      for (int w = 0; w < RED_WAVES; ++w) t += sh[w][k];
      v[k] = t;
    }
    finish_system<V>(o, s, v);
    if (o.advance && o.step && s == 0) o.step[0] += 1;
  }
}

}  // namespace tsx

using namespace tsx;

extern "C" {

static int chunks_for(int32_t max_system_atoms) {
  return max_system_atoms > 0 ? (max_system_atoms + RED_CH - 1) / RED_CH : 1;
}

int64_t tsx_reduce_ws_bytes(int32_t max_system_atoms, int32_t n_systems) {
  return 64 + (int64_t)sizeof(double) * RED_MAXV * chunks_for(max_system_atoms) *
                  (n_systems > 0 ? n_systems : 1);
}

static int seg_launch(int v, const double* e_atom, const double* w_atom, const double* cell_col,
                      const int32_t* sys_start, int32_t max_system_atoms, int32_t n_systems,
                      double* energy, double* stress, double* stresses, int64_t* step,
                      int32_t advance, int32_t* gate_clear, void* ws, hipStream_t st) {
  TSX_REQUIRE(n_systems > 0, "n_systems must be positive");
  TSX_REQUIRE(max_system_atoms >= 0, "max_system_atoms must be >= 0");
  TSX_REQUIRE(ws != nullptr, "workspace of tsx_reduce_ws_bytes required");
  TSX_REQUIRE(!advance || step, "step counter required to advance");
  double* part = (double*)((char*)ws + 64);
  SegOut o{energy, stress, cell_col, gate_clear, step, advance};
  const int chunks = chunks_for(max_system_atoms);
  const int64_t nblk = (int64_t)n_systems * chunks;
  TSX_REQUIRE(nblk < (1ll << 31), "too many reduction blocks");
  if (v == 1)
    k_seg_reduce<1><<<(int)nblk, RED_THREADS, 0, st>>>(e_atom, nullptr, sys_start, chunks,
                                                        n_systems, part, o, nullptr);
  else
    k_seg_reduce<7><<<(int)nblk, RED_THREADS, 0, st>>>(e_atom, w_atom, sys_start, chunks,
                                                        n_systems, part, o, stresses);
  TSX_CHECK_LAUNCH();
  if (chunks > 1) {
    if (v == 1)
      k_seg_finish<1><<<n_systems, RED_THREADS, 0, st>>>(sys_start, chunks, n_systems, part, o);
    else
      k_seg_finish<7><<<n_systems, RED_THREADS, 0, st>>>(sys_start, chunks, n_systems, part, o);
    TSX_CHECK_LAUNCH();
  }
  return TSX_OK;
}

int tsx_system_reduce(const double* e_atom, const double* w_atom, const double* cell_col,
                      const int32_t* sys_start, int32_t max_system_atoms, int32_t n_systems,
                      double* energy, double* stress, double* stresses, void* ws, void* stream) {
  TSX_REQUIRE(!(stress || stresses) || (w_atom && cell_col), "stress needs w_atom and cell");
  return seg_launch(w_atom ? 7 : 1, e_atom, w_atom, cell_col, sys_start, max_system_atoms,
                    n_systems, energy, stress, stresses, nullptr, 0, nullptr, ws,
                    (hipStream_t)stream);
}

int tsx_md_reduce(const double* e_atom, const int32_t* sys_start, int32_t max_system_atoms,
                  int32_t n_systems, double* energy, int64_t* step, int32_t advance,
                  int32_t* gate_clear, void* ws, void* stream) {
  return seg_launch(1, e_atom, nullptr, nullptr, sys_start, max_system_atoms, n_systems, energy,
                    nullptr, nullptr, step, advance, gate_clear, ws, (hipStream_t)stream);
}

}  // extern "C"
