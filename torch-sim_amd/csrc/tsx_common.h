// Shared device/host helpers for the tsx HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>

#include "tsx.h"

namespace tsx {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);

#define TSX_HIP(call)                                                      \
  do {                                                                     \
    hipError_t _e = (call);                                                \
    if (_e != hipSuccess) {                                                \
      ::tsx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                       __FILE__, __LINE__);                                \
      return TSX_ERR_HIP;                                                  \
    }                                                                      \
  } while (0)

#define TSX_CHECK_LAUNCH() TSX_HIP(hipGetLastError())

#define TSX_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      ::tsx::set_error(__VA_ARGS__);  \
      return TSX_ERR_ARG;             \
    }                                 \
  } while (0)

inline int blocks_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  return (int)(b < 1 ? 1 : b);
}

// XCD-aware block order: consecutive hardware blocks are dealt round-robin
// to the 8 XCDs of the chip (each with its own L2).  Renumbering so that
// every XCD processes one contiguous range of rows keeps the neighbour
// positions a block gathers in that XCD's L2.
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x, b = blockIdx.x;
  const int q = G >> 3, r = G & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// ------------------------------------------------------ neighbour words
// One 32-bit word per list entry: the neighbour's index within its system in
// the low jbits, then three sbits-wide two's-complement image-shift fields
// (S'_x, S'_y, S'_z), sbits = (32 - jbits) / 3.  A zero shift is a zero high
// part; each field decodes with one signed bitfield extract.
__device__ __forceinline__ uint32_t nl_word(int j_local, int sx, int sy, int sz, int jbits,
                                            int sbits) {
  const uint32_t m = (1u << sbits) - 1u;
  return (uint32_t)j_local | ((((uint32_t)sx & m) | (((uint32_t)sy & m) << sbits) |
                               (((uint32_t)sz & m) << (2 * sbits)))
                              << jbits);
}
__device__ __forceinline__ int nl_shift(uint32_t w, int jbits, int sbits, int k) {
  return __builtin_amdgcn_sbfe((int)w, (unsigned)(jbits + k * sbits), (unsigned)sbits);
}

// ------------------------------------------------------ wave64 reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// inclusive scan of one int per lane over the wave
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-wide sum of a double (blockDim.x a multiple of 64, <= 1024).
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < nw; ++k) t += sh[k];  // fixed order: deterministic
  __syncthreads();
  return t;
}

// ------------------------------------------------------------ 3x3 (FP64)
// Row-major 3x3 helpers.  m[3*r + c].
__device__ __host__ __forceinline__ double det33(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Inverse via cofactors; returns the determinant.
__device__ __host__ __forceinline__ double inv33(const double* m, double* o) {
  const double c00 = m[4] * m[8] - m[5] * m[7];
  const double c01 = m[5] * m[6] - m[3] * m[8];
  const double c02 = m[3] * m[7] - m[4] * m[6];
  const double d = m[0] * c00 + m[1] * c01 + m[2] * c02;
  const double r = 1.0 / d;
  o[0] = c00 * r;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * r;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * r;
  o[3] = c01 * r;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * r;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * r;
  o[6] = c02 * r;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * r;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * r;
  return d;
}

__device__ __host__ __forceinline__ void mm33(const double* a, const double* b, double* o) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
  for (int k = 0; k < 9; ++k) o[k] = t[k];
}

__device__ __host__ __forceinline__ void tr33(const double* a, double* o) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[3 * r + c] = a[3 * c + r];
  for (int k = 0; k < 9; ++k) o[k] = t[k];
}

// ----------------------------------------------------- Philox4x32-10 RNG
// Counter-based generator (Salmon et al., SC'11): one 128-bit block per
// (key, counter).  Used for the Langevin noise so that the stream depends
// only on (seed, step, original atom id), not on thread order.
struct u4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Uniform double in (0, 1] from two 32-bit words (53 random bits).
__device__ __forceinline__ double u01_53(uint32_t a, uint32_t b) {
  const uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  return ((double)(v & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
}

// Three standard normals for (seed, step, atom) via Box-Muller in FP64.
__device__ __forceinline__ void normal3(uint64_t seed, uint64_t step, uint32_t atom,
                                        double* z) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  u4 r0 = philox4x32_10(u4{atom, (uint32_t)step, (uint32_t)(step >> 32), 0u}, k0, k1);
  u4 r1 = philox4x32_10(u4{atom, (uint32_t)step, (uint32_t)(step >> 32), 1u}, k0, k1);
  const double u1 = u01_53(r0.x, r0.y), u2 = u01_53(r0.z, r0.w);
  const double u3 = u01_53(r1.x, r1.y), u4v = u01_53(r1.z, r1.w);
  const double a = sqrt(-2.0 * log(u1));
  const double b = sqrt(-2.0 * log(u3));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  z[0] = a * c;
  z[1] = a * s;
  double s2, c2;
  sincospi(2.0 * u4v, &s2, &c2);
  z[2] = b * c2;
}

}  // namespace tsx
