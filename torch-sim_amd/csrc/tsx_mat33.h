// 3x3 FP64 matrix functions for the per-system cell-filter updates.
//
// Row-major double[9].  Everything is __host__ __device__ and single-thread:
// one thread handles one system's 3x3 algebra.  The same header is compiled
// as plain C++ by the CPU test harness, which checks these routines against
// scipy.linalg.
//
//   m33_expm           exp(A): scaling and squaring with the degree-13 Pade
//                      approximant (Higham, SIAM J. Matrix Anal. Appl. 26,
//                      2005, Algorithm 2.3 coefficients).
//   m33_expm_frechet   (exp(A), L(A, E)): the same Pade / squaring recurrence
//                      evaluated in the algebra of block upper-triangular
//                      matrices [[X, Y], [0, X]], whose (1,2) block of
//                      exp([[A, E], [0, A]]) is the Frechet derivative L(A, E)
//                      (Najfeld and Havel, Adv. Appl. Math. 16, 1995).
//   m33_logm           principal log: inverse scaling and squaring -- repeated
//                      Denman-Beavers square roots until ||A - I||_1 <= 1/4,
//                      then log(I + X) by 8-point Gauss-Legendre quadrature of
//                      int_0^1 X (I + tX)^-1 dt (the [8/8] Pade approximant,
//                      Higham, SIAM J. Matrix Anal. Appl. 22, 2001).
#pragma once

#include <math.h>

namespace tsx {

struct M33 {
  double a[9];
};

__host__ __device__ inline M33 m33_zero() {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = 0.0;
  return r;
}
__host__ __device__ inline M33 m33_eye(double s = 1.0) {
  M33 r = m33_zero();
  r.a[0] = r.a[4] = r.a[8] = s;
  return r;
}
__host__ __device__ inline M33 m33_load(const double* p) {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = p[k];
  return r;
}
__host__ __device__ inline void m33_store(const M33& m, double* p) {
  for (int k = 0; k < 9; ++k) p[k] = m.a[k];
}
__host__ __device__ inline M33 m33_mul(const M33& x, const M33& y) {
  M33 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = x.a[3 * i] * y.a[j] + x.a[3 * i + 1] * y.a[3 + j] + x.a[3 * i + 2] * y.a[6 + j];
  return r;
}
__host__ __device__ inline M33 m33_add(const M33& x, const M33& y) {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = x.a[k] + y.a[k];
  return r;
}
__host__ __device__ inline M33 m33_sub(const M33& x, const M33& y) {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = x.a[k] - y.a[k];
  return r;
}
__host__ __device__ inline M33 m33_scale(const M33& x, double s) {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = x.a[k] * s;
  return r;
}
// s1 * x + s2 * y
__host__ __device__ inline M33 m33_axpby(double s1, const M33& x, double s2, const M33& y) {
  M33 r;
  for (int k = 0; k < 9; ++k) r.a[k] = s1 * x.a[k] + s2 * y.a[k];
  return r;
}
__host__ __device__ inline M33 m33_t(const M33& x) {
  M33 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = x.a[3 * j + i];
  return r;
}
__host__ __device__ inline double m33_det(const M33& m) {
  const double* a = m.a;
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// Inverse by cofactors.
__host__ __device__ inline M33 m33_inv(const M33& m) {
  const double* a = m.a;
  M33 r;
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double inv_d = 1.0 / (a[0] * c00 + a[1] * c01 + a[2] * c02);
  r.a[0] = c00 * inv_d;
  r.a[1] = (a[2] * a[7] - a[1] * a[8]) * inv_d;
  r.a[2] = (a[1] * a[5] - a[2] * a[4]) * inv_d;
  r.a[3] = c01 * inv_d;
  r.a[4] = (a[0] * a[8] - a[2] * a[6]) * inv_d;
  r.a[5] = (a[2] * a[3] - a[0] * a[5]) * inv_d;
  r.a[6] = c02 * inv_d;
  r.a[7] = (a[1] * a[6] - a[0] * a[7]) * inv_d;
  r.a[8] = (a[0] * a[4] - a[1] * a[3]) * inv_d;
  return r;
}
// Max column sum.
__host__ __device__ inline double m33_norm1(const M33& m) {
  double best = 0.0;
  for (int j = 0; j < 3; ++j) {
    const double s = fabs(m.a[j]) + fabs(m.a[3 + j]) + fabs(m.a[6 + j]);
    best = s > best ? s : best;
  }
  return best;
}
__host__ __device__ inline double m33_frob(const M33& m) {
  double s = 0.0;
  for (int k = 0; k < 9; ++k) s += m.a[k] * m.a[k];
  return sqrt(s);
}
__host__ __device__ inline double m33_dot(const M33& x, const M33& y) {
  double s = 0.0;
  for (int k = 0; k < 9; ++k) s += x.a[k] * y.a[k];
  return s;
}
__host__ __device__ inline void m33_apply(const M33& m, const double* v, double* o) {
  const double x = v[0], y = v[1], z = v[2];
  o[0] = m.a[0] * x + m.a[1] * y + m.a[2] * z;
  o[1] = m.a[3] * x + m.a[4] * y + m.a[5] * z;
  o[2] = m.a[6] * x + m.a[7] * y + m.a[8] * z;
}

// Block upper-triangular pair [[x, y], [0, x]].
struct Dual33 {
  M33 x, y;
};
__host__ __device__ inline Dual33 d33_mul(const Dual33& p, const Dual33& q) {
  return Dual33{m33_mul(p.x, q.x), m33_add(m33_mul(p.x, q.y), m33_mul(p.y, q.x))};
}
__host__ __device__ inline Dual33 d33_axpby(double s1, const Dual33& p, double s2, const Dual33& q) {
  return Dual33{m33_axpby(s1, p.x, s2, q.x), m33_axpby(s1, p.y, s2, q.y)};
}
// q^-1 p for block upper-triangular q (q.x invertible).
__host__ __device__ inline Dual33 d33_solve(const Dual33& q, const Dual33& p) {
  const M33 qi = m33_inv(q.x);
  const M33 rx = m33_mul(qi, p.x);
  const M33 ry = m33_mul(qi, m33_sub(p.y, m33_mul(q.y, rx)));
  return Dual33{rx, ry};
}

// Degree-13 Pade approximant of exp on the dual algebra, for ||a.x||_1 <= theta_13.
__host__ __device__ inline Dual33 d33_pade13(const Dual33& a) {
  // Pade coefficients b_k / b_0 (b_0 = 64764752532480000), so that the
  // approximant of exp(0) is exactly I
  const double b0 = 64764752532480000.0;
  const double b[14] = {1.0,
                        32382376266240000.0 / b0,
                        7771770303897600.0 / b0,
                        1187353796428800.0 / b0,
                        129060195264000.0 / b0,
                        10559470521600.0 / b0,
                        670442572800.0 / b0,
                        33522128640.0 / b0,
                        1323241920.0 / b0,
                        40840800.0 / b0,
                        960960.0 / b0,
                        16380.0 / b0,
                        182.0 / b0,
                        1.0 / b0};
  const Dual33 a2 = d33_mul(a, a);
  const Dual33 a4 = d33_mul(a2, a2);
  const Dual33 a6 = d33_mul(a4, a2);
  const Dual33 eye{m33_eye(), m33_zero()};
  // u = a [a6 (b13 a6 + b11 a4 + b9 a2) + b7 a6 + b5 a4 + b3 a2 + b1 I]
  Dual33 t = d33_axpby(b[13], a6, b[11], a4);
  t = d33_axpby(1.0, t, b[9], a2);
  t = d33_mul(a6, t);
  t = d33_axpby(1.0, t, b[7], a6);
  t = d33_axpby(1.0, t, b[5], a4);
  t = d33_axpby(1.0, t, b[3], a2);
  t = d33_axpby(1.0, t, b[1], eye);
  const Dual33 u = d33_mul(a, t);
  // v = a6 (b12 a6 + b10 a4 + b8 a2) + b6 a6 + b4 a4 + b2 a2 + b0 I
  Dual33 v = d33_axpby(b[12], a6, b[10], a4);
  v = d33_axpby(1.0, v, b[8], a2);
  v = d33_mul(a6, v);
  v = d33_axpby(1.0, v, b[6], a6);
  v = d33_axpby(1.0, v, b[4], a4);
  v = d33_axpby(1.0, v, b[2], a2);
  v = d33_axpby(1.0, v, b[0], eye);
  return d33_solve(d33_axpby(1.0, v, -1.0, u), d33_axpby(1.0, v, 1.0, u));
}

__host__ __device__ inline int expm_squarings(double norm1) {
  const double theta13 = 5.371920351148152;
  if (!(norm1 > theta13)) return 0;
  const int s = (int)ceil(log2(norm1 / theta13));
  return s < 0 ? 0 : (s > 60 ? 60 : s);
}

// (exp(A), L(A, E)).
__host__ __device__ inline Dual33 m33_expm_frechet(const M33& a, const M33& e) {
  const int s = expm_squarings(m33_norm1(a));
  const double sc = ldexp(1.0, -s);
  Dual33 r = d33_pade13(Dual33{m33_scale(a, sc), m33_scale(e, sc)});
  for (int k = 0; k < s; ++k) r = d33_mul(r, r);
  return r;
}

__host__ __device__ inline M33 m33_expm(const M33& a) {
  return m33_expm_frechet(a, m33_zero()).x;
}

// Principal square root by the scaled product Denman-Beavers iteration
// (Higham, "Functions of Matrices", 2008, eq. 6.17 with determinantal scaling).
//   M_0 = Y_0 = A,  g_k = |det M_k|^(-1/6) while M_k is far from I, else 1,
//   Y_{k+1} = (g_k / 2) Y_k (I + g_k^-2 M_k^-1),
//   M_{k+1} = (I + (g_k^2 M_k + g_k^-2 M_k^-1) / 2) / 2   ->  I.
__host__ __device__ inline M33 m33_sqrtm(const M33& a) {
  M33 m = a, y = a;
  for (int it = 0; it < 60; ++it) {
    const double dist = m33_frob(m33_sub(m, m33_eye()));
    if (dist <= 1e-15) break;
    double g = 1.0;
    if (dist > 1e-2) {
      const double dm = fabs(m33_det(m));
      if (dm > 0.0) g = pow(dm, -1.0 / 6.0);
    }
    const double g2 = g * g;
    const M33 mi = m33_inv(m);
    y = m33_mul(m33_scale(y, 0.5 * g), m33_axpby(1.0, m33_eye(), 1.0 / g2, mi));
    m = m33_axpby(0.5, m33_eye(), 0.25, m33_axpby(g2, m, 1.0 / g2, mi));
  }
  return y;
}

__host__ __device__ inline M33 m33_logm(const M33& a_in) {
  // Gauss-Legendre nodes and weights on [-1, 1] (8 points)
  const double xg[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267,
                        0.9602898564975363};
  const double wg[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745,
                        0.1012285362903763};
  M33 a = a_in;
  int k = 0;
  while (m33_norm1(m33_sub(a, m33_eye())) > 0.25 && k < 40) {
    a = m33_sqrtm(a);
    ++k;
  }
  const M33 x = m33_sub(a, m33_eye());
  M33 acc = m33_zero();
  for (int q = 0; q < 8; ++q) {
    const double t = 0.5 * (1.0 + (q < 4 ? -xg[3 - q] : xg[q - 4]));
    const double w = 0.5 * (q < 4 ? wg[3 - q] : wg[q - 4]);
    const M33 r = m33_mul(x, m33_inv(m33_add(m33_eye(), m33_scale(x, t))));
    acc = m33_axpby(1.0, acc, w, r);
  }
  return m33_scale(acc, ldexp(1.0, k));
}

}  // namespace tsx
