// Host build of the device 3x3 matrix functions, for CPU tests against scipy.
#define __host__
#define __device__
#include "tsx_mat33.h"

using namespace tsx;

extern "C" {
void c_expm(const double* a, double* o) { m33_store(m33_expm(m33_load(a)), o); }
void c_logm(const double* a, double* o) { m33_store(m33_logm(m33_load(a)), o); }
void c_sqrtm(const double* a, double* o) { m33_store(m33_sqrtm(m33_load(a)), o); }
void c_expm_frechet(const double* a, const double* e, double* ea, double* l) {
  const Dual33 r = m33_expm_frechet(m33_load(a), m33_load(e));
  m33_store(r.x, ea);
  m33_store(r.y, l);
}
}
