// hes_host.hpp -- host-side helpers shared by the libhes translation units:
// error reporting, HIP error checks and material constants.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>

#include "hes_device.hpp"
#include "../../include/hes.h"

namespace hes {

inline thread_local std::string g_err;

inline int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e__ = (expr);                                                       \
    if (e__ != hipSuccess) {                                                       \
      ::hes::g_err = std::string(#expr) + ": " + hipGetErrorString(e__);           \
      return HES_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

// Barton2009 constants (rho0, c0, cv, t0, b0, alpha, beta, gamma) -> MatDev
// with the derived quantities (EquationsOfState.jl:90-115).
inline MatDev make_mat(const double* p) {
  MatDev m;
  std::memset(&m, 0, sizeof(m));
  m.rho0 = p[0];
  m.irho0 = 1.0 / p[0];
  m.cv = p[2];
  m.t0 = p[3];
  double b0 = p[4];
  m.alpha = p[5];
  m.beta = p[6];
  m.gamma = p[7];
  m.b0sq = b0 * b0;  // EquationsOfState.jl:111-112
  m.k0 = p[1] * p[1] - (4.0 / 3.0) * (b0 * b0);
  m.hk0a2 = 0.5 * m.k0 / (m.alpha * m.alpha);
  m.k0_2a = m.k0 / (2.0 * m.alpha);
  m.k0_a = m.k0 / m.alpha;
  m.cvt0 = m.cv * m.t0;
  m.icvt0 = 1.0 / m.cvt0;
  m.hb0sq = 0.5 * m.b0sq;
  m.half_a = 0.5 * m.alpha;
  m.half_b = 0.5 * m.beta;
  m.half_g = 0.5 * m.gamma;
  m.ac_t = 3.0 + 2.0 * m.beta;
  m.ac_s = 2.0 + m.beta;
  m.ac_e = m.gamma * (1.0 + m.gamma);
  m.ac_w = m.beta * (1.0 + m.beta);
  m.fast = (m.alpha == 1.0 && m.beta == 3.0 && m.gamma == 2.0) ? 1 : 0;
  return m;
}

inline bool mat_ok(const double* p) {
  for (int k = 0; k < 8; ++k)
    if (!std::isfinite(p[k])) return false;
  return p[0] > 0 && p[2] > 0 && p[3] > 0 && p[5] != 0;
}

// Copies and fills that belong to a handle run on the handle's stream and are
// waited for: a plain hipMemcpy / hipMemset goes to the null stream, which is
// not ordered with the handle's non-blocking stream.
inline hipError_t copy_on(hipStream_t s, void* dst, const void* src, size_t bytes,
                          hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

inline hipError_t zero_on(hipStream_t s, void* dst, size_t bytes) {
  hipError_t e = hipMemsetAsync(dst, 0, bytes, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

}  // namespace hes
