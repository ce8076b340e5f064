// hes_adjoint_common.hpp -- device pieces shared by the adjoint of one
// single-phase Rusanov run (hes_adjoint.hip) and of a batch of them
// (hes_badjoint.hip): deterministic reductions, the CFL argmax record, the
// gradient of the sound speed, and the reverse of one chunk step.
#pragma once
#include <hip/hip_runtime.h>

#include "hes_adjoint_gen.hpp"
#include "hes_step.hpp"

// A batch and a single problem must evolve bit for bit alike, so both
// adjoint sources take the same eigenvalue iteration.
static_assert(HES_EIG_SECULAR == 1, "the adjoint's sources define HES_EIG_SECULAR 1");

namespace hes_adj {

using namespace hes;

constexpr int NQ = 14;
constexpr int NP = 13;     // differentiable primitives of a state: u(3), A(9), S
constexpr int JL = 16;     // dlam/dQ (14), dlam/db0, dlam/dc0

__device__ __forceinline__ double wave_sum(double v) {
  // xor butterfly: every lane ends with the same bits (each level adds the
  // same two operands on both partners)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block sum of v over NT threads (thread 0 gets the result).
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[NT / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int k = 1; k < NT / 64; ++k) s += red[k];
  return s;
}

// (value, index) with the larger value, ties to the smaller index.
__device__ __forceinline__ void arg_better(double& v, long long& i, double ov, long long oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

// Block argmax of (v, idx) over NT threads; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_argmax(double& v, long long& idx) {
  __shared__ double rv[NT / 64];
  __shared__ long long ri[NT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o);
    long long oi = __shfl_xor(idx, o);
    arg_better(v, idx, ov, oi);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    rv[w] = v;
    ri[w] = idx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) arg_better(v, idx, rv[k], ri[k]);
  }
}

constexpr long long NO_CELL = 0x7fffffffffffffffLL;

// Max CFL speed of a state (into lmax, as double bits) and, optionally, the
// per-block argmax (value, smallest index) of the owned cells' speeds.
template <int NT>
__device__ __forceinline__ void record_speeds(double lam, bool owned, long long i,
                                              unsigned long long* lmax, double* bv,
                                              long long* bi) {
  double v = owned ? lam : -1.0;
  long long idx = owned ? i : NO_CELL;
  block_argmax<NT>(v, idx);
  if (threadIdx.x == 0) {
    if (lmax && v > 0.0) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
      // read first: most blocks find a larger maximum already there (all
      // blocks of a step would otherwise contend for one address)
      if (bits > __hip_atomic_load(lmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(lmax, bits);
    }
    if (bv) {
      bv[blockIdx.x] = v;
      bi[blockIdx.x] = idx;
    }
  }
}

__device__ __forceinline__ double lmax_of(const unsigned long long* h, long long s) {
  return __longlong_as_double((long long)h[s]);
}

// Sound speed c = sqrt(rho(Omega)) of a symmetric acoustic tensor (O11 O12 O13
// O22 O23 O33) and dc/dO (off-diagonal entries counted twice, as each stands
// for two matrix entries).  rho(Omega) is the eigenvalue of largest modulus,
// found by sym3_extreme as the forward step finds it; its eigenvector is the largest cross
// product of two rows of Omega - mu I.
__device__ __forceinline__ double speed_and_grad(const double O[6], double g[6]) {
  const double a = O[0], b = O[3], c = O[5], d = O[1], e = O[2], f = O[4];
  const double mu = sym3_extreme(O);
  const double sgn = mu < 0.0 ? -1.0 : 1.0;
  const double rad = fabs(mu);
  const double A = a - mu, B = b - mu, C = c - mu;
  // rows (A d e), (d B f), (e f C)
  double x1 = d * f - e * B, y1 = e * d - A * f, z1 = A * B - d * d;
  double x2 = d * C - e * f, y2 = e * e - A * C, z2 = A * f - d * e;
  double x3 = B * C - f * f, y3 = f * e - d * C, z3 = d * f - B * e;
  double n1 = x1 * x1 + y1 * y1 + z1 * z1;
  double n2 = x2 * x2 + y2 * y2 + z2 * z2;
  double n3 = x3 * x3 + y3 * y3 + z3 * z3;
  double vx = x1, vy = y1, vz = z1, nn = n1;
  if (n2 > nn) { vx = x2; vy = y2; vz = z2; nn = n2; }
  if (n3 > nn) { vx = x3; vy = y3; vz = z3; nn = n3; }
  const double cs = rad > 0.0 ? sqrt_pos(rad) : 0.0;  // as sound_speed forms it
  // dc/dmu = sgn / (2c), dmu/dO = v v^T / |v|^2
  const double k = nn > 0.0 ? sgn * 0.5 * rcp(cs) * rcp(nn) : 0.0;
  g[0] = k * vx * vx;
  g[1] = 2.0 * k * vx * vy;
  g[2] = 2.0 * k * vx * vz;
  g[3] = k * vy * vy;
  g[4] = 2.0 * k * vy * vz;
  g[5] = k * vz * vz;
  return cs;
}


// (value, index) argmax over one wave, ties to the smaller index; every lane
// ends with the result.
__device__ __forceinline__ void wave_argmax(double& v, long long& idx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o);
    long long oi = __shfl_xor(idx, o);
    arg_better(v, idx, ov, oi);
  }
}

// Reverse of the single-phase Rusanov chunk step (chunk_step: lane l holds
// cell i = first + l - 1 of one problem, lanes 0 and 63 are halo cells).
//   q        state s on every lane (halo lanes outside the problem hold a
//            clamped copy; their cotangent is zero)
//   qb       in: Qbar of state s+1 with its lmax term; out: the face part of
//            Qbar of state s
//   interior / owned  as chunk_step defines them
//   c        the stored sound speed of the lane's cell in state s (as the
//            forward step read it: the face speeds below are its own)
//   hdt      dtdx_s / 2
//   gin      (b0, c0, rho0, cv, t0)
//   ib       out: the cell-VJP part (ib[0..13] of Qbar_s, ib[14], ib[15] the
//            lane's d/d(b0, c0))
//   dts      out: the lane's share of dtdxbar_s (owned lanes)
// With W_i = Qbar_i of state s+1 on interior cells and g_f = W_right(f) -
// W_left(f): Qbar_i += dtdx/2 (s_R g_R - s_L g_L), Fbar_i = dtdx/2 (g_L + g_R),
// and the face speed s_f = max(lam_i, lam_i+1) passes its cotangent
// -dtdx/2 g_f . (Q_i+1 - Q_i) to the larger side (the left one on ties).
__device__ __forceinline__ void reverse_chunk(const double* q, double* qb, bool interior,
                                              bool owned, double c, double hdt, const double* gin,
                                              double* ib, double& dts) {
  double in[hes_gen::CELL_IN];
#pragma unroll
  for (int r = 0; r < NQ; ++r) in[r] = q[r];
#pragma unroll
  for (int r = 0; r < 5; ++r) in[NQ + r] = gin[r];
  double out[hes_gen::CELL_OUT];
#pragma unroll
  for (int r = 0; r < hes_gen::CELL_IN; ++r) ib[r] = 0.0;
  // seeds of the cell VJP from the primal: first those of the flux F and u1;
  // the acoustic tensor is formed, and its seeds set, only in waves where a
  // face speed has a cotangent (lbar != 0 on some lane)
  double lbar = 0.0;
  auto seed = [&](const double* o, double* ob) -> bool {
    const double u1 = o[20];
    const double lam = fabs(u1) + c;
    const double lam_r = lane_down(lam), lam_l = lane_up(lam);
    const double sR = fmax(lam, lam_r), sL = fmax(lam_l, lam);
    double sdR = 0.0, sdL = 0.0, fdot = 0.0;
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      const double w = interior ? qb[r] : 0.0;
      const double wr = lane_down(w), wl = lane_up(w);
      const double qr = lane_down(q[r]), ql = lane_up(q[r]);
      const double gR = wr - w, gL = w - wl;
      const double gs = gL + gR;
      ob[r] = hdt * gs;
      fdot += o[r] * gs;
      sdR += gR * (qr - q[r]);
      sdL += gL * (q[r] - ql);
      qb[r] += hdt * (sR * gR - sL * gL);
    }
    lbar = (lam >= lam_r ? -hdt * sdR : 0.0) + (lam > lam_l ? -hdt * sdL : 0.0);
    ob[20] = u1 < 0.0 ? -lbar : lbar;
    dts = owned ? 0.5 * fdot - 0.5 * sR * sdR : 0.0;
    return __any(lbar != 0.0);
  };
  auto seed_acoustic = [&](const double* o, double* ob) {
    // d lam / d Omega = d c / d Omega (the top eigenvector's outer product)
    double gO[6];
    (void)speed_and_grad(o + NQ, gO);
#pragma unroll
    for (int k = 0; k < 6; ++k) ob[NQ + k] = lbar * gO[k];
  };
  hes_gen::cell_fwd_vjp(in, out, seed, seed_acoustic, ib);
}

}  // namespace hes_adj
