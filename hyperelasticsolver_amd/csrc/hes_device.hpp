// hes_device.hpp -- gfx950 device math for the Barton-2009 hyperelastic
// finite-volume step (one phase block at a time).
//
// Everything here is FP64 and lives in registers.  The reference evaluates the
// stress and the acoustic tensor with nested ForwardDiff duals
// (EquationsOfState.jl:179-190, 223-246); this file uses closed forms of the
// same derivatives instead:
//
//   energy  e(S,G) = k0/(2a^2)(s_a-1)^2 + cv t0 s_g (x-1) + b0^2/2 s_b (I1^2/3 - I2)
//           with s_k = I3^(k/2), x = exp(S/cv)              (EquationsOfState.jl:118-137)
//   stress  sigma = -2 rho H,  H = kappa (G^2 - I1/3 G) + c3 I,
//           kappa = b0^2 s_b / 2,
//           c3 = k0/(2a) (s_a-1) s_a + (g/2) E_th + (b/2) W  (= I3 de/dI3)
//           (the product G * dedG of EquationsOfState.jl:183-188 for symmetric G)
//   acoustic Omega_ij = (1/rho) d sigma(n,i) / dA along dA = e_j (A^T n)^T, giving
//           Omega_ij = 2 n_j (H n)_i - 2 (n^T dH_j)_i with dG_j = -(G e_j n^T + n e_j^T G)
//           (EquationsOfState.jl:226-243 contracted with n)
//
// The entropy never has to be formed on the step path: S = cv log(x) and the
// stress and acoustic tensor need only x = exp(S/cv), which cons->prim yields
// directly as the clamped argument of the logarithm (EquationsOfState.jl:150-155).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hes {

// Material constants of Barton2009 (EquationsOfState.jl:71-116), with the
// derived quantities the kernels use precomputed on the host.
struct MatDev {
  double rho0, irho0, cv, t0, b0sq, k0, alpha, beta, gamma;
  double hk0a2;   // k0 / (2 alpha^2)      : cold-energy prefactor
  double k0_2a;   // k0 / (2 alpha)        : in c3
  double k0_a;    // k0 / alpha            : in O11 of the acoustic tensor
  double cvt0;    // cv * t0
  double icvt0;   // 1 / (cv t0)
  double hb0sq;   // b0^2 / 2
  double half_a, half_b, half_g;  // alpha/2, beta/2, gamma/2
  double ac_t, ac_s;  // 3 + 2 beta, 2 + beta  : acoustic tensor, shear-coupling terms
  double ac_e, ac_w;  // gamma (1 + gamma), beta (1 + beta) : O11's E_th and W terms
  int fast;       // alpha == 1, beta == 3, gamma == 2 (no pow needed)
  int pad;
};

// State of one phase derived from its conserved block.
struct Thermo {
  double rho;            // true density
  double den;            // frac * rho
  double u[3];           // velocity
  double E;              // specific total energy
  double A[9];           // distortion tensor, column-major
  double G11, G22, G33, G12, G13, G23;  // Finger tensor (A A^T)^-1
  double i1, i2, i3;     // invariants of G
  double sa, sb, sg;     // I3^(alpha/2), I3^(beta/2), I3^(gamma/2)
  double W;              // shear energy
  double Eth;            // thermal energy cv t0 s_g (x - 1)
  double x;              // exp(S/cv), clamped at 1e-6 as the entropy law does
  int bad;               // non-positive det(rho A) or non-finite state
};

// 1/x from the hardware reciprocal refined by two Newton steps (error within
// about an ulp; the step does not need IEEE-rounded division).
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// Square root of a positive normal x from the hardware reciprocal square root
// (about 2^-24): one Goldschmidt step for sqrt(x) and 1/(2 sqrt(x)), then one
// correction of the root.  Measured on gfx950 equal to the IEEE square root
// on 2^20 arguments over 13 decades; without the scaling that the IEEE
// sequence spends on denormal and infinite arguments (x = 0 gives NaN).
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double e = fma(-h, g, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  return fma(fma(-g, g, x), h, g);
}

// Cube root of a positive normal v whose single-precision value is normal
// (about 1e-37 to 1e38): an estimate from the hardware's single-precision
// log2 and exp2 (a few 2^-23 relative), then one Halley step
// y (y^3 + 2 v) / (2 y^3 + v), which cubes the relative error, so the
// result is within a few ulps.
__device__ __forceinline__ double cbrt_pos(double v) {
  const float e = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)v) * (1.0f / 3.0f));
  const double y = (double)e;
  const double y3 = y * y * y;
  return y * fma(2.0, v, y3) * rcp(fma(2.0, y3, v));
}

__device__ __forceinline__ double det3cm(const double* a) {
  // column-major a(i,j) = a[i + 3j]
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
         a[6] * (a[1] * a[5] - a[4] * a[2]);
}

// Density of a phase block q = [a r, a r u(3), a r E, a r A(9)]: rho =
// sqrt(det(a r A) / a^3 / rho0), as cons2prim forms it (the stored a r is not
// read), den = a rho and iden = 1 / den; bad: non-positive det.
template <bool ONE_FRAC>
__device__ __forceinline__ void phase_density(const MatDev& m, double frac, const double* q,
                                              double& rho, double& den, double& iden, int& bad) {
  const double detP = det3cm(q + 5);
  double rho2;
  if (ONE_FRAC) {
    rho2 = detP * m.irho0;
  } else {
    double ifr = rcp(frac);
    rho2 = detP * (ifr * ifr * ifr) * m.irho0;
  }
  bad = !(detP > 0.0);
  rho = sqrt_pos(rho2);  // as thermo forms it
  den = ONE_FRAC ? rho : frac * rho;
  iden = rcp(den);
}

// Normal velocity u1 of a phase block (the u[0] that thermo computes).
template <bool ONE_FRAC>
__device__ __forceinline__ double phase_u1(const MatDev& m, double frac, const double* q) {
  double rho, den, iden;
  int bad;
  phase_density<ONE_FRAC>(m, frac, q, rho, den, iden, bad);
  return q[1] * iden;
}

// Finger tensor, invariants and energy split of a phase.
// q points at [a r, a r u(3), a r E, a r A(9)] (the phase block without a).
// HyperelasticityMPh.jl:106-124 (cons2prim) + Strains.jl:26-52 + EquationsOfState.jl:139-156.
// FAST: 1 = alpha 1, beta 3, gamma 2 (no pow), 0 = general exponents,
// -1 = decide from m.fast at run time.
//
// G = (A A^T)^-1 is formed from P = a rho A without dividing A out:
// G = den^2 adj(P P^T) / det(P)^2 with den = a rho, and since
// rho^2 = det(P) / (a^3 rho0) (cons2prim's density), den^2 / det(P)^2 =
// (1 / (den rho0))^2 / a^2.  Likewise det A = rho0 / rho, so
// I3 = det(G) = (rho / rho0)^2.  t.A is formed only where a caller reads it.
template <bool ONE_FRAC, int FAST = -1>
__device__ __forceinline__ void thermo(const MatDev& m, double frac, const double* q, Thermo& t) {
  const double* P = q + 5;
  const double detP = det3cm(P);
  double ifr = 1.0, rho2;
  if (ONE_FRAC) {
    rho2 = detP * m.irho0;
  } else {
    ifr = rcp(frac);
    rho2 = detP * (ifr * ifr * ifr) * m.irho0;
  }
  t.bad = !(detP > 0.0);
  t.rho = sqrt_pos(rho2);  // NaN for a non-positive det, flagged above
  t.den = ONE_FRAC ? t.rho : frac * t.rho;
  const double iden = rcp(t.den);
  t.u[0] = q[1] * iden;
  t.u[1] = q[2] * iden;
  t.u[2] = q[3] * iden;
  t.E = q[4] * iden;
  double e = t.E - 0.5 * (t.u[0] * t.u[0] + t.u[1] * t.u[1] + t.u[2] * t.u[2]);
#pragma unroll
  for (int k = 0; k < 9; ++k) t.A[k] = P[k] * iden;
  // P P^T (symmetric) and its adjugate
  double B11 = P[0] * P[0] + P[3] * P[3] + P[6] * P[6];
  double B22 = P[1] * P[1] + P[4] * P[4] + P[7] * P[7];
  double B33 = P[2] * P[2] + P[5] * P[5] + P[8] * P[8];
  double B12 = P[0] * P[1] + P[3] * P[4] + P[6] * P[7];
  double B13 = P[0] * P[2] + P[3] * P[5] + P[6] * P[8];
  double B23 = P[1] * P[2] + P[4] * P[5] + P[7] * P[8];
  const double ir = iden * m.irho0;
  const double sc = ONE_FRAC ? ir * ir : (ir * ir) * (ifr * ifr);  // den^2 / det(P)^2
  t.G11 = (B22 * B33 - B23 * B23) * sc;
  t.G22 = (B11 * B33 - B13 * B13) * sc;
  t.G33 = (B11 * B22 - B12 * B12) * sc;
  t.G12 = (B13 * B23 - B12 * B33) * sc;
  t.G13 = (B12 * B23 - B13 * B22) * sc;
  t.G23 = (B12 * B13 - B11 * B23) * sc;
  t.i1 = t.G11 + t.G22 + t.G33;
  double trG2 = t.G11 * t.G11 + t.G22 * t.G22 + t.G33 * t.G33 +
                2.0 * (t.G12 * t.G12 + t.G13 * t.G13 + t.G23 * t.G23);
  t.i2 = 0.5 * (t.i1 * t.i1 - trG2);
  const double s1 = t.rho * m.irho0;  // I3^(1/2) = 1 / det A = rho / rho0
  t.i3 = s1 * s1;
  double ith;  // 1 / (cv t0 s_g)
  if (FAST == 1 || (FAST == -1 && m.fast)) {
    t.sa = s1;
    t.sb = t.i3 * s1;
    t.sg = t.i3;
    // 1 / I3 = (rho0 / rho)^2 = (rho0 a iden)^2
    const double r0i = ONE_FRAC ? m.rho0 * iden : m.rho0 * (frac * iden);
    ith = (r0i * r0i) * m.icvt0;
  } else {
    double li3 = log(t.i3);
    t.sa = exp(m.half_a * li3);
    t.sb = exp(m.half_b * li3);
    t.sg = exp(m.half_g * li3);
    ith = rcp(m.cvt0 * t.sg);
  }
  t.W = m.hb0sq * t.sb * (t.i1 * t.i1 * (1.0 / 3.0) - t.i2);
  double sam1 = t.sa - 1.0;
  double Uc = m.hk0a2 * sam1 * sam1;
  double x = (e - t.W - Uc) * ith + 1.0;
  if (x < 1e-6) x = 1e-6;
  t.x = x;
  t.Eth = m.cvt0 * t.sg * (x - 1.0);
  t.bad |= !(e == e) | !(t.rho < 1e300);
}

// Entropy S = cv log x (EquationsOfState.jl:155).
__device__ __forceinline__ double entropy_of(const MatDev& m, const Thermo& t) {
  return m.cv * log(t.x);
}

// c3 = I3 de/dI3 at fixed entropy.
__device__ __forceinline__ double c3_of(const MatDev& m, const Thermo& t) {
  return m.k0_2a * (t.sa - 1.0) * t.sa + m.half_g * t.Eth + m.half_b * t.W;
}

// First row of the Cauchy stress sigma (= first column; sigma is symmetric).
__device__ __forceinline__ void stress_row1(const MatDev& m, const Thermo& t, double s[3]) {
  double G2_11 = t.G11 * t.G11 + t.G12 * t.G12 + t.G13 * t.G13;
  double G2_12 = t.G11 * t.G12 + t.G12 * t.G22 + t.G13 * t.G23;
  double G2_13 = t.G11 * t.G13 + t.G12 * t.G23 + t.G13 * t.G33;
  double kap = m.hb0sq * t.sb;
  double c3 = c3_of(m, t);
  double i13 = t.i1 * (1.0 / 3.0);
  double m2r = -2.0 * t.rho;
  s[0] = m2r * (kap * (G2_11 - i13 * t.G11) + c3);
  s[1] = m2r * (kap * (G2_12 - i13 * t.G12));
  s[2] = m2r * (kap * (G2_13 - i13 * t.G13));
}

// Full stress, column-major 9-vector.
__device__ __forceinline__ void stress_full(const MatDev& m, const Thermo& t, double s[9]) {
  double G[9] = {t.G11, t.G12, t.G13, t.G12, t.G22, t.G23, t.G13, t.G23, t.G33};
  double kap = m.hb0sq * t.sb;
  double c3 = c3_of(m, t);
  double i13 = t.i1 * (1.0 / 3.0);
  double m2r = -2.0 * t.rho;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double g2 = G[i] * G[3 * j] + G[i + 3] * G[3 * j + 1] + G[i + 6] * G[3 * j + 2];
      double h = kap * (g2 - i13 * G[i + 3 * j]) + (i == j ? c3 : 0.0);
      s[i + 3 * j] = m2r * h;
    }
}

// Temperature de/dS = t0 s_g x (HyperelasticityMPh.jl:212).
__device__ __forceinline__ double temperature_of(const MatDev& m, const Thermo& t) {
  return m.t0 * t.sg * t.x;
}

// Acoustic tensor for n = e1 (the 1D solver's normal), symmetric: returns
// O11 O12 O13 O22 O23 O33.  The stress is -2 rho H with H = kappa T + c3 I,
// T = G^2 - (I1/3) G and kappa = b0^2/2 I3^(beta/2) (stress_row1); the
// acoustic tensor of EquationsOfState.jl:223-246 differentiates it along n,
// with G' = -(e1 G_1. + G_.1 e1^T) and the invariants' and I3's chain rules.
// Collecting the terms by hand gives
//   O11 = 2 kappa ((3 + 2 beta) T11 + 4/3 G11^2)
//         + s_a (k0/alpha (s_a - 1) + k0 (2 s_a - 1)) + gamma (1 + gamma) E_th
//         + beta (1 + beta) W
//   O1k = 2 kappa ((2 + beta) T1k + 4/3 G11 G1k)          k = 2, 3
//   Ojk = 2 kappa (Tjk + G11 Gjk + 1/3 G1j G1k)           j, k = 2, 3
// (the term-by-term derivative, which this replaces, agrees to round-off).
__device__ __forceinline__ void acoustic_e1(const MatDev& m, const Thermo& t, double O[6]) {
  const double G11 = t.G11, G22 = t.G22, G33 = t.G33, G12 = t.G12, G13 = t.G13, G23 = t.G23;
  const double i13 = t.i1 * (1.0 / 3.0);
  const double T11 = G11 * G11 + G12 * G12 + G13 * G13 - i13 * G11;
  const double T12 = G11 * G12 + G12 * G22 + G13 * G23 - i13 * G12;
  const double T13 = G11 * G13 + G12 * G23 + G13 * G33 - i13 * G13;
  const double T22 = G12 * G12 + G22 * G22 + G23 * G23 - i13 * G22;
  const double T23 = G12 * G13 + G22 * G23 + G23 * G33 - i13 * G23;
  const double T33 = G13 * G13 + G23 * G23 + G33 * G33 - i13 * G33;
  const double k2 = m.b0sq * t.sb;  // 2 kappa
  const double g43 = (4.0 / 3.0) * G11;
  const double sa = t.sa;
  O[0] = k2 * (m.ac_t * T11 + g43 * G11) +
         (sa * (m.k0_a * (sa - 1.0) + m.k0 * (2.0 * sa - 1.0)) + m.ac_e * t.Eth + m.ac_w * t.W);
  O[1] = k2 * (m.ac_s * T12 + g43 * G12);
  O[2] = k2 * (m.ac_s * T13 + g43 * G13);
  const double g12 = (1.0 / 3.0) * G12, g13 = (1.0 / 3.0) * G13;
  O[3] = k2 * (T22 + G11 * G22 + g12 * G12);
  O[4] = k2 * (T23 + G11 * G23 + g12 * G13);
  O[5] = k2 * (T33 + G11 * G33 + g13 * G13);
}

// Acoustic tensor for a general direction n, column-major 3x3.
//   fixed_e == false: acoustic(eos, ent, F, n) (EquationsOfState.jl:223-246):
//                     entropy held, rho = rho0/det A varies with A.
//   fixed_e == true : acoustic(eos, P, n) (EquationsOfState.jl:192-221): the
//                     stress is stress(eos, den, e_int, F) with den and e_int
//                     held, so the entropy follows A (unless the entropy law's
//                     clamp is active, `clamped`) and there is no density term.
__device__ __forceinline__ void acoustic_n(const MatDev& m, const Thermo& t, const double n[3],
                                           double O[9], bool fixed_e = false,
                                           bool clamped = false) {
  double G[9] = {t.G11, t.G12, t.G13, t.G12, t.G22, t.G23, t.G13, t.G23, t.G33};
  double G2[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      G2[i + 3 * j] = G[i] * G[3 * j] + G[i + 3] * G[3 * j + 1] + G[i + 6] * G[3 * j + 2];
  double gn[3], g2n[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    gn[i] = G[i] * n[0] + G[i + 3] * n[1] + G[i + 6] * n[2];
    g2n[i] = G2[i] * n[0] + G2[i + 3] * n[1] + G2[i + 6] * n[2];
  }
  double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  double nGn = n[0] * gn[0] + n[1] * gn[1] + n[2] * gn[2];
  double kap = m.hb0sq * t.sb;
  double c3 = c3_of(m, t);
  double i1 = t.i1, i13 = i1 * (1.0 / 3.0);
  double Hn[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) Hn[i] = kap * (g2n[i] - i13 * gn[i]) + c3 * n[i];
  double hs = m.hb0sq * t.sb;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double nj = n[j];
    double di1 = -2.0 * gn[j];
    double di2 = i1 * di1 + 2.0 * g2n[j];
    double dkap = -m.beta * kap * nj;
    double dW = -m.beta * t.W * nj + hs * ((2.0 / 3.0) * i1 * di1 - di2);
    double dsa = -m.alpha * t.sa * nj;
    double dEth;
    if (fixed_e && !clamped) {
      double dUc = 2.0 * m.hk0a2 * (t.sa - 1.0) * dsa;
      dEth = -(dW + dUc);
    } else {
      dEth = -m.gamma * t.Eth * nj;
    }
    double dc3 = m.k0_2a * (2.0 * t.sa - 1.0) * dsa + m.half_g * dEth + m.half_b * dW;
    double di13 = di1 * (1.0 / 3.0);
    double rterm = fixed_e ? 0.0 : 2.0 * nj;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double ndG = -(gn[j] * n[i] + nn * G[i + 3 * j]);
      double ndG2 = -(g2n[j] * n[i] + nGn * G[i + 3 * j] + gn[j] * gn[i] + nn * G2[i + 3 * j]);
      double ndH = dkap * (g2n[i] - i13 * gn[i]) + kap * (ndG2 - di13 * gn[i] - i13 * ndG) +
                   dc3 * n[i];
      O[i + 3 * j] = rterm * Hn[i] - 2.0 * ndH;
    }
  }
}

// Largest root of t^3 - c2 t^2 + c1 t - c0 (all roots real) by Newton from an
// upper bound; the iteration decreases monotonically onto the root.
__device__ __forceinline__ double cubic_top_root(double c2, double c1, double c0, double ub,
                                                 double lb) {
  double t = ub;
#pragma unroll 1
  for (int it = 0; it < 60; ++it) {
    double f = ((t - c2) * t + c1) * t - c0;
    double fp = (3.0 * t - 2.0 * c2) * t + c1;
    if (!(fp > 0.0)) break;
    double tn = t - f * rcp(fp);
    if (!(tn < t)) break;
    t = tn < lb ? lb : tn;
  }
  return t;
}

// Smallest root, Newton from a lower bound (increases monotonically).
__device__ __forceinline__ double cubic_bottom_root(double c2, double c1, double c0, double lb,
                                                    double ub) {
  double t = lb;
#pragma unroll 1
  for (int it = 0; it < 60; ++it) {
    double f = ((t - c2) * t + c1) * t - c0;
    double fp = (3.0 * t - 2.0 * c2) * t + c1;
    if (!(fp > 0.0)) break;
    double tn = t - f * rcp(fp);
    if (!(tn > t)) break;
    t = tn > ub ? ub : tn;
  }
  return t;
}

// 1/x to about 2^-50 relative: the hardware reciprocal refined by one Newton
// step, for iterations that correct their own rounding.
__device__ __forceinline__ double rcp1(double x) {
  double r = __builtin_amdgcn_rcp(x);
  return fma(fma(-x, r, 1.0), r, r);
}

// HES_EIG_SECULAR = 1 leaves out sym3_extreme's shifted-cubic path, so that
// every state takes the secular-function iteration.  The adjoint's sources set
// it: their launches (a 2^18-cell problem runs a wave or two per SIMD) are
// bound by the latency of one wave's dependent chain, which the secular
// iteration keeps shorter (two steps of one reciprocal each, against a
// fixed-point step and two Newton steps), while the solver's and the batch's
// steps are bound by instruction throughput, where the cubic path issues
// about a third fewer.
#ifndef HES_EIG_SECULAR
#define HES_EIG_SECULAR 0
#endif

// Eigenvalue of largest modulus of a symmetric 3x3 given as O11 O12 O13 O22
// O23 O33.  HyperelasticityMPh.jl:263 takes sqrt(|eig|) of all three; only the
// largest modulus is ever consumed (main.jl:210, NumFluxes.jl:90-91).
//
// Write O = [a w^T; w D] with w = (O12, O13), D the trailing 2x2 block and
// M = a I - D.  Since adj(M + mu I) = adj(M) + mu I for a 2x2, the top
// eigenvalue is a + mu with mu the root in [0, inf) of the shifted cubic
//   h(mu) = det((a + mu) I - O) = mu^3 + T mu^2 + s mu - n0,
//   T = tr M, s = det M - |w|^2, n0 = w^T adj(M) w,
// whose coefficients carry no cancellation when a stands apart from D (the
// acoustic tensor's longitudinal entry above its shear block).  With T > 0
// and s > 0, h is increasing and convex on [0, inf), so that root is the
// only one there, and n0 / s (Newton from 0) bounds it from above.  One
// fixed-point step mu = n0 / (mu^2 + T mu + s) from that bound lands below
// the root, and Newton steps from there stay in [0, n0 / s], on which
// h'' / 2 <= 3 n0 / s + T and h' >= s.  A step st from mu, taken with slope
// h'(mu), then leaves an error of at most (3 n0 / s + T) h'(mu) st^2 / s^2,
// which is tested against 2^-52 (a + mu).  States with a 5% random distortion
// take two Newton steps, strongly sheared ones three or four.  A state with
// T <= 0 or s <= 0 (or that does not settle in ten steps) goes to the secular
// function
//   g(l) = l - a - w^T (l - D)^-1 w
// by Newton from l = a when a exceeds lDu = max(O22, O33) + |O23|, an upper
// bound of D's eigenvalues (g increases and is concave there, with g(a) <= 0,
// so the iterates rise monotonically onto the root; a step from l leaves an
// error of at most (p^2 + q^2) / det^2 / (l - lDu) step^2, p, q, det as
// below), and otherwise to Newton on the characteristic cubic from an upper
// bound.  A bottom eigenvalue of larger modulus needs tr O < top (it is
// tr O - top - (middle eigenvalue) >= tr O - 2 top) and a Gershgorin lower
// bound lb < -|top|; Newton from lb upwards then finds it.
__device__ __forceinline__ double sym3_extreme(const double O[6]) {
  const double a = O[0], b = O[3], c = O[5], d = O[1], e = O[2], f = O[4];
  double top = a;
  bool done = false;
  if constexpr (!HES_EIG_SECULAR) {
    const double amb = a - b, amc = a - c, T = amb + amc;
    const double d2 = d * d, e2 = e * e;
    const double s = fma(amb, amc, -(f * f)) - (d2 + e2);
    const double n0 = fma(amc, d2, fma(amb, e2, (2.0 * f) * d * e));
    if (T > 0.0 && s > 0.0) {
      const double is = rcp1(s);
      const double mu0 = n0 * is;                     // >= the root
      const double hb = fma(3.0, mu0, T) * (is * is);
      double mu = n0 * rcp1(fma(mu0 + T, mu0, s));    // <= the root
      double st = 0.0, hp = 1.0;
      auto newton = [&]() {
        const double h = fma(fma(mu + T, mu, s), mu, -n0);
        hp = fma(fma(3.0, mu, 2.0 * T), mu, s);
        st = -h * rcp1(hp);
        mu += st;
      };
      newton();
      newton();
      top = a + mu;
      done = hb * hp * (st * st) <= 0x1p-52 * top;  // false for NaN and overflow
#pragma unroll 1
      for (int it = 0; it < 8 && !done; ++it) {
        newton();
        top = a + mu;
        done = hb * hp * (st * st) <= 0x1p-52 * top;
      }
    }
  }
  if (!done) {
    const double lDu = fmax(b, c) + fabs(f);
    top = a;
    if (a > lDu) {
      const double ff = f * f, fd = f * d, fe = f * e;
#pragma unroll 1
      for (int it = 0; it < 24; ++it) {
        const double x = top - b, y = top - c;
        const double det = x * y - ff;       // det(l - D) > 0
        const double p = y * d + fe, q = x * e + fd;  // adj(l - D) w
        const double h = (top - a) * det - (d * p + e * q);  // g(l) det
        const double pq2 = p * p + q * q, det2 = det * det;
        const double st = -h * det * rcp1(det2 + pq2);  // -g / g'
        const double lo = top;
        top += st;
        if (!(pq2 * (st * st) > 0x1p-52 * lo * (lo - lDu) * det2)) {
          done = true;
          break;
        }
      }
    }
    if (!done) {
      const double r1 = fabs(d) + fabs(e), r2 = fabs(d) + fabs(f), r3 = fabs(e) + fabs(f);
      const double c2 = a + b + c;
      const double c1 = a * b - d * d + a * c - e * e + b * c - f * f;
      const double c0 = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
      const double ub = fmax(a + r1, fmax(b + r2, c + r3));  // Gershgorin
      const double lb = fmin(a - r1, fmin(b - r2, c - r3));
      top = cubic_top_root(c2, c1, c0, ub, lb);
    }
  }
  double mu = top;
  if (!(a + b + c >= top)) {
    const double r1 = fabs(d) + fabs(e), r2 = fabs(d) + fabs(f), r3 = fabs(e) + fabs(f);
    const double lb = fmin(a - r1, fmin(b - r2, c - r3));
    if (lb < -fabs(top)) {
      const double c2 = a + b + c;
      const double c1 = a * b - d * d + a * c - e * e + b * c - f * f;
      const double c0 = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
      const double bot = cubic_bottom_root(c2, c1, c0, lb, fmax(a + r1, fmax(b + r2, c + r3)));
      if (fabs(bot) > fabs(top)) mu = bot;
    }
  }
  return mu;
}

// Spectral radius of a symmetric 3x3 (O11 O12 O13 O22 O23 O33).
__device__ __forceinline__ double spectral_radius_sym(const double O[6]) {
  return fabs(sym3_extreme(O));
}

// All three eigenvalues of a symmetric 3x3 (descending): the largest by Newton,
// the other two from the deflated quadratic.
__device__ __forceinline__ void eig3_sym(const double O[6], double w[3]) {
  double a = O[0], b = O[3], c = O[5], d = O[1], e = O[2], f = O[4];
  double c2 = a + b + c;
  double c1 = a * b - d * d + a * c - e * e + b * c - f * f;
  double c0 = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
  double r1 = fabs(d) + fabs(e), r2 = fabs(d) + fabs(f), r3 = fabs(e) + fabs(f);
  double ub = fmax(a + r1, fmax(b + r2, c + r3));
  double lb = fmin(a - r1, fmin(b - r2, c - r3));
  double l1 = cubic_top_root(c2, c1, c0, ub, lb);
  double l3 = cubic_bottom_root(c2, c1, c0, lb, ub);
  double l2 = c2 - l1 - l3;
  w[0] = l1;
  w[1] = l2;
  w[2] = l3;
}

// Sound speed c_max = sqrt(rho(Omega)) of a phase for n = e1.
__device__ __forceinline__ double sound_speed(const MatDev& m, const Thermo& t) {
  double O[6];
  acoustic_e1(m, t, O);
  const double rad = spectral_radius_sym(O);
  return rad > 0.0 ? sqrt_pos(rad) : 0.0;
}

// Physical flux of a phase (HyperelasticityMPh.jl:146-175) into f[14]
// = flux block without the volume-fraction entry (which is 0).  P is the
// phase's conserved a rho vec(A), so den A = P in the distortion rows.
__device__ __forceinline__ void phase_flux(const MatDev& m, double frac, const Thermo& t,
                                           const double* P, double f[14]) {
  double s[3];
  stress_row1(m, t, s);
  double st0 = frac * s[0], st1 = frac * s[1], st2 = frac * s[2];
  const double* u = t.u;
  double du1 = t.den * u[0];
  f[0] = du1;
  f[1] = du1 * u[0] - st0;
  f[2] = du1 * u[1] - st1;
  f[3] = du1 * u[2] - st2;
  f[4] = du1 * t.E - (u[0] * st0 + u[1] * st1 + u[2] * st2);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f[5 + 3 * c] = 0.0;  // row 1: u1 A_1c - u1 A_1c
    f[6 + 3 * c] = u[0] * P[1 + 3 * c] - u[1] * P[3 * c];
    f[7 + 3 * c] = u[0] * P[2 + 3 * c] - u[2] * P[3 * c];
  }
}

// Energy e(S, G) of a phase given A and x = exp(S/cv) (prim->cons).
__device__ __forceinline__ double energy_from_A(const MatDev& m, const double A[9], double x) {
  double B11 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6];
  double B22 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7];
  double B33 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
  double B12 = A[0] * A[1] + A[3] * A[4] + A[6] * A[7];
  double B13 = A[0] * A[2] + A[3] * A[5] + A[6] * A[8];
  double B23 = A[1] * A[2] + A[4] * A[5] + A[7] * A[8];
  double C11 = B22 * B33 - B23 * B23;
  double C12 = B13 * B23 - B12 * B33;
  double C13 = B12 * B23 - B13 * B22;
  double detB = B11 * C11 + B12 * C12 + B13 * C13;
  double idetB = 1.0 / detB;
  double G11 = C11 * idetB, G22 = (B11 * B33 - B13 * B13) * idetB;
  double G33 = (B11 * B22 - B12 * B12) * idetB;
  double G12 = C12 * idetB, G13 = C13 * idetB, G23 = (B12 * B13 - B11 * B23) * idetB;
  double i1 = G11 + G22 + G33;
  double trG2 = G11 * G11 + G22 * G22 + G33 * G33 + 2.0 * (G12 * G12 + G13 * G13 + G23 * G23);
  double i2 = 0.5 * (i1 * i1 - trG2);
  double i3 = idetB;
  double li3 = log(i3);
  double sa = exp(m.half_a * li3), sb = exp(m.half_b * li3), sg = exp(m.half_g * li3);
  double sam1 = sa - 1.0;
  return m.hk0a2 * sam1 * sam1 + m.cvt0 * sg * (x - 1.0) +
         m.hb0sq * sb * (i1 * i1 * (1.0 / 3.0) - i2);
}

// Invariants of a general (column-major) G and the energy/entropy laws as
// functions of G (EquationsOfState.jl:118-156), for the point-wise API.
__device__ __forceinline__ void invariants_gen(const double* g, double inv[3]) {
  double tr = g[0] + g[4] + g[8];
  double trg2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) trg2 += g[i + 3 * k] * g[k + 3 * i];
  inv[0] = tr;
  inv[1] = 0.5 * (tr * tr - trg2);
  inv[2] = det3cm(g);
}

__device__ __forceinline__ double energy_G(const MatDev& m, double S, const double* g) {
  double inv[3];
  invariants_gen(g, inv);
  double li3 = log(inv[2]);
  double sa = exp(m.half_a * li3), sb = exp(m.half_b * li3), sg = exp(m.half_g * li3);
  double sam1 = sa - 1.0;
  return m.hk0a2 * sam1 * sam1 + m.cvt0 * sg * (exp(S / m.cv) - 1.0) +
         m.hb0sq * sb * (inv[0] * inv[0] * (1.0 / 3.0) - inv[1]);
}

__device__ __forceinline__ double entropy_G(const MatDev& m, double e_int, const double* g) {
  double inv[3];
  invariants_gen(g, inv);
  double li3 = log(inv[2]);
  double sa = exp(m.half_a * li3), sb = exp(m.half_b * li3), sg = exp(m.half_g * li3);
  double sam1 = sa - 1.0;
  double S = e_int - m.hb0sq * sb * (inv[0] * inv[0] * (1.0 / 3.0) - inv[1]) -
             m.hk0a2 * sam1 * sam1;
  S = S / (m.cvt0 * sg) + 1.0;
  if (S < 1e-6) S = 1e-6;
  return log(S) * m.cv;
}

// Thermo state of a phase from a distortion A and x = exp(S/cv) (rho = rho0/det A),
// for the point-wise stress/acoustic API.
__device__ __forceinline__ void thermo_from_A(const MatDev& m, const double* A, double x,
                                              Thermo& t) {
  double detA = det3cm(A);
  t.bad = !(detA > 0.0);
  t.rho = m.rho0 / detA;
  t.den = t.rho;
#pragma unroll
  for (int k = 0; k < 9; ++k) t.A[k] = A[k];
  t.u[0] = t.u[1] = t.u[2] = 0.0;
  t.E = 0.0;
  double B11 = A[0] * A[0] + A[3] * A[3] + A[6] * A[6];
  double B22 = A[1] * A[1] + A[4] * A[4] + A[7] * A[7];
  double B33 = A[2] * A[2] + A[5] * A[5] + A[8] * A[8];
  double B12 = A[0] * A[1] + A[3] * A[4] + A[6] * A[7];
  double B13 = A[0] * A[2] + A[3] * A[5] + A[6] * A[8];
  double B23 = A[1] * A[2] + A[4] * A[5] + A[7] * A[8];
  double C11 = B22 * B33 - B23 * B23;
  double C12 = B13 * B23 - B12 * B33;
  double C13 = B12 * B23 - B13 * B22;
  double detB = B11 * C11 + B12 * C12 + B13 * C13;
  double idetB = 1.0 / detB;
  t.G11 = C11 * idetB;
  t.G22 = (B11 * B33 - B13 * B13) * idetB;
  t.G33 = (B11 * B22 - B12 * B12) * idetB;
  t.G12 = C12 * idetB;
  t.G13 = C13 * idetB;
  t.G23 = (B12 * B13 - B11 * B23) * idetB;
  t.i1 = t.G11 + t.G22 + t.G33;
  double trG2 = t.G11 * t.G11 + t.G22 * t.G22 + t.G33 * t.G33 +
                2.0 * (t.G12 * t.G12 + t.G13 * t.G13 + t.G23 * t.G23);
  t.i2 = 0.5 * (t.i1 * t.i1 - trG2);
  t.i3 = idetB;
  double li3 = log(t.i3);
  t.sa = exp(m.half_a * li3);
  t.sb = exp(m.half_b * li3);
  t.sg = exp(m.half_g * li3);
  t.W = m.hb0sq * t.sb * (t.i1 * t.i1 * (1.0 / 3.0) - t.i2);
  t.x = x;
  t.Eth = m.cvt0 * t.sg * (x - 1.0);
}

}  // namespace hes
