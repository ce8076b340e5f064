// hes_faces.hpp -- per-cell and per-face building blocks of the time step:
// the physical flux and wave-speed bounds of a cell, the two-phase
// non-conservative column b^p, its path integrals, and the Rusanov / HLL
// path-conservative face fluxes (NumFluxes.jl:25-132).
#pragma once
#include "hes_device.hpp"

namespace hes {

// Quadrature on [0,1]: (x + 1)/2 and w/2 of the 6-point Gauss-Legendre and
// Gauss-Lobatto rules on [-1,1] (NumFluxes.jl:36-37, 94-95).
#define HES_GL_X0 0.93246951420315202781
#define HES_GL_X1 0.66120938646626451366
#define HES_GL_X2 0.23861918608319690863
#define HES_GL_W0 0.17132449237917034504
#define HES_GL_W1 0.36076157304813860757
#define HES_GL_W2 0.46791393457269104739
#define HES_LO_X1 0.76505532392946469285
#define HES_LO_X2 0.28523151648064509632
#define HES_LO_W0 0.066666666666666666667
#define HES_LO_W1 0.37847495629784698032
#define HES_LO_W2 0.55485837703548635301

static __constant__ double c_gl_s[6] = {(-HES_GL_X0 + 1.0) / 2.0, (-HES_GL_X1 + 1.0) / 2.0,
                                 (-HES_GL_X2 + 1.0) / 2.0, (HES_GL_X2 + 1.0) / 2.0,
                                 (HES_GL_X1 + 1.0) / 2.0,  (HES_GL_X0 + 1.0) / 2.0};
static __constant__ double c_gl_w[6] = {HES_GL_W0 / 2.0, HES_GL_W1 / 2.0, HES_GL_W2 / 2.0,
                                 HES_GL_W2 / 2.0, HES_GL_W1 / 2.0, HES_GL_W0 / 2.0};
static __constant__ double c_lo_s[6] = {(-1.0 + 1.0) / 2.0,        (-HES_LO_X1 + 1.0) / 2.0,
                                 (-HES_LO_X2 + 1.0) / 2.0, (HES_LO_X2 + 1.0) / 2.0,
                                 (HES_LO_X1 + 1.0) / 2.0,  (1.0 + 1.0) / 2.0};
static __constant__ double c_lo_w[6] = {HES_LO_W0 / 2.0, HES_LO_W1 / 2.0, HES_LO_W2 / 2.0,
                                 HES_LO_W2 / 2.0, HES_LO_W1 / 2.0, HES_LO_W0 / 2.0};

enum { FLUX_RUSANOV = 0, FLUX_HLL = 1 };

template <int NPH>
struct NComp { static constexpr int value = NPH == 1 ? 14 : 30; };

// Thermo state of phase p of a state vector in the model layout.
template <int NPH, int FAST = -1>
__device__ __forceinline__ void phase_thermo(const MatDev* mat, const double* q, int p, Thermo& t) {
  if (NPH == 1) thermo<true, FAST>(mat[0], 1.0, q, t);
  else thermo<false, FAST>(mat[p], q[15 * p], q + 15 * p + 1, t);
}

// Flux F(Q) of a cell and the wave-speed bounds of its eigenvalues
// [min_p(u1 - c_p), max_p(u1 + c_p)] from the stored sound speeds cc[p]
// (NOC, one phase: from the cell's state instead).
template <int NPH, int FAST = -1, bool NOC = false>
__device__ __forceinline__ void cell_flux(const MatDev* mat, const double* q, const double* cc,
                                          double* f, double& smin, double& smax, int& bad) {
  if (NPH == 1) {
    Thermo t;
    thermo<true, FAST>(mat[0], 1.0, q, t);
    phase_flux(mat[0], 1.0, t, q + 5, f);
    // NOC: the sound speed from this thermodynamic state (the same bits as
    // the stored one, which the step that wrote the state computed so)
    const double c0 = NOC ? sound_speed(mat[0], t) : cc[0];
    smin = t.u[0] - c0;
    smax = t.u[0] + c0;
    bad = t.bad;
  } else {
    bad = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      Thermo t;
      thermo<false, FAST>(mat[p], q[15 * p], q + 15 * p + 1, t);
      f[15 * p] = 0.0;
      phase_flux(mat[p], q[15 * p], t, q + 15 * p + 6, f + 15 * p + 1);
      double lo = t.u[0] - cc[p], hi = t.u[0] + cc[p];
      smin = p == 0 ? lo : fmin(smin, lo);
      smax = p == 0 ? hi : fmax(smax, hi);
      bad |= t.bad;
    }
  }
}

// Sound speeds of a cell (per phase) and its CFL speed max_p(|u1_p| + c_p)
// (main.jl:209-210).
template <int NPH, int FAST = -1>
__device__ __forceinline__ double cell_speeds(const MatDev* mat, const double* q, double* cc,
                                              int& bad) {
  double lam = 0.0;
  bad = 0;
#pragma unroll
  for (int p = 0; p < NPH; ++p) {
    Thermo t;
    phase_thermo<NPH, FAST>(mat, q, p, t);
    double c = sound_speed(mat[NPH == 1 ? 0 : p], t);
    cc[p] = c;
    lam = fmax(lam, fabs(t.u[0]) + c);
    bad |= t.bad | !(c < 1e300);
  }
  return lam;
}

// Column 1 of the non-conservative matrix of each phase block at state q
// (HyperelasticityMPh.jl:178-231 with omega = 0, k = (1/2, 1/2), beta = 0).
// Rows 1, 3:5 and 6 are common to both blocks; rows 7:15 are per phase.
struct BCol {
  double uI0;       // row 1: u_I,1
  double sI[3];     // rows 3:5: sigma_I(:,1)
  double sIu;       // row 6: sigma_I(:,1) . u_I
  double a[2][9];   // rows 7:15 of block p
};

template <int FAST = -1>
__device__ __forceinline__ void bcol2(const MatDev* mat, const double* q, BCol& b) {
  Thermo t0, t1;
  thermo<false, FAST>(mat[0], q[0], q + 1, t0);
  thermo<false, FAST>(mat[1], q[15], q + 16, t1);
  double s0[3], s1[3];
  stress_row1(mat[0], t0, s0);
  stress_row1(mat[1], t1, s1);
  double th0 = temperature_of(mat[0], t0), th1 = temperature_of(mat[1], t1);
  // sigma_I = (k2 th2 K1 + k1 th1 K2) / (k1 th1 + k2 th2), K_p = sigma_p
  double iden = rcp(0.5 * th0 + 0.5 * th1);
  double w0 = 0.5 * th1 * iden, w1 = 0.5 * th0 * iden;
  double uI[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uI[k] = 0.5 * t0.u[k] + 0.5 * t1.u[k];
    b.sI[k] = w0 * s0[k] + w1 * s1[k];
  }
  b.uI0 = uI[0];
  b.sIu = b.sI[0] * uI[0] + b.sI[1] * uI[1] + b.sI[2] * uI[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const Thermo& t = p == 0 ? t0 : t1;
    double dv0 = uI[0] - t.u[0], dv1 = uI[1] - t.u[1], dv2 = uI[2] - t.u[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double ra = t.rho * t.A[3 * c];
      double atdv = t.A[3 * c] * dv0 + t.A[1 + 3 * c] * dv1 + t.A[2 + 3 * c] * dv2;
      b.a[p][3 * c + 0] = ra * t.u[0] + t.rho * atdv;
      b.a[p][3 * c + 1] = ra * t.u[1];
      b.a[p][3 * c + 2] = ra * t.u[2];
    }
  }
}

// Integral of B(psi) dpsi/ds along psi(s) = qa (1 - s) + qb s over [0,1]
// (NumFluxes.jl:97-107 / 39-47).  B has one non-zero column per phase block,
// so the integral is (sum_q w_q b^p(psi_q)) times the volume-fraction jump.
template <bool LOBATTO, int FAST = -1>
__device__ void bint2(const MatDev* mat, const double* qa, const double* qb, double* out) {
  double da0 = qb[0] - qa[0], da1 = qb[15] - qa[15];
  if (da0 == 0.0 && da1 == 0.0) {
#pragma unroll
    for (int r = 0; r < 30; ++r) out[r] = 0.0;
    return;
  }
  double acc_u = 0.0, acc_s0 = 0.0, acc_s1 = 0.0, acc_s2 = 0.0, acc_su = 0.0;
  double acc_a[18];
#pragma unroll
  for (int r = 0; r < 18; ++r) acc_a[r] = 0.0;
#pragma unroll 1
  for (int k = 0; k < 6; ++k) {
    double s = LOBATTO ? c_lo_s[k] : c_gl_s[k];
    double w = LOBATTO ? c_lo_w[k] : c_gl_w[k];
    double oms = 1.0 - s;
    double psi[30];
#pragma unroll
    for (int r = 0; r < 30; ++r) psi[r] = qa[r] * oms + qb[r] * s;
    BCol b;
    bcol2<FAST>(mat, psi, b);
    acc_u += w * b.uI0;
    acc_s0 += w * b.sI[0];
    acc_s1 += w * b.sI[1];
    acc_s2 += w * b.sI[2];
    acc_su += w * b.sIu;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      acc_a[r] += w * b.a[0][r];
      acc_a[9 + r] += w * b.a[1][r];
    }
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double da = p == 0 ? da0 : da1;
    double* o = out + 15 * p;
    o[0] = acc_u * da;
    o[1] = 0.0;
    o[2] = acc_s0 * da;
    o[3] = acc_s1 * da;
    o[4] = acc_s2 * da;
    o[5] = acc_su * da;
#pragma unroll
    for (int r = 0; r < 9; ++r) o[6 + r] = acc_a[9 * p + r] * da;
  }
}

// Rusanov / LxF face (NumFluxes.jl:25-60): conservative flux and the common
// half path integral D- = D+.
template <int NPH, int FAST = -1>
__device__ __forceinline__ void face_rusanov(const MatDev* mat, const double* ql, const double* qr,
                                             const double* fl, const double* fr, double lam,
                                             double* cons, double* dmp) {
  constexpr int NC = NComp<NPH>::value;
  double hl = 0.5 * lam;
#pragma unroll
  for (int r = 0; r < NC; ++r) cons[r] = 0.5 * (fl[r] + fr[r]) - hl * (qr[r] - ql[r]);
  if (NPH == 2) {
    bint2<true, FAST>(mat, ql, qr, dmp);
#pragma unroll
    for (int r = 0; r < NC; ++r) dmp[r] = 0.5 * dmp[r];
  }
}

// HLL wave-speed bounds and the coefficients of D- and D+ (NumFluxes.jl:88-92,
// 111-131): s_l = min(0, min eig(Q_m), min eig(Q_l)), s_r = max(0, max eig(Q_m),
// max eig(Q_r)); D- = cl J + clr dQ, D+ = cr J - clr dQ.
struct HllCoef {
  double s_l, s_r, isd, cl, cr, clr;
};

__device__ __forceinline__ HllCoef hll_coef(double mn_m, double mx_m, double smin_l,
                                            double smax_r) {
  HllCoef c;
  c.s_l = fmin(0.0, fmin(mn_m, smin_l));
  c.s_r = fmax(0.0, fmax(mx_m, smax_r));
  const double sd = c.s_r - c.s_l;
  c.isd = rcp(sd);
  c.cl = -c.s_l * c.isd;
  c.cr = c.s_r * c.isd;
  c.clr = c.s_l * c.s_r * c.isd;
  return c;
}

// Eigenvalue range of the mid state Q_m = (Q_l + Q_r) / 2 of a face, formed
// from the sum S = Q_l + Q_r = 2 Q_m with volume fraction 2 a_m (one phase:
// 2): thermo's density takes det(S) / (2 a_m)^3, and every power of two
// scales out exactly, so the state and its range are Q_m's to the bit.
template <int NPH, int FAST = -1>
__device__ __forceinline__ void mid_speed_range(const MatDev* mat, const double* ql,
                                                const double* qr, double& mn_m, double& mx_m) {
  constexpr int NC = NComp<NPH>::value;
  double qs[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) qs[r] = ql[r] + qr[r];
#pragma unroll
  for (int p = 0; p < NPH; ++p) {
    Thermo t;
    if (NPH == 1) thermo<false, FAST>(mat[0], 2.0, qs, t);
    else thermo<false, FAST>(mat[p], qs[15 * p], qs + 15 * p + 1, t);
    const double c = sound_speed(mat[NPH == 1 ? 0 : p], t);
    const double lo = t.u[0] - c, hi = t.u[0] + c;
    mn_m = p == 0 ? lo : fmin(mn_m, lo);
    mx_m = p == 0 ? hi : fmax(mx_m, hi);
  }
}

// HLL face (NumFluxes.jl:70-132): D- and D+ (the conservative part is zero).
// smin_l = min eig of the left cell, smax_r = max eig of the right cell.  When
// coef is given it receives the coefficients used.
template <int NPH, int FAST = -1>
__device__ __forceinline__ void face_hll(const MatDev* mat, const double* ql, const double* qr,
                                         const double* fl, const double* fr, double smin_l,
                                         double smax_r, double* dm, double* dp,
                                         HllCoef* coef = nullptr) {
  constexpr int NC = NComp<NPH>::value;
  double mn_m, mx_m;
  mid_speed_range<NPH, FAST>(mat, ql, qr, mn_m, mx_m);
  const HllCoef hc = hll_coef(mn_m, mx_m, smin_l, smax_r);
  if (coef) *coef = hc;
  const double s_l = hc.s_l, s_r = hc.s_r, isd = hc.isd, cl = hc.cl, cr = hc.cr, clr = hc.clr;
  if (NPH == 1) {
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      double J = fr[r] - fl[r];
      double dq = qr[r] - ql[r];
      dm[r] = cl * J + clr * dq;
      dp[r] = cr * J - clr * dq;
    }
  } else {
    // path_int = B_int(Q_l, Q_r) + F(Q_r) - F(Q_l)
    double J[NC];
    bint2<false, FAST>(mat, ql, qr, J);
#pragma unroll
    for (int r = 0; r < NC; ++r) J[r] = J[r] + (fr[r] - fl[r]);
    // Q_hll; the volume-fraction rows are written as a_l + (s_r (a_r - a_l) - P)/(s_r - s_l),
    // algebraically the same and exact when a_l == a_r with no path jump.
    double qh[NC];
#pragma unroll
    for (int r = 0; r < NC; ++r) qh[r] = (qr[r] * s_r - ql[r] * s_l - J[r]) * isd;
    qh[0] = ql[0] + (s_r * (qr[0] - ql[0]) - J[0]) * isd;
    qh[15] = ql[15] + (s_r * (qr[15] - ql[15]) - J[15]) * isd;
    double b[NC];
#pragma unroll
    for (int r = 0; r < NC; ++r) J[r] = fr[r] - fl[r];
    bint2<false, FAST>(mat, ql, qh, b);
#pragma unroll
    for (int r = 0; r < NC; ++r) J[r] = J[r] + b[r];
    bint2<false, FAST>(mat, qh, qr, b);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      double Jr = J[r] + b[r];
      double dq = qr[r] - ql[r];
      dm[r] = cl * Jr + clr * dq;
      dp[r] = cr * Jr - clr * dq;
    }
  }
}

}  // namespace hes
