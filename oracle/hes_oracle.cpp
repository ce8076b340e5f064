// hes_oracle.cpp -- CPU restatement of the reference's finite-volume time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU path is checked
// against; it is never part of the product.  Only tests/, the smoke entry point
// and the CPU-baseline leg of the benchmark may load the library built from it.
//
// It restates, operation by operation, the Julia reference:
//   SimpleLA.jl:51-96            inv_, det_, tr_
//   Strains.jl:26-52             finger, invariants
//   EquationsOfState.jl:90-156   Barton2009 constructor, energy, entropy
//   EquationsOfState.jl:162-246  stress (both methods), acoustic (both methods)
//   HyperelasticityMPh.jl:63-266 prim2cons, cons2prim, flux, noncons_flux, get_eigvals
//   NumFluxes.jl:25-132          lxf / lxf_pathcons, hll / hll_pathcons
//   main.jl:43-60, 202-241       update_cell and the time loop (CFL, frozen end cells)
//
// Third-party arithmetic the reference calls, restated here:
//   * ForwardDiff.gradient / jacobian / derivative -> forward-mode dual numbers
//     (Dual<T,N> below; nesting Dual<Dual<double,9>,9> reproduces the
//     gradient-inside-jacobian structure of acoustic()).
//   * LinearAlgebra.det on a 3x3 -> LU with partial pivoting (LAPACK getrf order).
//   * LinearAlgebra.eigvals -> cyclic Jacobi on the symmetric part of the 3x3
//     acoustic tensor (the tensor is symmetric up to round-off; tests pin the
//     result against LAPACK dgeev through numpy).
//   * FastGaussQuadrature gausslegendre(6)/gausslobatto(6) -> nodes and weights
//     as correctly rounded constants (tests pin them against numpy/mpmath).
// Bit-level agreement with Julia (ForwardDiff op order, LAPACK rounding) is not
// claimed; agreement is to a few ulps per operation.
//
// The single-phase path (EOS, cons2prim/prim2cons, flux, get_eigvals, lxf and
// the time loop) is templated on its scalar type, so that ForwardDiff.gradient
// of a misfit through the whole run is restated too (ora_adjoint_dual: 28
// tangents carried by Dual<double,28> beneath the EOS's nested duals; the
// eigenvalues of a dual matrix take first-order perturbation partials).  The
// double instantiations are the same operations as before, bit for bit.
//
// Layout: a phase block has 15 components [a, a*r, a*r*u(3), a*r*E, a*r*A(9)],
// A column-major (HyperelasticityMPh.jl:80-84).  Single-phase states have 14
// components: the phase block without the volume fraction (a == 1).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace ora {

// ---------------------------------------------------------------------------
// Forward-mode dual numbers (ForwardDiff restatement)
// ---------------------------------------------------------------------------
template <class T, int N>
struct Dual {
  T v;
  T d[N];
  Dual() : v(0.0) { for (int k = 0; k < N; ++k) d[k] = T(0.0); }
  Dual(double c) : v(c) { for (int k = 0; k < N; ++k) d[k] = T(0.0); }
  explicit Dual(const T& c, int) : v(c) { for (int k = 0; k < N; ++k) d[k] = T(0.0); }
};

inline double val(double x) { return x; }
template <class T, int N>
inline double val(const Dual<T, N>& x) { return val(x.v); }

template <class T, int N>
inline Dual<T, N> lift(const T& c) { return Dual<T, N>(c, 0); }

// lift_to<T, U>::of(u): u (of a scalar type U at some nesting level of T) as a
// T with zero partials at every level above U.
template <class T, class U> struct lift_to;
template <class T> struct lift_to<T, T> {
  static T of(const T& u) { return u; }
};
template <class V, int N, class U> struct lift_to<Dual<V, N>, U> {
  static Dual<V, N> of(const U& u) { return Dual<V, N>(lift_to<V, U>::of(u), 0); }
};
template <class V, int N> struct lift_to<Dual<V, N>, Dual<V, N>> {
  static Dual<V, N> of(const Dual<V, N>& u) { return u; }
};

// A material constant inside an expression of type T: a double stays a
// double (so the double instantiations keep their mixed double operations);
// a differentiated constant (the parameter tangents of the time-loop duals
// below) is lifted to T.
template <class T> inline const double& K(const double& p) { return p; }
template <class T, class U> inline T K(const U& p) { return lift_to<T, U>::of(p); }

#define DUAL_T template <class T, int N>
DUAL_T inline Dual<T, N> operator+(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r(a.v + b.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator-(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r(a.v - b.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator-(const Dual<T, N>& a) {
  Dual<T, N> r(-a.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r(a.v * b.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
  // ForwardDiff: value a/b, partials (a' * b - a * b') / b^2
  Dual<T, N> r(a.v / b.v, 0);
  T b2 = b.v * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / b2;
  return r;
}
// mixed with double
DUAL_T inline Dual<T, N> operator+(const Dual<T, N>& a, double b) {
  Dual<T, N> r = a; r.v = a.v + b; return r;
}
DUAL_T inline Dual<T, N> operator+(double a, const Dual<T, N>& b) {
  Dual<T, N> r = b; r.v = a + b.v; return r;
}
DUAL_T inline Dual<T, N> operator-(const Dual<T, N>& a, double b) {
  Dual<T, N> r = a; r.v = a.v - b; return r;
}
DUAL_T inline Dual<T, N> operator-(double a, const Dual<T, N>& b) {
  Dual<T, N> r(a - b.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator*(const Dual<T, N>& a, double b) {
  Dual<T, N> r(a.v * b, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
DUAL_T inline Dual<T, N> operator*(double a, const Dual<T, N>& b) {
  Dual<T, N> r(a * b.v, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
DUAL_T inline Dual<T, N> operator/(const Dual<T, N>& a, double b) {
  Dual<T, N> r(a.v / b, 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
  return r;
}
DUAL_T inline Dual<T, N> operator/(double a, const Dual<T, N>& b) {
  Dual<T, N> r(a / b.v, 0);
  T b2 = b.v * b.v;
  for (int k = 0; k < N; ++k) r.d[k] = -(a * b.d[k]) / b2;
  return r;
}
DUAL_T inline bool operator<(const Dual<T, N>& a, double b) { return val(a) < b; }
DUAL_T inline Dual<T, N>& operator+=(Dual<T, N>& a, const Dual<T, N>& b) { a = a + b; return a; }

inline double sqrt_(double x) { return std::sqrt(x); }
inline double exp_(double x) { return std::exp(x); }
inline double log_(double x) { return std::log(x); }
inline double pow_(double x, double p) { return std::pow(x, p); }
inline double abs_(double x) { return std::fabs(x); }

DUAL_T inline Dual<T, N> sqrt_(const Dual<T, N>& a) {
  T s = sqrt_(a.v);
  Dual<T, N> r(s, 0);
  T den = 2.0 * s;
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / den;
  return r;
}
DUAL_T inline Dual<T, N> exp_(const Dual<T, N>& a) {
  T e = exp_(a.v);
  Dual<T, N> r(e, 0);
  for (int k = 0; k < N; ++k) r.d[k] = e * a.d[k];
  return r;
}
DUAL_T inline Dual<T, N> log_(const Dual<T, N>& a) {
  Dual<T, N> r(log_(a.v), 0);
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / a.v;
  return r;
}
DUAL_T inline Dual<T, N> pow_(const Dual<T, N>& a, double p) {
  // x^p for a Float64 exponent: value v^p, partials p * v^(p-1) * dx
  Dual<T, N> r(pow_(a.v, p), 0);
  T dp = p * pow_(a.v, p - 1.0);
  for (int k = 0; k < N; ++k) r.d[k] = dp * a.d[k];
  return r;
}
DUAL_T inline Dual<T, N> abs_(const Dual<T, N>& a) { return val(a) < 0.0 ? -a : a; }

template <class T> inline T sq(const T& x) { return x * x; }  // Base.literal_pow(^, x, Val(2))

// ---------------------------------------------------------------------------
// SimpleLA.jl:51-96 and the 3x3 generic matmul order
// Matrices are column-major 9-vectors: m(i,j) = m[i + 3j] (0-based).
// ---------------------------------------------------------------------------
#define M(a, i, j) (a)[(i) + 3 * (j)]

template <class T>
inline void matmul3(const T* a, const T* b, T* c) {
  // LinearAlgebra matmul3x3: c(i,j) = a(i,1)b(1,j) + a(i,2)b(2,j) + a(i,3)b(3,j)
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      M(c, i, j) = M(a, i, 0) * M(b, 0, j) + M(a, i, 1) * M(b, 1, j) + M(a, i, 2) * M(b, 2, j);
}

template <class T>
inline T det_(const T* m) {  // SimpleLA.jl:84-89
  return M(m, 0, 0) * (M(m, 1, 1) * M(m, 2, 2) - M(m, 2, 1) * M(m, 1, 2)) -
         M(m, 0, 1) * (M(m, 1, 0) * M(m, 2, 2) - M(m, 1, 2) * M(m, 2, 0)) +
         M(m, 0, 2) * (M(m, 1, 0) * M(m, 2, 1) - M(m, 1, 1) * M(m, 2, 0));
}

template <class T>
inline void inv_(const T* m, T* r) {  // SimpleLA.jl:51-79
  T dt = det_(m);
  T invdet = 1.0 / dt;
  M(r, 0, 0) = (M(m, 1, 1) * M(m, 2, 2) - M(m, 2, 1) * M(m, 1, 2)) * invdet;
  M(r, 0, 1) = (M(m, 0, 2) * M(m, 2, 1) - M(m, 0, 1) * M(m, 2, 2)) * invdet;
  M(r, 0, 2) = (M(m, 0, 1) * M(m, 1, 2) - M(m, 0, 2) * M(m, 1, 1)) * invdet;
  M(r, 1, 0) = (M(m, 1, 2) * M(m, 2, 0) - M(m, 1, 0) * M(m, 2, 2)) * invdet;
  M(r, 1, 1) = (M(m, 0, 0) * M(m, 2, 2) - M(m, 0, 2) * M(m, 2, 0)) * invdet;
  M(r, 1, 2) = (M(m, 1, 0) * M(m, 0, 2) - M(m, 0, 0) * M(m, 1, 2)) * invdet;
  M(r, 2, 0) = (M(m, 1, 0) * M(m, 2, 1) - M(m, 2, 0) * M(m, 1, 1)) * invdet;
  M(r, 2, 1) = (M(m, 2, 0) * M(m, 0, 1) - M(m, 0, 0) * M(m, 2, 1)) * invdet;
  M(r, 2, 2) = (M(m, 0, 0) * M(m, 1, 1) - M(m, 1, 0) * M(m, 0, 1)) * invdet;
}

template <class T>
inline T tr_(const T* m) { return M(m, 0, 0) + M(m, 1, 1) + M(m, 2, 2); }  // SimpleLA.jl:94-96

// LinearAlgebra.det on a 3x3: LU factorisation with partial pivoting.
template <class T>
inline T det_lu(const T* m) {
  T a[9];
  for (int k = 0; k < 9; ++k) a[k] = m[k];
  double sign = 1.0;
  for (int j = 0; j < 3; ++j) {
    int p = j;
    double best = abs_(val(M(a, j, j)));
    for (int i = j + 1; i < 3; ++i)
      if (abs_(val(M(a, i, j))) > best) { best = abs_(val(M(a, i, j))); p = i; }
    if (p != j) {
      for (int c = 0; c < 3; ++c) std::swap(M(a, j, c), M(a, p, c));
      sign = -sign;
    }
    if (val(M(a, j, j)) != 0.0) {
      for (int i = j + 1; i < 3; ++i) {
        T l = M(a, i, j) / M(a, j, j);
        M(a, i, j) = l;
        for (int c = j + 1; c < 3; ++c) M(a, i, c) = M(a, i, c) - l * M(a, j, c);
      }
    }
  }
  T d = M(a, 0, 0) * M(a, 1, 1) * M(a, 2, 2);
  return sign < 0 ? -d : d;
}

// ---------------------------------------------------------------------------
// Strains.jl:26-52
// ---------------------------------------------------------------------------
template <class T>
inline void finger(const T* a, T* g) {  // g = vec(inv_(A * A^T))
  T at[9], aat[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M(at, i, j) = M(a, j, i);
  matmul3(a, at, aat);
  inv_(aat, g);
}

template <class T>
inline void invariants(const T* g, T* inv) {
  T g2[9];
  matmul3(g, g, g2);
  inv[0] = tr_(g);
  inv[1] = 0.5 * (sq(tr_(g)) - tr_(g2));
  inv[2] = det_(g);
}

// ---------------------------------------------------------------------------
// EquationsOfState.jl:71-156  (Barton2009)
// ---------------------------------------------------------------------------
// P: the type of the two differentiable constants (double, or a dual carrying
// their tangents); rho0, cv, t0 and the exponents are plain numbers.
template <class P>
struct BartonT {
  double rho0;
  P c0;
  double cv, t0;
  P b0;
  double alpha, beta, gamma;
  P b0sq, k0;
};
using Barton2009 = BartonT<double>;

inline Barton2009 make_barton(const double* p) {  // EquationsOfState.jl:90-115
  Barton2009 e;
  e.rho0 = p[0]; e.c0 = p[1]; e.cv = p[2]; e.t0 = p[3];
  e.b0 = p[4]; e.alpha = p[5]; e.beta = p[6]; e.gamma = p[7];
  e.b0sq = e.b0 * e.b0;
  e.k0 = e.c0 * e.c0 - (4.0 / 3.0) * (e.b0 * e.b0);
  return e;
}

template <class T, class E>
inline T energy(const E& e, const T& S, const T* G) {  // :118-137
  T i[3];
  invariants(G, i);
  T U = 0.5 * K<T>(e.k0) / (e.alpha * e.alpha) * sq(pow_(i[2], 0.5 * e.alpha) - 1.0) +
        e.cv * e.t0 * pow_(i[2], 0.5 * e.gamma) * (exp_(S / e.cv) - 1.0);
  T W = 0.5 * K<T>(e.b0sq) * pow_(i[2], 0.5 * e.beta) * (sq(i[0]) / 3.0 - i[1]);
  return U + W;
}

template <class T, class E>
inline T entropy(const E& e, const T& e_int, const T* G) {  // :139-156
  T i[3];
  invariants(G, i);
  T S = e_int - 0.5 * K<T>(e.b0sq) * pow_(i[2], 0.5 * e.beta) * (sq(i[0]) / 3.0 - i[1]) -
        0.5 * K<T>(e.k0) / (e.alpha * e.alpha) * sq(pow_(i[2], 0.5 * e.alpha) - 1.0);
  S = (S / (e.cv * e.t0 * pow_(i[2], 0.5 * e.gamma)) + 1.0);
  if (S < 1e-6) S = T(1e-6);
  return log_(S) * e.cv;
}

// dedG = gradient(G -> energy(eos, S, G), G) at fixed S, 9 independent entries
template <class T, class E>
inline void dedG_grad(const E& e, const T& S, const T* G, T* dedG) {
  typedef Dual<T, 9> D;
  D Gd[9];
  for (int k = 0; k < 9; ++k) {
    Gd[k] = lift<T, 9>(G[k]);
    Gd[k].d[k] = T(1.0);
  }
  D en = energy<D>(e, lift<T, 9>(S), Gd);
  for (int k = 0; k < 9; ++k) dedG[k] = en.d[k];
}

// stress(eos, ent, F) :179-190
// (ent is held fixed: a double, or a scalar of a lower nesting level of T)
template <class T, class E, class S>
inline void stress_ent(const E& e, const S& ent, const T* F, T* sig) {
  T den = e.rho0 / det_lu(F);
  T G[9], dedG[9], mG[9];
  finger(F, G);
  dedG_grad<T>(e, lift_to<T, S>::of(ent), G, dedG);
  for (int k = 0; k < 9; ++k) mG[k] = (-2.0 * den) * G[k];
  matmul3(mG, dedG, sig);
}

// stress(eos, den, e_int, F) :162-176
template <class T, class E>
inline void stress_den(const E& e, const T& den, const T& e_int, const T* F, T* sig) {
  T G[9], dedG[9], mG[9];
  finger(F, G);
  T S = entropy(e, e_int, G);
  dedG_grad<T>(e, S, G, dedG);
  for (int k = 0; k < 9; ++k) mG[k] = (-2.0 * den) * G[k];
  matmul3(mG, dedG, sig);
}

// acoustic(eos, ent, F, n) :223-246 ; jacobian of stress w.r.t. F (nested duals)
template <class T, class E>
inline void acoustic_ent(const E& e, const T& ent, const T* F, const double* n, T* ac) {
  typedef Dual<T, 9> D;
  D Fd[9], sd[9];
  for (int k = 0; k < 9; ++k) {
    Fd[k] = lift<T, 9>(F[k]);
    Fd[k].d[k] = T(1.0);
  }
  stress_ent<D>(e, ent, Fd, sd);
  T den = e.rho0 / det_lu(F);
  T inv_den = 1.0 / den;
  // dTdF[m,i,j,l] = d sig(m,i) / d F(j,l) = sd[m + 3i].d[j + 3l]
  for (int k = 0; k < 9; ++k) ac[k] = T(0.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m)
            M(ac, i, j) += inv_den * sd[m + 3 * i].d[j + 3 * l] * M(F, k, l) * n[m] * n[k];
}

// acoustic(eos, P, n) :192-221 (den = P[1]*P[2], stress from e_int)
inline void acoustic_prim(const Barton2009& e, const double* P, const double* n, double* ac) {
  typedef Dual<double, 9> D;
  double den = P[0] * P[1];
  double ent = P[5];
  const double* F = P + 6;
  double G[9];
  finger(F, G);
  double e_int = energy(e, ent, G);
  D Fd[9], sd[9];
  for (int k = 0; k < 9; ++k) {
    Fd[k] = D(F[k]);
    Fd[k].d[k] = 1.0;
  }
  stress_den<D>(e, D(den), D(e_int), Fd, sd);
  double inv_den = 1.0 / den;
  for (int k = 0; k < 9; ++k) ac[k] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m)
            M(ac, i, j) += inv_den * sd[m + 3 * i].d[j + 3 * l] * M(F, k, l) * n[m] * n[k];
}

// Eigenvalues of a 3x3 whose antisymmetric part is round-off: cyclic Jacobi on
// the symmetric part.  V (optional) receives the accumulated rotation, whose
// column i is the unit eigenvector of w[i].
inline void eigvals3(const double* a, double* w, double (*V)[3] = nullptr) {
  if (V)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  double s[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s[i][j] = 0.5 * (M(a, i, j) + M(a, j, i));
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = std::fabs(s[0][1]) + std::fabs(s[0][2]) + std::fabs(s[1][2]);
    double scale = std::fabs(s[0][0]) + std::fabs(s[1][1]) + std::fabs(s[2][2]);
    if (off <= 1e-300 || off <= 1e-18 * scale) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (s[p][q] == 0.0) continue;
        double theta = (s[q][q] - s[p][p]) / (2.0 * s[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 3; ++k) {  // columns
          double skp = s[k][p], skq = s[k][q];
          s[k][p] = c * skp - sn * skq;
          s[k][q] = sn * skp + c * skq;
        }
        for (int k = 0; k < 3; ++k) {  // rows
          double spk = s[p][k], sqk = s[q][k];
          s[p][k] = c * spk - sn * sqk;
          s[q][k] = sn * spk + c * sqk;
        }
        if (V)
          for (int k = 0; k < 3; ++k) {  // V <- V J (the column rotation)
            double vkp = V[k][p], vkq = V[k][q];
            V[k][p] = c * vkp - sn * vkq;
            V[k][q] = sn * vkp + c * vkq;
          }
      }
  }
  w[0] = s[0][0]; w[1] = s[1][1]; w[2] = s[2][2];
}

// Eigenvalues of a matrix of first-order duals: the values as above, the
// partials of each simple eigenvalue from first-order perturbation theory,
// d(w_i) = v_i^T dA v_i with v_i its unit eigenvector (the derivative that
// forward-mode differentiation through a converged symmetric eigensolver
// yields).
template <int N>
inline void eigvals3(const Dual<double, N>* a, Dual<double, N>* w) {
  double av[9], wv[3], V[3][3];
  for (int k = 0; k < 9; ++k) av[k] = a[k].v;
  eigvals3(av, wv, V);
  for (int i = 0; i < 3; ++i) {
    w[i] = Dual<double, N>(wv[i], 0);
    for (int n = 0; n < N; ++n) {
      double d = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) d += V[j][i] * V[k][i] * M(a, j, k).d[n];
      w[i].d[n] = d;
    }
  }
}

// ---------------------------------------------------------------------------
// HyperelasticityMPh.jl:63-266 (per phase, 15 components)
// ---------------------------------------------------------------------------
template <class T, class E>
inline void prim2cons_phase(const E& e, const T* P, T* Q) {  // :66-87
  T frac = P[0], true_den = P[1];
  T den = frac * true_den;
  const T* vel = P + 2;
  T ent = P[5];
  const T* def_grad = P + 6;
  T G[9];
  finger(def_grad, G);
  T e_int = energy(e, ent, G);
  T e_kin = (sq(vel[0]) + sq(vel[1]) + sq(vel[2])) / 2.0;
  T e_total = e_int + e_kin;
  Q[0] = frac;
  Q[1] = den;
  for (int k = 0; k < 3; ++k) Q[2 + k] = den * vel[k];
  Q[5] = den * e_total;
  for (int k = 0; k < 9; ++k) Q[6 + k] = den * def_grad[k];
}

template <class T>
struct PhasePrimsT {
  T frac, true_den, den, vel[3], e_total, e_int, def_grad[9];
};
using PhasePrims = PhasePrimsT<double>;

template <class T, class E>
inline void prims_phase(const E& e, const T* Q, PhasePrimsT<T>& p) {  // :106-121
  p.frac = Q[0];
  T FQ[9];
  for (int k = 0; k < 9; ++k) FQ[k] = Q[6 + k] / p.frac;
  p.true_den = sqrt_(det_lu(FQ) / e.rho0);
  p.den = p.frac * p.true_den;
  for (int k = 0; k < 3; ++k) p.vel[k] = Q[2 + k] / p.den;
  p.e_total = Q[5] / p.den;
  T e_kin = (sq(p.vel[0]) + sq(p.vel[1]) + sq(p.vel[2])) / 2.0;
  p.e_int = p.e_total - e_kin;
  for (int k = 0; k < 9; ++k) p.def_grad[k] = Q[6 + k] / p.den;
}

template <class T, class E>
inline void cons2prim_phase(const E& e, const T* Q, T* P) {  // :106-133
  PhasePrimsT<T> p;
  prims_phase(e, Q, p);
  T G[9];
  finger(p.def_grad, G);
  T ent = entropy(e, p.e_int, G);
  P[0] = p.frac;
  P[1] = p.true_den;
  for (int k = 0; k < 3; ++k) P[2 + k] = p.vel[k];
  P[5] = ent;
  for (int k = 0; k < 9; ++k) P[6 + k] = p.def_grad[k];
}

template <class T, class E>
inline void flux_phase(const E& e, const T* Q, T* F) {  // :146-175
  PhasePrimsT<T> p;
  prims_phase(e, Q, p);
  T G[9], sig[9], strs[9];
  finger(p.def_grad, G);
  T ent = entropy(e, p.e_int, G);
  stress_ent<T>(e, ent, p.def_grad, sig);
  for (int k = 0; k < 9; ++k) strs[k] = p.frac * sig[k];
  const T* v = p.vel;
  const T* A = p.def_grad;
  F[0] = T(0.0);
  F[1] = p.den * v[0];
  for (int k = 0; k < 3; ++k) F[2 + k] = p.den * v[0] * v[k] - strs[3 * k];  // strs[1:3:end]
  F[5] = p.den * v[0] * p.e_total - (v[0] * strs[0] + v[1] * strs[3] + v[2] * strs[6]);
  // den .* (vel[1] .* A - vec(vel * transpose(A[1:3:end])))
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      T outer = v[r] * A[3 * c];  // vel[r] * A(1,c)
      F[6 + r + 3 * c] = p.den * (v[0] * A[r + 3 * c] - outer);
    }
}

// get_eigvals(eos::T, Q, n) :258-266 -> [u.n + c_k (k=1..3), u.n - c_k]
template <class T, class E>
inline void eigvals_phase(const E& e, const T* Q, const double* n, T* out) {
  T P[15];
  cons2prim_phase(e, Q, P);
  T ac[9], w[3];
  acoustic_ent(e, P[5], P + 6, n, ac);
  eigvals3(ac, w);
  T spd = P[2] * n[0] + P[3] * n[1] + P[4] * n[2];
  for (int k = 0; k < 3; ++k) {
    T c = sqrt_(abs_(w[k]));
    out[k] = spd + c;
    out[3 + k] = spd - c;
  }
}

// noncons_flux(eos, Q) :178-250 for two phases -> dense 30x30 (column-major)
inline void noncons_flux2(const Barton2009* eos, const double* Q, double* B) {
  const int nph = 2;
  PhasePrims p[2];
  double G[2][9], ent[2], strs[2][9], temp[2];
  for (int q = 0; q < nph; ++q) {
    prims_phase(eos[q], Q + 15 * q, p[q]);
    finger(p[q].def_grad, G[q]);
    ent[q] = entropy(eos[q], p[q].e_int, G[q]);
    double sig[9];
    stress_ent<double>(eos[q], ent[q], p[q].def_grad, sig);
    for (int k = 0; k < 9; ++k) strs[q][k] = p[q].frac * sig[k];
    // temp = derivative(S -> energy(eos, S, G), ent)
    typedef Dual<double, 1> D1;
    D1 Sd(ent[q]);
    Sd.d[0] = 1.0;
    D1 Gd[9];
    for (int k = 0; k < 9; ++k) Gd[k] = D1(G[q][k]);
    temp[q] = energy<D1>(eos[q], Sd, Gd).d[0];
  }
  const double omega = 0.0;
  const double kk[2] = {0.5, 1.0 - 0.5};
  const double beta[2] = {0.0, 0.0};
  double vel_i[3];
  for (int k = 0; k < 3; ++k) vel_i[k] = kk[0] * p[0].vel[k] + kk[1] * p[1].vel[k];
  double K[2][9];
  for (int q = 0; q < nph; ++q) {
    double trs = tr_(strs[q]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double idm = (i == j) ? 1.0 : 0.0;
        double inner = omega / 3.0 * trs * idm + (1.0 - omega) * M(strs[q], i, j);
        M(K[q], i, j) = 1.0 / p[q].frac * inner + beta[q] * idm;
      }
  }
  double strs_i[9];
  double dn = kk[0] * temp[0] + kk[1] * temp[1];
  for (int k = 0; k < 9; ++k) strs_i[k] = (kk[1] * temp[1] * K[0][k] + kk[0] * temp[0] * K[1][k]) / dn;

  std::memset(B, 0, sizeof(double) * 900);
  for (int q = 0; q < nph; ++q) {
    double blk[15];
    for (int k = 0; k < 15; ++k) blk[k] = 0.0;
    blk[0] = vel_i[0];
    for (int k = 0; k < 3; ++k) blk[2 + k] = M(strs_i, k, 0);
    blk[5] = M(strs_i, 0, 0) * vel_i[0] + M(strs_i, 1, 0) * vel_i[1] + M(strs_i, 2, 0) * vel_i[2];
    for (int i = 0; i < 9; i += 3)
      for (int k = 0; k < 3; ++k)
        blk[6 + i + k] = omega * p[q].true_den / 3.0 * (vel_i[0] - p[q].vel[0]) * p[q].def_grad[i + k] +
                         p[q].true_den * p[q].def_grad[i] * p[q].vel[k];
    // B[7:3:15, 1] += (1-omega)*true_den .* transpose(reshape(A)) * (vel_i - vel)
    double dv[3];
    for (int k = 0; k < 3; ++k) dv[k] = vel_i[k] - p[q].vel[k];
    for (int c = 0; c < 3; ++c) {
      const double* A = p[q].def_grad;
      double atdv = A[0 + 3 * c] * dv[0] + A[1 + 3 * c] * dv[1] + A[2 + 3 * c] * dv[2];
      blk[6 + 3 * c] += (1.0 - omega) * p[q].true_den * atdv;
    }
    int off = 15 * q;
    for (int r = 0; r < 15; ++r) B[(off + r) + 30 * (off + 0)] = blk[r];
  }
}

// ---------------------------------------------------------------------------
// Quadrature on [0,1] (FastGaussQuadrature, mapped as NumFluxes.jl:37,95)
// ---------------------------------------------------------------------------
static const double GL6_X[6] = {-0.93246951420315202781, -0.66120938646626451366,
                                -0.23861918608319690863, 0.23861918608319690863,
                                0.66120938646626451366,  0.93246951420315202781};
static const double GL6_W[6] = {0.17132449237917034504, 0.36076157304813860757,
                                0.46791393457269104739, 0.46791393457269104739,
                                0.36076157304813860757, 0.17132449237917034504};
static const double GLO6_X[6] = {-1.0, -0.76505532392946469285, -0.28523151648064509632,
                                 0.28523151648064509632, 0.76505532392946469285, 1.0};
static const double GLO6_W[6] = {0.066666666666666666667, 0.37847495629784698032,
                                 0.55485837703548635301,  0.55485837703548635301,
                                 0.37847495629784698032,  0.066666666666666666667};

template <class P>
struct ModelT {
  int nph;             // 1 (14-component single phase) or 2 (30 components)
  BartonT<P> eos[2];
};
using Model = ModelT<double>;

// Bint over the straight path Q_a -> Q_b with the given rule (two phases only)
inline void B_int2(const Model& m, const double* Qa, const double* Qb, const double* X,
                   const double* Wt, double* out) {
  double acc[30];
  for (int r = 0; r < 30; ++r) acc[r] = 0.0;
  double dv[30];
  for (int r = 0; r < 30; ++r) dv[r] = -Qa[r] + Qb[r];  // derivative of the path
  static thread_local std::vector<double> Bm(900);
  for (int q = 0; q < 6; ++q) {
    double s = (X[q] + 1.0) / 2.0;
    double w = Wt[q] / 2.0;
    double u[30];
    for (int r = 0; r < 30; ++r) u[r] = Qa[r] * (1.0 - s) + Qb[r] * s;
    noncons_flux2(m.eos, u, Bm.data());
    for (int r = 0; r < 30; ++r) {
      double sum = 0.0;
      for (int c = 0; c < 30; ++c) sum += (w * Bm[r + 30 * c]) * dv[c];
      acc[r] = (q == 0) ? sum : acc[r] + sum;
    }
  }
  for (int r = 0; r < 30; ++r) out[r] = acc[r];
}

// Phase-block helpers on the model's state layout.  The single-phase state is
// embedded into a 15-component block with a == 1 for the per-phase functions.
template <class P>
inline int ncomp(const ModelT<P>& m) { return m.nph == 1 ? 14 : 30; }

template <class T, class P>
inline void to_blocks(const ModelT<P>& m, const T* Q, T* Qb) {
  if (m.nph == 1) {
    Qb[0] = T(1.0);
    for (int k = 0; k < 14; ++k) Qb[1 + k] = Q[k];
  } else {
    for (int k = 0; k < 30; ++k) Qb[k] = Q[k];
  }
}
template <class T, class P>
inline void from_blocks(const ModelT<P>& m, const T* Qb, T* Q) {
  if (m.nph == 1) {
    for (int k = 0; k < 14; ++k) Q[k] = Qb[1 + k];
  } else {
    for (int k = 0; k < 30; ++k) Q[k] = Qb[k];
  }
}

template <class T, class P>
inline void flux_mph(const ModelT<P>& m, const T* Qb, T* Fb) {  // :140-144
  for (int q = 0; q < m.nph; ++q) flux_phase(m.eos[q], Qb + 15 * q, Fb + 15 * q);
}

template <class T, class P>
inline void get_eigvals(const ModelT<P>& m, const T* Qb, T* out) {  // :252-256
  const double n[3] = {1.0, 0.0, 0.0};
  for (int q = 0; q < m.nph; ++q) eigvals_phase(m.eos[q], Qb + 15 * q, n, out + 6 * q);
}

// lxf(eos, Q_l, Q_r, lambda) NumFluxes.jl:25-60 ; block layout in/out
// (the two-phase path integral in double precision only: the time-loop duals
// below run the single-phase model)
template <class T, class P>
inline void lxf(const ModelT<P>& m, const T* Ql, const T* Qr, const T& lambda, T* cons, T* dm,
                T* dp) {
  int nb = 15 * m.nph;
  T Fl[30], Fr[30];
  flux_mph(m, Ql, Fl);
  flux_mph(m, Qr, Fr);
  for (int r = 0; r < nb; ++r) cons[r] = 0.5 * (Fl[r] + Fr[r]) - 0.5 * lambda * (Qr[r] - Ql[r]);
  if constexpr (std::is_same<T, double>::value) {
    if (m.nph == 2) {
      double fp[30];
      B_int2(m, Ql, Qr, GLO6_X, GLO6_W, fp);
      for (int r = 0; r < nb; ++r) { dp[r] = (1.0 / 2.0) * fp[r]; dm[r] = (1.0 / 2.0) * fp[r]; }
      return;
    }
  }
  // one phase: the path integral multiplies the volume-fraction jump, which is 0
  for (int r = 0; r < nb; ++r) { dp[r] = T(0.0); dm[r] = T(0.0); }
}

// hll(eos, Q_l, Q_r, eigvals) NumFluxes.jl:70-132 ; eig_l/eig_r have 6*nph entries
inline void hll(const Model& m, const double* Ql, const double* Qr, const double* eig_l,
                const double* eig_r, double* cons, double* dm, double* dp) {
  int nb = 15 * m.nph, ne = 6 * m.nph;
  double Qm[30], em[12];
  for (int r = 0; r < nb; ++r) Qm[r] = 0.5 * (Ql[r] + Qr[r]);
  get_eigvals(m, Qm, em);
  double mn_m = em[0], mx_m = em[0], mn_l = eig_l[0], mx_r = eig_r[0];
  for (int k = 1; k < ne; ++k) {
    mn_m = std::min(mn_m, em[k]); mx_m = std::max(mx_m, em[k]);
    mn_l = std::min(mn_l, eig_l[k]); mx_r = std::max(mx_r, eig_r[k]);
  }
  double s_l = std::min(0.0, std::min(mn_m, mn_l));
  double s_r = std::max(0.0, std::max(mx_m, mx_r));
  double Fl[30], Fr[30];
  flux_mph(m, Ql, Fl);
  flux_mph(m, Qr, Fr);
  double path_int[30], Bs[30];
  if (m.nph == 2) B_int2(m, Ql, Qr, GL6_X, GL6_W, Bs);
  else for (int r = 0; r < nb; ++r) Bs[r] = 0.0;
  for (int r = 0; r < nb; ++r) path_int[r] = Bs[r] + Fr[r] - Fl[r];
  double Qh[30] = {0};
  for (int r = 0; r < nb; ++r) Qh[r] = (Qr[r] * s_r - Ql[r] * s_l - path_int[r]) / (s_r - s_l);
  double b1[30], b2[30];
  if (m.nph == 2) {
    B_int2(m, Ql, Qh, GL6_X, GL6_W, b1);
    B_int2(m, Qh, Qr, GL6_X, GL6_W, b2);
  } else {
    for (int r = 0; r < nb; ++r) { b1[r] = 0.0; b2[r] = 0.0; }
  }
  double sd = s_r - s_l;
  for (int r = 0; r < nb; ++r) {
    double J = Fr[r] - Fl[r] + b1[r] + b2[r];
    double dq = Qr[r] - Ql[r];
    dm[r] = -s_l / sd * J + s_l * s_r / sd * dq;
    dp[r] = s_r / sd * J - s_l * s_r / sd * dq;
    cons[r] = 0.0;
  }
}

enum { FLUX_RUSANOV = 0, FLUX_HLL = 1, FLUX_LXF = 2 };

// max as Base.max / std::max order it: b if a < b, else a (by value)
template <class T> inline T max_(const T& a, const T& b) { return val(a) < val(b) ? b : a; }

// main.jl:202-241 time loop; Q is ncomp x nx column-major (component fastest).
// Stops after max_steps steps or once t >= t_end (checked before each step).
// T: double, or a first-order dual carrying tangents of the initial state and
// of the material constants through every step (Rusanov and Lax-Friedrichs,
// single phase; the dual's time t and dt_hist record values).
template <class T, class P>
inline int run(const ModelT<P>& m, int flux, int nx, T* Q, double cfl, double dx, int max_steps,
               double t_end, double* t_io, int* steps_out, double* dt_hist, int nthreads) {
  const int nb = 15 * m.nph, nc = ncomp(m), ne = 6 * m.nph;
  if constexpr (!std::is_same<T, double>::value) {
    if (m.nph != 1 || flux == FLUX_HLL) return -1;
  }
  std::vector<T> Q0((size_t)nb * nx), Q1((size_t)nb * nx), eig((size_t)ne * nx), lam(nx);
  for (int i = 0; i < nx; ++i) to_blocks(m, Q + (size_t)nc * i, Q0.data() + (size_t)nb * i);
  double t = *t_io;
  int steps = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  while (steps < max_steps && t < t_end) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
      T* ev = eig.data() + (size_t)ne * i;
      get_eigvals(m, Q0.data() + (size_t)nb * i, ev);
      T mx = T(0.0);
      for (int k = 0; k < ne; ++k) mx = max_(mx, abs_(ev[k]));
      lam[i] = mx;
    }
    T lmax = lam[0];
    for (int i = 1; i < nx; ++i) lmax = max_(lmax, lam[i]);
    T dt = cfl * dx / lmax;
    if (dt_hist) dt_hist[steps] = val(dt);
    t += val(dt);
    steps += 1;
    for (int r = 0; r < nb; ++r) {
      Q1[r] = Q0[r];
      Q1[(size_t)nb * (nx - 1) + r] = Q0[(size_t)nb * (nx - 1) + r];
    }
    T dtdx = dt / dx;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i) {
      const T* Ql = Q0.data() + (size_t)nb * (i - 1);
      const T* Qc = Q0.data() + (size_t)nb * i;
      const T* Qr = Q0.data() + (size_t)nb * (i + 1);
      T Fl[30], dml[30], NFl[30], Fr[30], NFr[30], dpr[30];
      if constexpr (std::is_same<T, double>::value) {
        if (flux == FLUX_HLL) {  // main.jl:56-59
          hll(m, Ql, Qc, eig.data() + (size_t)ne * (i - 1), eig.data() + (size_t)ne * i, Fl, dml,
              NFl);
          hll(m, Qc, Qr, eig.data() + (size_t)ne * i, eig.data() + (size_t)ne * (i + 1), Fr, NFr,
              dpr);
        }
      }
      if (flux != FLUX_HLL) {  // Rusanov, main.jl:46-53 (commented variants)
        // Rusanov: lambda_l, lambda_r = max over the face's two cells (main.jl:47-48);
        // Lax-Friedrichs: lambda_l, lambda_r = 1 / dtdx, 1 / dtdx (main.jl:50)
        T lam_l = flux == FLUX_LXF ? 1 / dtdx : max_(lam[i - 1], lam[i]);
        T lam_r = flux == FLUX_LXF ? 1 / dtdx : max_(lam[i], lam[i + 1]);
        lxf(m, Ql, Qc, lam_l, Fl, dml, NFl);
        lxf(m, Qc, Qr, lam_r, Fr, NFr, dpr);
      }
      T* out = Q1.data() + (size_t)nb * i;
      for (int r = 0; r < nb; ++r) out[r] = Qc[r] - dtdx * ((Fr[r] - Fl[r]) + (NFr[r] + NFl[r]));
    }
    std::swap(Q0, Q1);
  }
  for (int i = 0; i < nx; ++i) from_blocks(m, Q0.data() + (size_t)nb * i, Q + (size_t)nc * i);
  *t_io = t;
  *steps_out = steps;
  return 0;
}

inline Model make_model(int nph, const double* mats) {
  Model m;
  m.nph = nph;
  m.eos[0] = make_barton(mats);
  m.eos[1] = make_barton(nph == 2 ? mats + 8 : mats);
  return m;
}

}  // namespace ora

// ---------------------------------------------------------------------------
// C entry points for the tests (ctypes).  mats: nph x 8 doubles
// (rho0, c0, cv, t0, b0, alpha, beta, gamma) as Barton2009's keyword arguments.
// ---------------------------------------------------------------------------
using namespace ora;
extern "C" {

void ora_inv(const double* m, double* r) { inv_(m, r); }
double ora_det(const double* m) { return det_(m); }
double ora_det_lu(const double* m) { return det_lu(m); }
void ora_finger(const double* a, double* g) { finger(a, g); }
void ora_invariants(const double* g, double* i) { invariants(g, i); }

double ora_energy(const double* mat, double S, const double* G) {
  return energy(make_barton(mat), S, G);
}
double ora_entropy(const double* mat, double e, const double* G) {
  return entropy(make_barton(mat), e, G);
}
void ora_stress(const double* mat, double ent, const double* F, double* sig) {
  stress_ent<double>(make_barton(mat), ent, F, sig);
}
void ora_stress_den(const double* mat, double den, double e_int, const double* F, double* sig) {
  stress_den<double>(make_barton(mat), den, e_int, F, sig);
}
void ora_acoustic(const double* mat, double ent, const double* F, const double* n, double* ac) {
  acoustic_ent(make_barton(mat), ent, F, n, ac);
}
void ora_acoustic_prim(const double* mat, const double* P, const double* n, double* ac) {
  acoustic_prim(make_barton(mat), P, n, ac);
}
void ora_eigvals3(const double* a, double* w) { eigvals3(a, w); }
// eigenvalues of a + s da and their derivatives in s at s = 0 (first-order duals)
void ora_eigvals3_dual(const double* a, const double* da, double* w, double* dw) {
  Dual<double, 1> ad[9], wd[3];
  for (int k = 0; k < 9; ++k) {
    ad[k] = Dual<double, 1>(a[k]);
    ad[k].d[0] = da[k];
  }
  eigvals3(ad, wd);
  for (int i = 0; i < 3; ++i) {
    w[i] = wd[i].v;
    dw[i] = wd[i].d[0];
  }
}
double ora_temperature(const double* mat, double S, const double* G) {
  typedef Dual<double, 1> D1;
  D1 Sd(S);
  Sd.d[0] = 1.0;
  D1 Gd[9];
  for (int k = 0; k < 9; ++k) Gd[k] = D1(G[k]);
  return energy<D1>(make_barton(mat), Sd, Gd).d[0];
}

// Model-level functions; states in the model layout (14 or 30 components).
void ora_prim2cons(int nph, const double* mats, const double* P, double* Q) {
  Model m = make_model(nph, mats);
  if (nph == 1) {
    double Pb[15], Qb[15];
    Pb[0] = 1.0;
    for (int k = 0; k < 14; ++k) Pb[1 + k] = P[k];
    prim2cons_phase(m.eos[0], Pb, Qb);
    for (int k = 0; k < 14; ++k) Q[k] = Qb[1 + k];
  } else {
    for (int q = 0; q < 2; ++q) prim2cons_phase(m.eos[q], P + 15 * q, Q + 15 * q);
  }
}
void ora_cons2prim(int nph, const double* mats, const double* Q, double* P) {
  Model m = make_model(nph, mats);
  double Qb[30], Pb[30];
  to_blocks(m, Q, Qb);
  for (int q = 0; q < nph; ++q) cons2prim_phase(m.eos[q], Qb + 15 * q, Pb + 15 * q);
  from_blocks(m, Pb, P);
}
void ora_flux(int nph, const double* mats, const double* Q, double* F) {
  Model m = make_model(nph, mats);
  double Qb[30], Fb[30];
  to_blocks(m, Q, Qb);
  flux_mph(m, Qb, Fb);
  from_blocks(m, Fb, F);
}
void ora_get_eigvals(int nph, const double* mats, const double* Q, double* out) {
  Model m = make_model(nph, mats);
  double Qb[30];
  to_blocks(m, Q, Qb);
  get_eigvals(m, Qb, out);
}
void ora_noncons_flux(const double* mats, const double* Q30, double* B900) {
  Model m = make_model(2, mats);
  noncons_flux2(m.eos, Q30, B900);
}
void ora_lxf(int nph, const double* mats, const double* Ql, const double* Qr, double lambda,
             double* cons, double* dm, double* dp) {
  Model m = make_model(nph, mats);
  double a[30], b[30], c[30], d[30], e[30];
  to_blocks(m, Ql, a);
  to_blocks(m, Qr, b);
  lxf(m, a, b, lambda, c, d, e);
  from_blocks(m, c, cons);
  from_blocks(m, d, dm);
  from_blocks(m, e, dp);
}
void ora_hll(int nph, const double* mats, const double* Ql, const double* Qr, const double* eig_l,
             const double* eig_r, double* cons, double* dm, double* dp) {
  Model m = make_model(nph, mats);
  double a[30], b[30], c[30], d[30], e[30];
  to_blocks(m, Ql, a);
  to_blocks(m, Qr, b);
  hll(m, a, b, eig_l, eig_r, c, d, e);
  from_blocks(m, c, cons);
  from_blocks(m, d, dm);
  from_blocks(m, e, dp);
}
void ora_quadrature(int which, double* x, double* w) {
  for (int k = 0; k < 6; ++k) {
    x[k] = which == 0 ? GL6_X[k] : GLO6_X[k];
    w[k] = which == 0 ? GL6_W[k] : GLO6_W[k];
  }
}
int ora_run(int nph, const double* mats, int flux, int nx, double* Q, double cfl, double dx,
            int max_steps, double t_end, double* t_io, int* steps_out, double* dt_hist,
            int nthreads) {
  Model m = make_model(nph, mats);
  return run(m, flux, nx, Q, cfl, dx, max_steps, t_end, t_io, steps_out, dt_hist, nthreads);
}

// Loss of the single-phase Rusanov run from two pieces and its gradient by
// forward-mode duals through the whole time loop (ForwardDiff.gradient of the
// misfit, restated): 28 tangents, the 13 primitives u(3), vec(A), S of each
// piece (rho = rho0 / det A) and the constants b0 (shear) and c0 (bulk).
// prims: 13 x 2 column-major; cells i with (i + 1/2) / nx > x_ifc take piece 2.
// loss = sum_i (rho_i - rho_t_i)^2 + (u1_i - u_t_i)^2 with rho, u1 of cons2prim.
// Returns the number of steps run, or -1.
int ora_adjoint_dual(const double* mats, const double* prims, double x_ifc, int nx, int steps,
                     double cfl, const double* rho_t, const double* u_t, int nthreads,
                     double* loss, double* grad) {
  typedef Dual<double, 28> D;
  ModelT<D> m;
  m.nph = 1;
  BartonT<D>& e = m.eos[0];
  e.rho0 = mats[0];
  e.c0 = D(mats[1]);
  e.c0.d[27] = 1.0;
  e.cv = mats[2];
  e.t0 = mats[3];
  e.b0 = D(mats[4]);
  e.b0.d[26] = 1.0;
  e.alpha = mats[5];
  e.beta = mats[6];
  e.gamma = mats[7];
  e.b0sq = e.b0 * e.b0;                                // EquationsOfState.jl:90-115
  e.k0 = e.c0 * e.c0 - (4.0 / 3.0) * (e.b0 * e.b0);
  m.eos[1] = e;
  D Qs[2][14];
  for (int j = 0; j < 2; ++j) {
    D x[13];
    for (int k = 0; k < 13; ++k) {
      x[k] = D(prims[13 * j + k]);
      x[k].d[13 * j + k] = 1.0;
    }
    D P[15], Q[15];
    P[0] = D(1.0);
    P[1] = e.rho0 / det_(x + 3);
    for (int k = 0; k < 3; ++k) P[2 + k] = x[k];
    P[5] = x[12];
    for (int k = 0; k < 9; ++k) P[6 + k] = x[3 + k];
    prim2cons_phase(e, P, Q);
    for (int k = 0; k < 14; ++k) Qs[j][k] = Q[1 + k];
  }
  std::vector<D> Q((size_t)14 * nx);
  for (int i = 0; i < nx; ++i) {
    const int j = (i + 0.5) / nx > x_ifc ? 1 : 0;
    for (int k = 0; k < 14; ++k) Q[(size_t)14 * i + k] = Qs[j][k];
  }
  double t = 0.0;
  int done = 0;
  if (run(m, FLUX_RUSANOV, nx, Q.data(), cfl, 1.0 / nx, steps, INFINITY, &t, &done, nullptr,
          nthreads) != 0)
    return -1;
  D L = D(0.0);
  for (int i = 0; i < nx; ++i) {
    D Qb[30];
    to_blocks(m, Q.data() + (size_t)14 * i, Qb);
    PhasePrimsT<D> p;
    prims_phase(e, Qb, p);
    L = L + (sq(p.true_den - rho_t[i]) + sq(p.vel[0] - u_t[i]));
  }
  *loss = L.v;
  for (int k = 0; k < 28; ++k) grad[k] = L.d[k];
  return done;
}

}  // extern "C"
