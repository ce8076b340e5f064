// hes_step.hpp -- the per-chunk time step shared by the single-problem and
// batched kernels: a 64-lane wave holds 64 consecutive cells of one problem
// (lanes 0 and 63 are halo cells), evaluates each cell's flux once and each
// face once, updates the 62 owned cells and returns their new CFL speeds.
#pragma once
#include "hes_faces.hpp"

namespace hes {

constexpr int TILE = 62;  // owned cells per 64-lane chunk

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// A read through the constant address space, for data that a kernel only
// reads (written by earlier launches): at a wave-uniform address it loads
// into scalar registers, and no wait for it waits for the wave's vector
// memory operations.
template <class T>
__device__ __forceinline__ T cload(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// Streaming access to the state: each element is read once and written once
// per step, so it bypasses cache allocation.
#ifndef HES_NT_LOAD
#define HES_NT_LOAD 1
#endif
#ifndef HES_NT_STORE
#define HES_NT_STORE 0
#endif
__device__ __forceinline__ double ld_stream(const double* p) {
#if HES_NT_LOAD
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void st_stream(double v, double* p) {
#if HES_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Neighbour values within a wave: lane_down(v) is lane + 1's v (lane 63 keeps
// its own), lane_up(v) is lane - 1's v (lane 0 keeps its own), as
// __shfl_down / __shfl_up by 1.  By default they are the wave-wide DPP shifts
// (wave_shl:1 / wave_shr:1, plain VALU moves); HES_DPP_SHIFT=0 selects the LDS
// crossbar (ds_bpermute) instead.
#ifndef HES_DPP_SHIFT
#define HES_DPP_SHIFT 1
#endif
#if HES_DPP_SHIFT
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_down(double v) { return dpp_shift<0x130>(v); }  // wave_shl:1
__device__ __forceinline__ double lane_up(double v) { return dpp_shift<0x138>(v); }    // wave_shr:1
// The same shifts writing zero into the lane that has no source (lane 63 for
// lane_down_z, lane 0 for lane_up_z) or whose source lane is inactive, so
// they are only used where every lane of the wave is active: never under a
// lane-dependent condition (a && or ?: whose other operand varies by lane).  With no old value to keep, the shifted
// copy goes straight into a new register: one DPP move per half, where
// lane_down / lane_up first copy the value (the original stays live).  For
// values whose boundary-lane result only reaches the halo lanes' own face and
// update, which are never stored.
template <int CTRL>
__device__ __forceinline__ double dpp_shift_z(double v) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_down_z(double v) { return dpp_shift_z<0x130>(v); }
__device__ __forceinline__ double lane_up_z(double v) { return dpp_shift_z<0x138>(v); }
#else
__device__ __forceinline__ double lane_down(double v) { return __shfl_down(v, 1); }
__device__ __forceinline__ double lane_up(double v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double lane_down_z(double v) { return __shfl_down(v, 1); }
__device__ __forceinline__ double lane_up_z(double v) { return __shfl_up(v, 1); }
#endif

// Block order across the 8 XCDs: the hardware deals consecutive blocks to
// XCDs round-robin; remap so each XCD walks a contiguous stretch of cells and
// the cache lines two neighbouring tiles share stay in one L2.
__device__ __forceinline__ long long xcd_block() {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned q = nb / 8;
  if (b < 8 * q) return (long long)(b % 8) * q + b / 8;
  return b;
}

// Launches of up to this many cells alternate their direction per step: odd
// steps walk the waves from the far end, so their first loads are the lines
// the previous step stored last, which the 256 MB memory-side cache may still
// hold.  Up to 2^23 cells (a 2 GB step) that saves 4-7% per step; from 2^24
// the cache holds too small a share of a step and alternating measured 0 to
// 2% slower.  The order changes no value.
constexpr long long ALT_MAX_CELLS = 1LL << 23;

#ifndef HES_STEP_WAVES_PER_EU
#define HES_STEP_WAVES_PER_EU 3
#endif

// Keeps a value's computation where it stands (emits no instruction).  The
// updated state is consumed only in the branch of interior cells; left alone,
// the compiler sinks the update arithmetic into that branch and every face
// value and shifted neighbour value stays live across it instead of one
// component at a time.
__device__ __forceinline__ void pin(double& v) { asm volatile("" : "+v"(v)); }

// Wave priority while a wave issues its state loads, and back to the default
// for its arithmetic: a newly started wave's loads then go out ahead of the
// FP64 work of the waves already computing, which keeps more loads in flight
// (single-phase step at 2^22 cells: 205 -> 200 us; two-phase HLL at 2^21:
// about 301 -> 292 us).  Used by the solver's step only: the batched step
// was 13% slower with it (0.82 -> 0.93 ms), wherever it was raised.
__device__ __forceinline__ void prio_loads() { __builtin_amdgcn_s_setprio(2); }
__device__ __forceinline__ void prio_compute() { __builtin_amdgcn_s_setprio(0); }


// One 64-lane chunk: lane l holds cell i = first + l - 1 (lanes 0 and 63 are
// halo cells); q and cc are its state and sound speeds.  Updates the owned
// interior cells, writes their new state and sound speeds, and returns the
// lane's CFL speed of the new state (0 for lanes that own nothing).  When
// lam_cur is given it receives the CFL speed of the lane's current state.
// Where a chunk reads and writes: the cells of one problem are [lo, hi) of the
// structure-of-arrays state; its first and last cells are frozen.
struct ChunkCtx {
  const MatDev* mat;
  const double* __restrict__ Qin;
  double* __restrict__ Qout;
  const double* __restrict__ Cin;
  double* __restrict__ Cout;
  int* status;
  long long pitch, lo, hi;
  double cfl, dx;
  double relax = 0.0;  // two phases: 1 / tau of the interphase velocity relaxation (0: none)
  // two phases, Rusanov: half path integrals of the faces with a volume-fraction
  // jump, precomputed by k_path2 (face i+1/2 at [r * pitch + i]); null: inline
  const double* __restrict__ dpath = nullptr;
  // Rusanov kernels: nonzero selects the Lax-Friedrichs face speed
  // lambda = 1 / dtdx = dx / dt on every face (main.jl:30-41, 50) in place of
  // max(lambda_i, lambda_i+1); the path integrals do not depend on it
  int lxf = 0;
  double relax_p = 0.0;  // two phases: 1 / tau of the pressure relaxation (0: none)
};

// Interphase velocity relaxation of a two-phase cell over a time dt (an
// optional source; the reference has none): the drag d(m1 u1)/dt = mu (u2 - u1)
// / tau, mu = m1 m2 / (m1 + m2), with m_p = a_p rho_p as cons2prim gives it.
// Solved exactly: the mixture velocity stays, the relative velocity decays by
// exp(-dt / tau).  The kinetic energy it dissipates goes to the two phases in
// equal halves, so mass, momentum and total energy are unchanged.
__device__ __forceinline__ void relax_velocity2(const MatDev* mat, double* q, double decay) {
  double m[2], u[2][3];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double* b = q + 15 * p;
    const double ia = rcp(b[0]);
    m[p] = b[0] * sqrt(det3cm(b + 6) * (ia * ia * ia) * mat[p].irho0);
    const double im = rcp(m[p]);
#pragma unroll
    for (int k = 0; k < 3; ++k) u[p][k] = b[2 + k] * im;
  }
  const double M = m[0] + m[1], iM = rcp(M);
  const double mu = m[0] * m[1] * iM;
  double dk = 0.0, ke[2] = {0.0, 0.0}, kn[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double U = (m[0] * u[0][k] + m[1] * u[1][k]) * iM;
    const double ur = u[1][k] - u[0][k], urn = ur * decay;
    const double v0 = U - m[1] * iM * urn, v1 = U + m[0] * iM * urn;
    dk += 0.5 * mu * (ur * ur - urn * urn);
    ke[0] += u[0][k] * u[0][k];
    ke[1] += u[1][k] * u[1][k];
    kn[0] += v0 * v0;
    kn[1] += v1 * v1;
    q[2 + k] = m[0] * v0;
    q[17 + k] = m[1] * v1;
  }
  q[5] += 0.5 * m[0] * (kn[0] - ke[0]) + 0.5 * dk;
  q[20] += 0.5 * m[1] * (kn[1] - ke[1]) + 0.5 * dk;
}

// Pressure -tr(sigma)/3 of a phase: with tr sigma = -2 rho (2 W + 3 c3),
// p = rho (4/3 W + 2 c3) = rho ((4/3 + beta) W + 2 k0/(2a) (s_a - 1) s_a + gamma E_th).
__device__ __forceinline__ double phase_pressure(const MatDev& m, const Thermo& t) {
  return t.rho * ((4.0 / 3.0 + m.beta) * t.W + 2.0 * m.k0_2a * (t.sa - 1.0) * t.sa +
                  m.gamma * t.Eth);
}

// The pressure of a phase after its volume fraction changes from a to a x at
// fixed mass, velocity and distortion shape (A -> x^(1/3) A), with its
// internal energy changed by de: rho -> rho / x, I3^(k/2) -> I3^(k/2) x^-k,
// W -> W x^-(beta + 4/3) (I1^2/3 - I2 scales as x^-4/3), and E_th from the
// energy balance e = W + U_cold + E_th (with the entropy law's clamp).
struct PhaseScale {
  double rho, W, sa, sg, e;
};
template <int FAST>
__device__ __forceinline__ double scaled_pressure(const MatDev& m, const PhaseScale& s, double x,
                                                  double de) {
  const double ix = rcp(x);
  double fa, fb, fg;  // x^-alpha, x^-beta x^-4/3, x^-gamma
  if (FAST == 1 || (FAST == -1 && m.fast)) {
    const double ix2 = ix * ix;
    fa = ix;
    fb = ix2 * ix2 * cbrt_pos(ix);
    fg = ix2;
  } else {
    const double lx = log(x);
    fa = exp(-m.alpha * lx);
    fb = exp(-(m.beta + 4.0 / 3.0) * lx);
    fg = exp(-m.gamma * lx);
  }
  const double rho = s.rho * ix, sa = s.sa * fa, sg = s.sg * fg, W = s.W * fb;
  const double sam1 = sa - 1.0;
  const double eth_min = m.cvt0 * sg * (1e-6 - 1.0);  // x_S >= 1e-6 (EquationsOfState.jl:152-154)
  const double Eth = fmax(s.e + de - W - m.hk0a2 * sam1 * sam1, eth_min);
  return rho * ((4.0 / 3.0 + m.beta) * W + 2.0 * m.k0_2a * sam1 * sa + m.gamma * Eth);
}

// Pressure relaxation of a two-phase cell over a time dt (an optional stiff
// source; the reference has none), solved implicitly:
//   a1' - a1 = (dt / tau) (p1' - p2') / Z,  Z = sum_k rho_k c0_k^2 / a_k,
// with p_k = -tr(sigma_k)/3 at the end of the step.  Z is the linearised
// stiffness -d(p1 - p2)/da1, so near equilibrium p1 - p2 decays as
// exp(-t / tau); as tau -> 0 the step ends in pressure equilibrium.  Each
// phase keeps its mass, momentum and distortion shape (a rho A scales by
// (a'/a)^(1/3), so rho' = rho a / a'), and the phases exchange the work of
// the interface pressure p_I = (p1 + p2) / 2 of the start of the step:
// a_k rho_k E_k changes by -p_I (a_k' - a_k).  Mixture mass, momentum and
// energy are unchanged, a1 + a2 is kept, and a state in pressure equilibrium
// is left alone.  The residual g(d) = d - (dt/tau)/Z (p1'(d) - p2'(d))
// increases with d = a1' - a1 and changes sign on (-a1, a2) (a phase
// compressed to nothing has unbounded pressure); its root is found by secant
// steps kept inside a shrinking bracket.
template <int FAST>
__device__ __forceinline__ void relax_pressure2(const MatDev* mat, double* q, double dtau) {
  PhaseScale ps[2];
  double p[2], Z = 0.0;
  const double a[2] = {q[0], q[15]};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    Thermo t;
    thermo<false, FAST>(mat[k], a[k], q + 15 * k + 1, t);
    ps[k] = {t.rho, t.W, t.sa, t.sg, t.E - 0.5 * (t.u[0] * t.u[0] + t.u[1] * t.u[1] + t.u[2] * t.u[2])};
    p[k] = phase_pressure(mat[k], t);
    Z += t.rho * (mat[k].k0 + (4.0 / 3.0) * mat[k].b0sq) * rcp(a[k]);
  }
  const double kc = dtau * rcp(Z);
  const double pI = 0.5 * (p[0] + p[1]);
  // de_k = -p_I (a_k' - a_k) / (a_k rho_k), with a_k rho_k from the density
  const double im0 = rcp(a[0] * ps[0].rho), im1 = rcp(a[1] * ps[1].rho);
  auto g = [&](double d) {
    const double p0 = scaled_pressure<FAST>(mat[0], ps[0], (a[0] + d) * rcp(a[0]), -pI * d * im0);
    const double p1 = scaled_pressure<FAST>(mat[1], ps[1], (a[1] - d) * rcp(a[1]), pI * d * im1);
    return d - kc * (p0 - p1);
  };
  const double g0 = -kc * (p[0] - p[1]);
  if (g0 == 0.0) return;  // pressure equilibrium: nothing moves
  // bracket [lo, hi] of the root; first guess from the linearised stiffness
  double lo = g0 < 0.0 ? 0.0 : -a[0] * (1.0 - 1e-12);
  double hi = g0 < 0.0 ? a[1] * (1.0 - 1e-12) : 0.0;
  double d0 = 0.0, gd0 = g0;
  double d1 = -g0 * rcp(1.0 + kc * Z);
  d1 = fmin(fmax(d1, lo), hi);
#pragma unroll 1
  for (int it = 0; it < 60; ++it) {
    const double gd1 = g(d1);
    if (gd1 == 0.0) break;
    if (gd1 < 0.0) lo = d1;
    else hi = d1;
    double dn = d1 - gd1 * (d1 - d0) * rcp(gd1 - gd0);
    if (!(dn > lo && dn < hi)) dn = 0.5 * (lo + hi);  // outside the bracket: bisect
    const bool done = fabs(dn - d1) <= 4e-16 * fmax(a[0], a[1]) || hi - lo <= 4e-16;
    d0 = d1;
    gd0 = gd1;
    d1 = dn;
    if (done) break;
  }
  const double d = d1;
  const double a0n = a[0] + d, a1n = a[1] - d;
  const double s0 = cbrt_pos(a0n * rcp(a[0])), s1 = cbrt_pos(a1n * rcp(a[1]));
  q[0] = a0n;
  q[15] = a1n;
  q[5] -= pI * d;
  q[20] += pI * d;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    q[6 + r] *= s0;
    q[21 + r] *= s1;
  }
}

// PIN (single phase): each updated component is pinned where it is computed
// (see pin): Rusanov 102 instead of 160 VGPRs, HLL 106 instead of 136, four
// waves per SIMD instead of three.  It pays on large grids (2^22 cells:
// Rusanov 203 -> 200 us, HLL 242 -> 229 us) and costs on small ones, whose
// few waves per SIMD want the unpinned schedule's instruction-level
// parallelism (2^18 cells, Rusanov: 19.6 -> 20.3 us).
// SRC (two phases): the optional relaxation sources after the flux update
// (velocity and pressure relaxation, each on when its rate is non-zero); a
// separate instantiation, so that the plain step carries none of their code.
// NOC (one phase): the input state's sound speed is computed from its state
// rather than read (cc is not used), and the new state's is not stored: 16 B
// less traffic per cell for about 170 more VALU instructions per wave.  The
// values are the same bits either way.  It pays for the memory-bound pinned
// Rusanov step (2^22 cells: 185.6-186.1 -> 182.1-182.8 us) and the batched
// Rusanov step (4096 problems, 16.2M cells: 823.5-824.5 -> 785.3-786.8 us)
// and costs the VALU-bound HLL one (197.6-197.9 -> 206.8-207.0 us), same box.
template <int NPH, int FLUX, int FAST, bool PATHBUF = false, bool PIN = false, bool SRC = false,
          bool NOC = false>
__device__ __forceinline__ double chunk_step(const ChunkCtx& a, long long i, const double* q,
                                             const double* cc, double lmax,
                                             double* lam_cur = nullptr) {
  constexpr int NC = NComp<NPH>::value;
  const int lane = threadIdx.x & 63;
  const long long pitch = a.pitch;
  const bool inr = (i >= a.lo) && (i < a.hi);

  // HLL (one phase, or two with the path terms precomputed): the mid-state
  // eigenvalue range of face i+1/2 first, while only the state is live (its
  // value does not depend on the order)
  constexpr bool HLL_STREAM = FLUX == FLUX_HLL && (NPH == 1 || PATHBUF);
  // two phases with the path terms precomputed: the face speeds come first
  // (from u1 and the stored sound speeds), then each phase's flux is formed
  // and consumed by the update in turn, so only one phase's flux is live
  constexpr bool PHASED = NPH == 2 && PATHBUF;
  // The mid state Q_m = (Q_i + Q_i+1) / 2 enters as the neighbours' sum
  // S = 2 Q_m (0.5 S is exact) with volume fraction 2 a_m (one phase: 2):
  // thermo's density takes det(S) / (2 a_m)^3, so every power of two scales
  // out exactly and the thermodynamic state, sound speed and range are those
  // of Q_m to the bit, without the 14 or 15 halvings per phase.
  double mn_m = 0.0, mx_m = 0.0;
  if (HLL_STREAM && PHASED) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      double qs[15];
#pragma unroll
      for (int r = 0; r < 15; ++r) qs[r] = q[15 * p + r] + lane_down(q[15 * p + r]);
      Thermo t;
      thermo<false, FAST>(a.mat[p], qs[0], qs + 1, t);
      const double c = sound_speed(a.mat[p], t);
      mn_m = p == 0 ? t.u[0] - c : fmin(mn_m, t.u[0] - c);
      mx_m = p == 0 ? t.u[0] + c : fmax(mx_m, t.u[0] + c);
    }
  } else if (HLL_STREAM && NPH == 1) {
    double qs[NC];
#pragma unroll
    for (int r = 0; r < NC; ++r) qs[r] = q[r] + lane_down(q[r]);
    Thermo t;
    thermo<false, FAST>(a.mat[0], 2.0, qs, t);
    const double c = sound_speed(a.mat[0], t);
    mn_m = t.u[0] - c;
    mx_m = t.u[0] + c;
  }

  double f[NC], smin, smax;
  int bad;
  if (PHASED) {
    // [min_p(u1 - c_p), max_p(u1 + c_p)] as cell_flux forms it
    bad = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const double u1 = phase_u1<false>(a.mat[p], q[15 * p], q + 15 * p + 1);
      smin = p == 0 ? u1 - cc[p] : fmin(smin, u1 - cc[p]);
      smax = p == 0 ? u1 + cc[p] : fmax(smax, u1 + cc[p]);
    }
  } else {
    cell_flux<NPH, FAST, NOC>(a.mat, q, cc, f, smin, smax, bad);
  }

  // dt is needed only from here on: the CFL maximum is read after the state
  // loads are in flight (main.jl:212, dt/dx of main.jl:225)
  // dt / dx = cfl / lmax, to about an ulp (dt itself, which the time sums,
  // is formed exactly in the step's epilogue)
  const double dtdx = a.cfl * rcp(lmax);

  double upd[NC];
  if (FLUX == FLUX_RUSANOV && NPH == 1) {
    // Rusanov face i+1/2: 1/2 (F_i + F_i+1) - 1/2 lam (Q_i+1 - Q_i) with
    // lam = max(lam_i, lam_i+1) (NumFluxes.jl:30, main.jl:47-48), split as
    // a_i + b_i+1 with a = F/2 + lam/2 Q and b = F/2 - lam/2 Q.  Each face sum
    // is formed from the same two rounded halves by both of its cells, so the
    // two cells see bit-identical face fluxes.
    const double lam_c = fmax(fabs(smin), fabs(smax));  // maximum(abs.(eigvals))
    // lambda = 1 / dtdx (main.jl:50), to about an ulp (dt itself stays exact)
    const double hg = a.lxf ? 0.5 * rcp(dtdx) : 0.0;
    const double lam_r = lane_down_z(lam_c), lam_l = lane_up_z(lam_c);
    const double hr = a.lxf ? hg : 0.5 * fmax(lam_c, lam_r);
    const double hl = a.lxf ? hg : 0.5 * fmax(lam_l, lam_c);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      double fa = 0.5 * f[r] + hr * q[r];
      double fb = 0.5 * f[r] - hl * q[r];
      double Fr = fa + lane_down_z(fb);
      double Fl = lane_up_z(fa) + fb;
      upd[r] = q[r] - dtdx * (Fr - Fl);  // main.jl:59 with zero path integrals
      if (PIN) pin(upd[r]);
    }
  } else if (FLUX == FLUX_RUSANOV && NPH == 2 && PATHBUF) {
    // Two phases with the half path integrals precomputed (k_path2): each
    // component's face flux is formed (as face_rusanov forms it) and consumed
    // in turn, so the neighbour state and the face arrays are never all live
    // at once; the path terms are added only in waves that have a face with a
    // volume-fraction jump.
    const double lam_c = fmax(fabs(smin), fabs(smax));  // maximum(abs.(eigvals))
    const double lam_n = fmax(fabs(lane_down(smin)), fabs(lane_down(smax)));
    const double hl = a.lxf ? 0.5 * rcp(dtdx) : 0.5 * fmax(lam_c, lam_n);  // main.jl:47-50
    // the shifts run with every lane active (not under the lane-dependent
    // conditions), so each lane sees its right neighbour's volume fractions
    const double a0n = lane_down(q[0]), a1n = lane_down(q[15]);
    const bool jump = !(((a0n - q[0]) == 0.0) & ((a1n - q[15]) == 0.0)) & (i >= a.lo) &
                      (i < a.hi - 1);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      Thermo t;
      thermo<false, FAST>(a.mat[p], q[15 * p], q + 15 * p + 1, t);
      bad |= t.bad;
      double fp[15];
      fp[0] = 0.0;
      phase_flux(a.mat[p], q[15 * p], t, q + 15 * p + 6, fp + 1);
#pragma unroll
      for (int k = 0; k < 15; ++k) {
        const int r = 15 * p + k;
        // zero-filling shifts (every lane is active): only lane 63's face and
        // lane 0's update see the zeros, and neither is stored
        const double qn = lane_down_z(q[r]), fn = lane_down_z(fp[k]);
        const double cons = 0.5 * (fp[k] + fn) - hl * (qn - q[r]);
        upd[r] = q[r] - dtdx * (cons - lane_up_z(cons));  // main.jl:59, conservative part
        pin(upd[r]);
      }
    }
    if (__any(jump)) {
      // only the lanes of jump faces load, all 30 rows in one masked region
      // (so the loads issue together), and each touches only the cache lines
      // of its own face; the shifts then run with every lane active (the
      // branch is wave-uniform) and lane 0's update is never stored
      double d[NC];
#pragma unroll
      for (int r = 0; r < NC; ++r) d[r] = 0.0;
      if (jump) {
#pragma unroll
        for (int r = 0; r < NC; ++r) d[r] = a.dpath[r * a.pitch + i];
      }
#pragma unroll
      for (int r = 0; r < NC; ++r) {
        upd[r] -= dtdx * (d[r] + lane_up_z(d[r]));  // D- of face i+1/2 + D+ of face i-1/2
        pin(upd[r]);
      }
    }
  } else if (HLL_STREAM) {
    // HLL: D-, D+ without path integrals, streamed per component (for one
    // phase that is all of NumFluxes.jl:70-132; for two phases it is exact for
    // faces without a volume-fraction jump, where all three path integrals of
    // NumFluxes.jl:111-131 vanish, and faces with a jump add the difference to
    // the full D-, D+ precomputed by k_path2_hll).
    const HllCoef hc = hll_coef(mn_m, mx_m, smin, lane_down(smax));
    bool jump = false;
    if (NPH == 2) {  // shifts with every lane active, as in the Rusanov branch
      const double a0n = lane_down(q[0]), a1n = lane_down(q[NPH == 2 ? 15 : 0]);
      jump = !(((a0n - q[0]) == 0.0) & ((a1n - q[NPH == 2 ? 15 : 0]) == 0.0)) & (i >= a.lo) &
             (i < a.hi - 1);
    }
    auto stream = [&](int r, double fr) {
      // one register holds the flux for both its uses (shifted to the left
      // neighbour and local): left to itself the compiler may form the two
      // from separate copies of the flux arithmetic, contracted differently,
      // and a uniform state then moves by an ulp (F(Q) - F(Q) must be 0)
      pin(fr);
      // zero-filling shifts (every lane is active here): lane 63's face and
      // lane 0's update read the zeros, and neither is stored
      const double J = lane_down_z(fr) - fr;
      const double dq = lane_down_z(q[r]) - q[r];
      // D- + D+ = J at every face (cl + cr = 1), so D-(i+1/2) + D+(i-1/2) =
      // H(i+1/2) - H(i-1/2) with H = F_i + D-(i+1/2), the HLL flux: two
      // fused products per face instead of four, and a uniform state still
      // gives H = F_i exactly
      const double h = fma(hc.cl, J, fma(hc.clr, dq, fr));
      upd[r] = q[r] - dtdx * (h - lane_up_z(h));  // conservative part is zero (NumFluxes.jl:82)
      if (PHASED || PIN) pin(upd[r]);
    };
    if (PHASED) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        Thermo t;
        thermo<false, FAST>(a.mat[p], q[15 * p], q + 15 * p + 1, t);
        bad |= t.bad;
        double fp[15];
        fp[0] = 0.0;
        phase_flux(a.mat[p], q[15 * p], t, q + 15 * p + 6, fp + 1);
#pragma unroll
        for (int k = 0; k < 15; ++k) stream(15 * p + k, fp[k]);
      }
    } else {
#pragma unroll
      for (int r = 0; r < NC; ++r) stream(r, f[r]);
    }
    if (NPH == 2 && __any(jump)) {  // loads and shifts as in the Rusanov branch, a phase at a time
#pragma unroll
      for (int p = 0; p < NPH; ++p) {
        double dm[15], dp[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) {
          dm[k] = 0.0;
          dp[k] = 0.0;
        }
        if (jump) {
#pragma unroll
          for (int k = 0; k < 15; ++k) {
            dm[k] = a.dpath[(15 * p + k) * a.pitch + i];
            dp[k] = a.dpath[(NC + 15 * p + k) * a.pitch + i];
          }
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) {
          const int r = 15 * p + k;
          upd[r] -= dtdx * (dm[k] + lane_up_z(dp[k]));
          pin(upd[r]);
        }
      }
    }
  } else {

  // right neighbour (cell i+1) from lane+1
  double qn[NC], fn[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    qn[r] = lane_down(q[r]);
    fn[r] = lane_down(f[r]);
  }
  double smin_n = lane_down(smin), smax_n = lane_down(smax);

  // face i+1/2 (this lane) and face i-1/2 (from lane-1)
  if (FLUX == FLUX_RUSANOV) {
    double lam_c = fmax(fabs(smin), fabs(smax));       // maximum(abs.(eigvals))
    double lam_n = fmax(fabs(smin_n), fabs(smax_n));
    double cons[NC], dmp[NC];
    face_rusanov<NPH, FAST>(a.mat, q, qn, f, fn, a.lxf ? rcp(dtdx) : fmax(lam_c, lam_n), cons,
                            dmp);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      double cl = lane_up(cons[r]);
      double nfr = NPH == 2 ? dmp[r] : 0.0;
      double nfl = NPH == 2 ? lane_up(dmp[r]) : 0.0;
      upd[r] = q[r] - dtdx * ((cons[r] - cl) + (nfr + nfl));  // main.jl:59
    }
  } else {
    double dm[NC], dp[NC];
    face_hll<NPH, FAST>(a.mat, q, qn, f, fn, smin, smax_n, dm, dp);
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      double dpl = lane_up(dp[r]);
      upd[r] = q[r] - dtdx * (dm[r] + dpl);  // conservative part is zero (NumFluxes.jl:82)
    }
  }
  }

  const bool owned = inr && lane >= 1 && lane <= TILE;
  const bool interior = owned && i > a.lo && i < a.hi - 1;
  if constexpr (NPH == 2 && SRC) {
    if (interior) {
      const double dt = a.cfl * a.dx / lmax;  // main.jl:212
      if (a.relax > 0.0) relax_velocity2(a.mat, upd, exp(-dt * a.relax));
      if (a.relax_p > 0.0) relax_pressure2<FAST>(a.mat, upd, dt * a.relax_p);
    }
  }
  double lam_new = 0.0;
  int bad2 = 0;
  if (interior) {
#pragma unroll
    for (int r = 0; r < NC; ++r) st_stream(upd[r], a.Qout + r * pitch + i);
    double cn[NPH];
    lam_new = cell_speeds<NPH, FAST>(a.mat, upd, cn, bad2);
    if (!NOC) {
#pragma unroll
      for (int p = 0; p < NPH; ++p) st_stream(cn[p], a.Cout + p * pitch + i);
    }
  } else if (owned) {  // frozen end cells (main.jl:219-220)
    lam_new = fmax(fabs(smin), fabs(smax));
    if (!NOC) {
#pragma unroll
      for (int p = 0; p < NPH; ++p) st_stream(cc[p], a.Cout + p * pitch + i);
    }
  }
  if (lam_cur) *lam_cur = fmax(fabs(smin), fabs(smax));
  int anybad = (owned && bad) || bad2;
  if (__any(anybad) && lane == 0) atomicOr(a.status, 1);
  return lam_new;
}

// Whether face i+1/2 (0 <= i < n - 1) of a two-phase state has a volume-
// fraction jump: the faces whose path integrals are not zero.  The step's own
// test (chunk_step) is the same expression.
__device__ __forceinline__ bool face_jump2(const double* __restrict__ Q, long long i,
                                           long long pitch) {
  const double a0 = Q[i], a0n = Q[i + 1], a1 = Q[15 * pitch + i], a1n = Q[15 * pitch + i + 1];
  return !(((a0n - a0) == 0.0) & ((a1n - a1) == 0.0));
}

// Offset of this lane's entry in a list that every wave of a block of NT
// threads appends to: one atomic per block on *count (waves that append into
// one counter one by one serialise at its cache line), the waves' shares in
// block order, the lanes' in lane order.  Every thread of the block must call
// it.  Returns the position of a flagged lane's entry.
template <int NT>
__device__ __forceinline__ unsigned long long block_append(bool flag, unsigned long long* count) {
  __shared__ unsigned long long share[NT / 64];
  __shared__ unsigned long long base;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(flag);
  if (lane == 0) share[w] = (unsigned long long)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0ull;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
      const unsigned long long c = share[k];
      share[k] = total;
      total += c;
    }
    base = total ? atomicAdd(count, total) : 0ull;
  }
  __syncthreads();
  return base + share[w] + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
}

// One quadrature node's share of a path integral (NumFluxes.jl:39-47, 97-107):
// w b^p(psi) da_p in the rows of phase block p, where psi is the state at the
// node, b^p the non-zero column of B's block p and da_p the volume-fraction
// jump of the path.  The integral is the sum over the nodes.
// ACC: add the share to out instead of writing it.
template <int FAST, bool ACC = false>
__device__ __forceinline__ void bnode2(const MatDev* mat, const double* psi, double w, double da0,
                                       double da1, double* out) {
  BCol b;
  bcol2<FAST>(mat, psi, b);
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double wd = w * (p == 0 ? da0 : da1);
    double* o = out + 15 * p;
    auto put = [&](int r, double v) { o[r] = ACC ? o[r] + v : v; };
    put(0, wd * b.uI0);
    if (!ACC) o[1] = 0.0;
    put(2, wd * b.sI[0]);
    put(3, wd * b.sI[1]);
    put(4, wd * b.sI[2]);
    put(5, wd * b.sIu);
#pragma unroll
    for (int r = 0; r < 9; ++r) put(6 + r, wd * b.a[p][r]);
  }
}

// Sum of v[30] over the 16 lanes of a lane group (every lane of the group active).
__device__ __forceinline__ void group16_sum(double* v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1)
#pragma unroll
    for (int r = 0; r < 30; ++r) v[r] += __shfl_xor(v[r], o, 16);
}

// v[g] and v[g + 15] for lane g of a group (g < 15), by selects: indexing a
// register array with a lane-dependent index would move it to scratch memory.
__device__ __forceinline__ void rows_of_lane(const double* v, int g, double& lo, double& hi) {
  lo = 0.0;
  hi = 0.0;
#pragma unroll
  for (int r = 0; r < 15; ++r) {
    lo = g == r ? v[r] : lo;
    hi = g == r ? v[r + 15] : hi;
  }
}

// The jump faces' path terms are computed by a group of lanes per face (lane
// g of the group), the quadrature nodes spread over the group's lanes: the
// nodes evaluate the same code on different data, which a wave runs in
// parallel for all its faces.

// Number of lanes per jump face in the Rusanov path-term kernels.
#ifndef HES_PATH_LANES
#define HES_PATH_LANES 2
#endif

// Rusanov: half path integral 1/2 int B dpsi of the jump face i+1/2
// (0 <= i < n - 1) of a two-phase state (D- = D+ of NumFluxes.jl:35-59): the
// 6 Gauss-Lobatto nodes over the group's GL lanes (node m = g + GL round), a
// group sum, and lane g stores rows g, g + GL, ...; only jump faces are
// written (the step reads only those).
template <int FAST, int GL = HES_PATH_LANES>
__device__ __forceinline__ void face_path2(const MatDev* mat, const double* __restrict__ Q,
                                           double* __restrict__ D, long long i, long long pitch,
                                           int g) {
  double d[30];
#pragma unroll
  for (int r = 0; r < 30; ++r) d[r] = 0.0;
#pragma unroll 1
  for (int round = 0; round * GL < 6; ++round) {
    const int m = g + GL * round;
    if (m < 6) {
      // the two cell states are read in every round (from L2) rather than
      // held across the rounds: an offset the compiler cannot see through
      // keeps it from hoisting the loads out of the loop
      int ofs = 0;
      asm volatile("" : "+v"(ofs));
      const double* __restrict__ ql = Q + i + ofs;
      const double s = c_lo_s[m], oms = 1.0 - s;
      double psi[30];
#pragma unroll
      for (int r = 0; r < 30; ++r) psi[r] = ql[r * pitch] * oms + ql[r * pitch + 1] * s;
      bnode2<FAST, true>(mat, psi, c_lo_w[m], ql[1] - ql[0], ql[15 * pitch + 1] - ql[15 * pitch],
                         d);
    }
  }
#pragma unroll
  for (int o = GL / 2; o > 0; o >>= 1)
#pragma unroll
    for (int r = 0; r < 30; ++r) d[r] += __shfl_xor(d[r], o, GL);
#pragma unroll
  for (int k = 0; k * GL < 30; ++k) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < GL; ++j)
      if (GL * k + j < 30) v = g == j ? d[GL * k + j] : v;
    const int r = GL * k + g;
    if (r < 30) D[r * pitch + i] = 0.5 * v;
  }
}

// Sum of v[30] over the GL lanes of a lane group (every lane of the group active).
template <int GL>
__device__ __forceinline__ void group_sum(double* v) {
#pragma unroll
  for (int o = GL / 2; o > 0; o >>= 1)
#pragma unroll
    for (int r = 0; r < 30; ++r) v[r] += __shfl_xor(v[r], o, GL);
}

// Number of lanes per jump face in the HLL path-term kernels.
#ifndef HES_PATH_HLL_LANES
#define HES_PATH_HLL_LANES 4
#endif

// HLL: for the jump face i+1/2, D-, D+ (NumFluxes.jl:70-132) minus their
// path-free part cl J + clr dQ, cr J - clr dQ, i.e. cl and cr times the two
// half-path integrals through the HLL state (rows 0..29: D-, rows 30..59: D+
// of face i+1/2).  C holds the stored sound speeds (2 rows), as the step uses
// them.  A group of GL lanes per face (lane g), 64 / GL faces per wave: the
// Gauss nodes of Q_l -> Q_r over the group's lanes (in rounds of GL); lanes 0
// and 1 the two cells' fluxes, then one phase each of the mid-state
// eigenvalue range; then the 12 nodes of Q_l -> Q_hll and Q_hll -> Q_r, again
// in rounds.  A wave runs each stage for all its faces at once, so with few
// lanes per face the stages that occupy only two lanes of a group cost
// little, and the node rounds keep every lane busy.
template <int FAST, int GL = HES_PATH_HLL_LANES>
__device__ __forceinline__ void face_path2_hll(const MatDev* mat, const double* __restrict__ Q,
                                               const double* __restrict__ C,
                                               double* __restrict__ D, long long i,
                                               long long pitch, int g) {
  static_assert(GL >= 2 && GL <= 16 && (GL & (GL - 1)) == 0, "lanes per face");
  double q[30], qn[30], bs[30];
#pragma unroll
  for (int r = 0; r < 30; ++r) {
    q[r] = Q[r * pitch + i];
    qn[r] = Q[r * pitch + i + 1];
    bs[r] = 0.0;
  }
  // B_int(Q_l, Q_r): node m = g + GL round
#pragma unroll 1
  for (int round = 0; round * GL < 6; ++round) {
    const int m = g + GL * round;
    if (m < 6) {
      const double s = c_gl_s[m], oms = 1.0 - s;
      double psi[30];
#pragma unroll
      for (int r = 0; r < 30; ++r) psi[r] = q[r] * oms + qn[r] * s;
      bnode2<FAST, true>(mat, psi, c_gl_w[m], qn[0] - q[0], qn[15] - q[15], bs);
    }
  }
  group_sum<GL>(bs);
  // eigenvalue range of the mid state (Q_l + Q_r) / 2, phase g on lane g
  double mlo = 0.0, mhi = 0.0;
  if (g < 2) {
    // from the neighbours' sum 2 Q_m, as the step forms it (the same bits)
    double qs[15];
#pragma unroll
    for (int r = 0; r < 15; ++r) qs[r] = (g == 0 ? q[r] : q[15 + r]) + (g == 0 ? qn[r] : qn[15 + r]);
    Thermo t;
    thermo<false, FAST>(mat[g], qs[0], qs + 1, t);
    const double c = sound_speed(mat[g], t);
    mlo = t.u[0] - c;
    mhi = t.u[0] + c;
  }
  // F(Q_l) on lane 0, F(Q_r) on lane 1, a phase at a time, with the speed
  // bounds cell_flux forms; each phase's flux difference joins the path
  // integral at once: J = B_int(Q_l, Q_r) + F(Q_r) - F(Q_l) (in bs)
  double lo = 0.0, hi = 0.0;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double fp[15];
#pragma unroll
    for (int r = 0; r < 15; ++r) fp[r] = 0.0;
    if (g < 2) {
      double qp[15];
#pragma unroll
      for (int r = 0; r < 15; ++r) qp[r] = g == 0 ? q[15 * p + r] : qn[15 * p + r];
      Thermo t;
      thermo<false, FAST>(mat[p], qp[0], qp + 1, t);
      phase_flux(mat[p], qp[0], t, qp + 6, fp + 1);
      const double cs = C[p * pitch + i + g];
      const double l = t.u[0] - cs, h = t.u[0] + cs;
      lo = p == 0 ? l : fmin(lo, l);
      hi = p == 0 ? h : fmax(hi, h);
    }
#pragma unroll
    for (int r = 0; r < 15; ++r) {
      const double fl = __shfl(fp[r], 0, GL), fr = __shfl(fp[r], 1, GL);
      bs[15 * p + r] = bs[15 * p + r] + (fr - fl);
    }
  }
  const double smin_l = __shfl(lo, 0, GL), smax_r = __shfl(hi, 1, GL);
  const double mn_m = fmin(__shfl(mlo, 0, GL), __shfl(mlo, 1, GL));
  const double mx_m = fmax(__shfl(mhi, 0, GL), __shfl(mhi, 1, GL));
  const HllCoef hc = hll_coef(mn_m, mx_m, smin_l, smax_r);
  // Q_hll from J, as face_hll forms it
  double qh[30];
#pragma unroll
  for (int r = 0; r < 30; ++r) qh[r] = (qn[r] * hc.s_r - q[r] * hc.s_l - bs[r]) * hc.isd;
  const double J0 = bs[0], J15 = bs[15];
  qh[0] = q[0] + (hc.s_r * (qn[0] - q[0]) - J0) * hc.isd;
  qh[15] = q[15] + (hc.s_r * (qn[15] - q[15]) - J15) * hc.isd;
  // B_int(Q_l, Q_hll) + B_int(Q_hll, Q_r): node m = g + GL round; nodes 0..5
  // are the first segment's
  double bb[30];
#pragma unroll
  for (int r = 0; r < 30; ++r) bb[r] = 0.0;
#pragma unroll 1
  for (int round = 0; round * GL < 12; ++round) {
    const int m = g + GL * round;
    if (m < 12) {
      const bool first = m < 6;
      const int k = first ? m : m - 6;
      const double s = c_gl_s[k], oms = 1.0 - s;
      double psi[30];
#pragma unroll
      for (int r = 0; r < 30; ++r)
        psi[r] = (first ? q[r] : qh[r]) * oms + (first ? qh[r] : qn[r]) * s;
      const double da0 = first ? qh[0] - q[0] : qn[0] - qh[0];
      const double da1 = first ? qh[15] - q[15] : qn[15] - qh[15];
      bnode2<FAST, true>(mat, psi, c_gl_w[k], da0, da1, bb);
    }
  }
  group_sum<GL>(bb);
  // lane g stores rows g, g + GL, g + 2 GL, ... (< 30)
#pragma unroll
  for (int k = 0; k * GL < 30; ++k) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < GL; ++j)
      if (GL * k + j < 30) v = g == j ? bb[GL * k + j] : v;
    const int r = GL * k + g;
    if (r < 30) {
      D[r * pitch + i] = hc.cl * v;
      D[(30 + r) * pitch + i] = hc.cr * v;
    }
  }
}

template <int NPH>
__device__ __forceinline__ void load_chunk(const ChunkCtx& a, long long i, double* q, double* cc) {
  constexpr int NC = NComp<NPH>::value;
  const long long ic = i < a.lo ? a.lo : (i >= a.hi ? a.hi - 1 : i);
#pragma unroll
  for (int r = 0; r < NC; ++r) q[r] = ld_stream(a.Qin + r * a.pitch + ic);
#pragma unroll
  for (int p = 0; p < NPH; ++p) cc[p] = ld_stream(a.Cin + p * a.pitch + ic);
}


}  // namespace hes
