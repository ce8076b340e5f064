// hes_kernels.hip -- gfx950 kernels of the finite-volume time step and the
// C-ABI of libhes.so (declared in include/hes.h).
//
// Time step (main.jl:202-241).  One launch per step:
//   k_step<NPH,FLUX>  reads Q^n and the per-cell sound speeds c^n, takes
//                     dt^n = cfl dx / max lambda^n from a ring slot written by
//                     the previous launch, evaluates each cell's flux once and
//                     each face once (main.jl evaluates every face twice),
//                     updates the interior cells, then computes c^{n+1} of the
//                     new state and max-reduces lambda^{n+1} for the next step.
// State is structure-of-arrays FP64: component k of cell i at Q[k*pitch + i].
// A 64-lane wave owns a tile of 62 cells plus one halo cell on each side, so
// neighbour exchange is a lane shuffle and needs no LDS or barrier.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <mutex>

#include "hes_device.hpp"
#include "hes_faces.hpp"
#include "hes_host.hpp"
#include "hes_step.hpp"
#include "../../include/hes.h"

using namespace hes;

#ifndef HES_STEP_BLOCK
#define HES_STEP_BLOCK 256
#endif
constexpr int BLOCK = HES_STEP_BLOCK;  // threads per block of the step kernel
#ifndef HES_STEP_WAVES_PER_EU
#define HES_STEP_WAVES_PER_EU 3
#endif
#ifndef HES_STEP2_WAVES_PER_EU
#define HES_STEP2_WAVES_PER_EU 2
#endif
// single phase: the pinned schedule (chunk_step's PIN) from this many cells
constexpr long long PIN_MIN_CELLS = 1LL << 21;

// ---------------------------------------------------------------------------
// Step kernel
// ---------------------------------------------------------------------------
struct StepArgs {
  const double* __restrict__ Qin;
  double* __restrict__ Qout;
  const double* __restrict__ Cin;
  double* __restrict__ Cout;
  unsigned long long* slots;  // ring of 3: max lambda of state s lives in slots[s % 3]
  double* tring;              // ring of 2: time of state s in tring[s & 1]
  double* dt_hist;            // optional: dt of each executed step
  int* status;                // [0] error bits, [1] executed steps
  long long n, pitch;
  double cfl, dx, t_end;
  double relax;               // two phases: 1 / tau of the velocity relaxation (0: none)
  double relax_p;             // two phases: 1 / tau of the pressure relaxation (0: none)
  const double* dpath;        // two phases: path terms of the jump faces (k_path2, k_path2_hll)
  unsigned long long* jcnt;   // two phases: jump-face count of step s in jcnt[s & 1]
  long long* jlist;           // two phases: the jump faces of the step (k_mark2)
  int lxf;                    // Rusanov kernels: Lax-Friedrichs face speed dx / dt (main.jl:50)
  int step;                   // step index modulo 6 (ring phase)
  int alt;                    // odd steps walk the tiles from the far end (n <= ALT_MAX_CELLS)
  int hist_idx;
  MatDev mat[2];
};


__device__ __forceinline__ void block_max_to_slot(double v, unsigned long long* slot) {
  double m;
  if (BLOCK == 64) {
    m = wave_max(v);
  } else {
    __shared__ double red[BLOCK / 64];
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    m = red[0];
#pragma unroll
    for (int k = 1; k < BLOCK / 64; ++k) m = fmax(m, red[k]);
  }
  if (threadIdx.x == 0) {
    if (m > 0.0) {
      unsigned long long bits = (unsigned long long)__double_as_longlong(m);
      unsigned long long cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (bits > cur) atomicMax(slot, bits);
    }
  }
}

// Time bookkeeping of a step: returns false if t >= t_end (no step).
__device__ __forceinline__ bool step_begin(const StepArgs& a, double& tcur) {
  tcur = a.tring[a.step & 1];
  if (!(tcur < a.t_end)) {  // main.jl:202 `while t < T`
    if (blockIdx.x == 0 && threadIdx.x == 0) a.tring[(a.step + 1) & 1] = tcur;
    return false;
  }
  return true;
}

// max lambda of the current state, max-reduced by the previous launch (a
// plain load: the previous kernel's atomics are visible at launch, and a plain
// load lets the compiler issue it together with the state loads).
__device__ __forceinline__ double step_lmax(const StepArgs& a) {
  return __longlong_as_double(*(const long long*)(a.slots + (a.step % 3)));
}

__device__ __forceinline__ void step_epilogue(const StepArgs& a, double lam_new, double tcur,
                                              double lmax) {
  block_max_to_slot(lam_new, a.slots + ((a.step + 1) % 3));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double dt = a.cfl * a.dx / lmax;  // main.jl:212
    a.tring[(a.step + 1) & 1] = tcur + dt;
    a.slots[(a.step + 2) % 3] = 0ull;
    if (a.dt_hist) a.dt_hist[a.hist_idx] = dt;
    atomicAdd(a.status + 1, 1);
  }
}

// One chunk per wave.  PIN: see chunk_step (single phase, Rusanov, large grids).
// SRC (two phases): with the relaxation sources (see chunk_step).
template <int NPH, int FLUX, int FAST, int PIN = 0, int SRC = 0>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(NPH == 1 ? HES_STEP_WAVES_PER_EU : HES_STEP2_WAVES_PER_EU))) void k_step(StepArgs a) {
  constexpr int NC = NComp<NPH>::value;
  const int lane = threadIdx.x & 63;
  long long wave = (xcd_block() * BLOCK + threadIdx.x) >> 6;
  // the same set of waves either way (see ALT_MAX_CELLS)
  if ((a.step & 1) && a.alt) wave = (long long)gridDim.x * (BLOCK / 64) - 1 - wave;
  const long long i = wave * TILE + lane - 1;
  double tcur;
  prio_loads();
  if (!step_begin(a, tcur)) return;
  const ChunkCtx ctx = {a.mat, a.Qin, a.Qout, a.Cin, a.Cout, a.status, a.pitch, 0, a.n, a.cfl, a.dx,
                        a.relax, a.dpath, a.lxf, a.relax_p};
  double q[NC], cc[NPH];
  load_chunk<NPH>(ctx, i, q, cc);
  const double lmax = step_lmax(a);
  prio_compute();
  constexpr bool PB = NPH == 2;  // path integrals from k_path2 / k_path2_hll
  // the pinned single-phase Rusanov step keeps no sound speeds (see chunk_step)
  constexpr bool NOC = NPH == 1 && FLUX == FLUX_RUSANOV && PIN != 0;
  double lam = chunk_step<NPH, FLUX, FAST, PB, PIN != 0, SRC != 0, NOC>(ctx, i, q, cc, lmax);
  step_epilogue(a, lam, tcur, lmax);
}

// Two phases: the faces with a volume-fraction jump of the step's input state,
// appended to a list (face index i of face i+1/2) whose length is counted in
// jcnt[step & 1]; the other slot is cleared here for the next step.  A step
// that is a no-op (t >= t_end) also clears its own slot: the host does not
// count it, so the next run starts at the same step index and slot.  One face
// per thread, one atomic per block of MARK_BLK threads.
constexpr int MARK_BLK = 256;
__global__ __launch_bounds__(MARK_BLK) void k_mark2(StepArgs a) {
  const long long i = (long long)blockIdx.x * MARK_BLK + threadIdx.x;
  const bool live = a.tring[a.step & 1] < a.t_end;  // uniform over the grid
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.jcnt[(a.step + 1) & 1] = 0ull;
    if (!live) a.jcnt[a.step & 1] = 0ull;
  }
  if (!live) return;  // the step is a no-op too (every thread of every block)
  const bool jump = i < a.n - 1 && face_jump2(a.Qin, i, a.pitch);
  const unsigned long long pos = block_append<MARK_BLK>(jump, a.jcnt + (a.step & 1));
  if (jump && pos < (unsigned long long)a.n) a.jlist[pos] = i;  // at most n - 1 faces
}

// The path terms of the listed jump faces over a fixed grid: a group of
// HES_PATH_LANES lanes per face, 64 / HES_PATH_LANES faces per wave (the faces
// without a jump never reach this kernel).
template <int FAST>
__global__ __launch_bounds__(64) void k_path2(StepArgs a) {
  if (!(a.tring[a.step & 1] < a.t_end)) return;
  const unsigned long long cnt = min(a.jcnt[a.step & 1], (unsigned long long)a.n - 1ull);
  constexpr int GL = HES_PATH_LANES;
  const unsigned long long stride = (unsigned long long)gridDim.x * (64 / GL);
  for (unsigned long long j = ((unsigned long long)blockIdx.x * 64 + threadIdx.x) / GL; j < cnt;
       j += stride)
    face_path2<FAST>(a.mat, a.Qin, (double*)a.dpath, a.jlist[j], a.pitch, threadIdx.x & (GL - 1));
}

// (the HLL form: HES_PATH_HLL_LANES lanes per face)
template <int FAST>
__global__ __launch_bounds__(64) void k_path2_hll(StepArgs a) {
  if (!(a.tring[a.step & 1] < a.t_end)) return;
  const unsigned long long cnt = min(a.jcnt[a.step & 1], (unsigned long long)a.n - 1ull);
  constexpr int GL = HES_PATH_HLL_LANES;
  const unsigned long long stride = (unsigned long long)gridDim.x * (64 / GL);
  for (unsigned long long j = ((unsigned long long)blockIdx.x * 64 + threadIdx.x) / GL; j < cnt;
       j += stride)
    face_path2_hll<FAST>(a.mat, a.Qin, a.Cin, (double*)a.dpath, a.jlist[j], a.pitch,
                         threadIdx.x & (GL - 1));
}

// Sound speeds of every cell of a state and the max CFL speed (main.jl:204-212).
template <int NPH>
__global__ __launch_bounds__(BLOCK) void k_speeds(const double* __restrict__ Q,
                                                  double* __restrict__ C, long long n,
                                                  long long pitch, MatDev m0, MatDev m1,
                                                  unsigned long long* slot, int* status) {
  constexpr int NC = NComp<NPH>::value;
  const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
  MatDev mat[2] = {m0, m1};
  double lam = 0.0;
  int bad = 0;
  if (i < n) {
    double q[NC], cc[NPH];
#pragma unroll
    for (int r = 0; r < NC; ++r) q[r] = Q[r * pitch + i];
    lam = cell_speeds<NPH>(mat, q, cc, bad);
#pragma unroll
    for (int p = 0; p < NPH; ++p) C[p * pitch + i] = cc[p];
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
  block_max_to_slot(lam, slot);
}

// AoS (component fastest, the Julia Matrix layout) <-> SoA.
__global__ void k_aos_to_soa(const double* __restrict__ in, double* __restrict__ out, int nc,
                             long long n, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < nc; ++r) out[r * pitch + i] = in[i * nc + r];
}
__global__ void k_soa_to_aos(const double* __restrict__ in, double* __restrict__ out, int nc,
                             long long n, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < nc; ++r) out[i * nc + r] = in[r * pitch + i];
}

// Piecewise-constant initial state: cell i (0-based, centre x = (i + 1/2) / n)
// takes state k = #{j : x > s_j} for sorted interfaces s_j.  With one interface
// at 1/2 this is main.jl:103 (cell i is left iff i < n/2).
__global__ void k_piecewise(const double* __restrict__ states, const double* __restrict__ ifc,
                            int nifc, int nc, double* __restrict__ Q, long long n,
                            long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = ((double)i + 0.5) / (double)n;
  int lo = 0, hi = nifc;  // first j with !(x > s_j)
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x > ifc[mid]) lo = mid + 1;
    else hi = mid;
  }
  const double* src = states + (long long)lo * nc;
  for (int r = 0; r < nc; ++r) Q[r * pitch + i] = src[r];
}

// ---------------------------------------------------------------------------
// Point-wise kernels behind the module API (one thread per point, AoS data)
// ---------------------------------------------------------------------------
enum PointOp {
  OP_CONS2PRIM = 0,
  OP_PRIM2CONS,
  OP_FLUX,
  OP_EIGVALS,
  OP_NONCONS,
  OP_MAXSPEED,
};

template <int NPH>
__global__ __launch_bounds__(64) void k_point(int op, const double* __restrict__ in,
                                              double* __restrict__ out, long long n, MatDev m0,
                                              MatDev m1, double n0, double n1, double n2,
                                              int* status) {
  constexpr int NC = NComp<NPH>::value;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  MatDev mat[2] = {m0, m1};
  double q[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) q[r] = in[i * NC + r];
  int bad = 0;
  if (op == OP_CONS2PRIM) {  // HyperelasticityMPh.jl:99-133
    double* o = out + i * NC;
    for (int p = 0; p < NPH; ++p) {
      Thermo t;
      phase_thermo<NPH>(mat, q, p, t);
      const MatDev& m = mat[NPH == 1 ? 0 : p];
      int b0 = NPH == 1 ? -1 : 15 * p;
      if (NPH == 2) o[b0] = q[b0];
      o[b0 + 1] = t.rho;
      o[b0 + 2] = t.u[0];
      o[b0 + 3] = t.u[1];
      o[b0 + 4] = t.u[2];
      o[b0 + 5] = entropy_of(m, t);
      for (int k = 0; k < 9; ++k) o[b0 + 6 + k] = t.A[k];
      bad |= t.bad;
    }
  } else if (op == OP_PRIM2CONS) {  // HyperelasticityMPh.jl:63-90
    double* o = out + i * NC;
    for (int p = 0; p < NPH; ++p) {
      const MatDev& m = mat[NPH == 1 ? 0 : p];
      int b0 = NPH == 1 ? -1 : 15 * p;
      double frac = NPH == 1 ? 1.0 : q[b0];
      double true_den = q[b0 + 1];
      double den = frac * true_den;
      const double* vel = q + b0 + 2;
      double S = q[b0 + 5];
      const double* A = q + b0 + 6;
      double e_int = energy_from_A(m, A, exp(S / m.cv));
      double e_kin = (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) / 2.0;
      if (NPH == 2) o[b0] = frac;
      o[b0 + 1] = den;
      for (int k = 0; k < 3; ++k) o[b0 + 2 + k] = den * vel[k];
      o[b0 + 5] = den * (e_int + e_kin);
      for (int k = 0; k < 9; ++k) o[b0 + 6 + k] = den * A[k];
    }
  } else if (op == OP_FLUX) {  // HyperelasticityMPh.jl:140-175
    double cc[2] = {0.0, 0.0}, f[NC], smin, smax;
    cell_flux<NPH>(mat, q, cc, f, smin, smax, bad);
    for (int r = 0; r < NC; ++r) out[i * NC + r] = f[r];
  } else if (op == OP_EIGVALS) {  // HyperelasticityMPh.jl:252-266 -> 6 per phase
    for (int p = 0; p < NPH; ++p) {
      Thermo t;
      phase_thermo<NPH>(mat, q, p, t);
      const double nv[3] = {n0, n1, n2};
      double O9[9], w[3];
      acoustic_n(mat[NPH == 1 ? 0 : p], t, nv, O9);
      double O[6] = {O9[0], 0.5 * (O9[1] + O9[3]), 0.5 * (O9[2] + O9[6]), O9[4],
                     0.5 * (O9[5] + O9[7]), O9[8]};
      eig3_sym(O, w);
      double spd = t.u[0] * n0 + t.u[1] * n1 + t.u[2] * n2;  // dot(P[3:5], n)
      double* o = out + i * 6 * NPH + 6 * p;
      for (int k = 0; k < 3; ++k) {
        double c = sqrt(fabs(w[k]));
        o[k] = spd + c;
        o[3 + k] = spd - c;
      }
      bad |= t.bad;
    }
  } else if (op == OP_NONCONS && NPH == 2) {  // HyperelasticityMPh.jl:178-250 -> 30x30
    BCol b;
    bcol2(mat, q, b);
    double* o = out + i * 900;
    for (int k = 0; k < 900; ++k) o[k] = 0.0;
    for (int p = 0; p < 2; ++p) {
      int off = 15 * p;
      double col[15];
      col[0] = b.uI0;
      col[1] = 0.0;
      col[2] = b.sI[0];
      col[3] = b.sI[1];
      col[4] = b.sI[2];
      col[5] = b.sIu;
      for (int r = 0; r < 9; ++r) col[6 + r] = b.a[p][r];
      for (int r = 0; r < 15; ++r) o[(off + r) + 30 * off] = col[r];
    }
  } else if (op == OP_MAXSPEED) {
    double cc[NPH];
    out[i] = cell_speeds<NPH>(mat, q, cc, bad);
  }
  if (bad) atomicOr(status, 1);
}

// EquationsOfState point functions: energy / entropy / stress / acoustic.
enum EosOp { EOS_ENERGY = 0, EOS_ENTROPY, EOS_STRESS, EOS_STRESS_DEN, EOS_ACOUSTIC, EOS_ACOUSTIC_PRIM,
             EOS_TEMPERATURE, EOS_FINGER, EOS_INVARIANTS };

__global__ __launch_bounds__(64) void k_eos(int op, MatDev m, const double* __restrict__ s, const double* __restrict__ s2,
                      const double* __restrict__ v9, const double* __restrict__ nv,
                      double* __restrict__ out, long long n, int* status) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int bad = 0;
  if (op == EOS_ENERGY) {  // energy(eos, S, G), EquationsOfState.jl:118
    out[i] = energy_G(m, s[i], v9 + 9 * i);
  } else if (op == EOS_ENTROPY) {  // entropy(eos, e_int, G), :139
    out[i] = entropy_G(m, s[i], v9 + 9 * i);
  } else if (op == EOS_TEMPERATURE) {  // d energy / dS
    const double* g = v9 + 9 * i;
    double inv[3];
    invariants_gen(g, inv);
    out[i] = m.t0 * exp(m.half_g * log(inv[2])) * exp(s[i] / m.cv);
  } else if (op == EOS_STRESS || op == EOS_ACOUSTIC) {  // stress(eos, ent, F) :179 / acoustic :223
    Thermo t;
    thermo_from_A(m, v9 + 9 * i, exp(s[i] / m.cv), t);
    bad = t.bad;
    if (op == EOS_STRESS) {
      stress_full(m, t, out + 9 * i);
    } else {
      acoustic_n(m, t, nv + 3 * i, out + 9 * i);
    }
  } else if (op == EOS_STRESS_DEN) {  // stress(eos, den, e_int, F) :162
    const double* F = v9 + 9 * i;
    double G[9];
    {
      Thermo t;
      thermo_from_A(m, F, 1.0, t);
      G[0] = t.G11; G[1] = t.G12; G[2] = t.G13;
      G[3] = t.G12; G[4] = t.G22; G[5] = t.G23;
      G[6] = t.G13; G[7] = t.G23; G[8] = t.G33;
    }
    double S = entropy_G(m, s2[i], G);
    Thermo t;
    thermo_from_A(m, F, exp(S / m.cv), t);
    t.rho = s[i];  // the density argument, as the reference uses it
    bad = t.bad;
    stress_full(m, t, out + 9 * i);
  } else if (op == EOS_ACOUSTIC_PRIM) {  // acoustic(eos, P, n) :192 ; v9 holds P (15)
    const double* P = v9 + 15 * i;
    const double* F = P + 6;
    // e_int = energy(eos, P[6], G); the stress then recomputes the entropy from
    // e_int, so x is the entropy law's argument (clamped at 1e-6).
    Thermo t;
    double x0 = exp(P[5] / m.cv);
    thermo_from_A(m, F, x0, t);
    double sam1 = t.sa - 1.0;
    double Uc = m.hk0a2 * sam1 * sam1;
    double e_int = Uc + t.Eth + t.W;
    double x = (e_int - t.W - Uc) / (m.cvt0 * t.sg) + 1.0;
    bool clamped = x < 1e-6;
    if (clamped) x = 1e-6;
    t.x = x;
    t.Eth = m.cvt0 * t.sg * (x - 1.0);
    t.rho = P[0] * P[1];  // den = P[1] * P[2], held fixed
    acoustic_n(m, t, nv + 3 * i, out + 9 * i, true, clamped);
    bad = t.bad;
  } else if (op == EOS_FINGER) {  // Strains.jl:26-32
    Thermo t;
    thermo_from_A(m, v9 + 9 * i, 1.0, t);
    double* g = out + 9 * i;
    g[0] = t.G11; g[1] = t.G12; g[2] = t.G13;
    g[3] = t.G12; g[4] = t.G22; g[5] = t.G23;
    g[6] = t.G13; g[7] = t.G23; g[8] = t.G33;
  } else if (op == EOS_INVARIANTS) {  // Strains.jl:46-52
    invariants_gen(v9 + 9 * i, out + 3 * i);
  }
  if (bad) atomicOr(status, 1);
}

// Face fluxes for arrays of face states (NumFluxes.lxf / hll).
template <int NPH, int FLUX>
__global__ __launch_bounds__(64) void k_face(const double* __restrict__ QL, const double* __restrict__ QR,
                       const double* __restrict__ aux, double* __restrict__ cons,
                       double* __restrict__ dm, double* __restrict__ dp, long long n, MatDev m0,
                       MatDev m1) {
  constexpr int NC = NComp<NPH>::value;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  MatDev mat[2] = {m0, m1};
  double ql[NC], qr[NC], fl[NC], fr[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    ql[r] = QL[i * NC + r];
    qr[r] = QR[i * NC + r];
  }
  double cc[2] = {0.0, 0.0}, s0, s1;
  int bad;
  cell_flux<NPH>(mat, ql, cc, fl, s0, s1, bad);
  cell_flux<NPH>(mat, qr, cc, fr, s0, s1, bad);
  if (FLUX == FLUX_RUSANOV) {
    double c[NC], d[NC];
    face_rusanov<NPH>(mat, ql, qr, fl, fr, aux[i], c, d);
    for (int r = 0; r < NC; ++r) {
      cons[i * NC + r] = c[r];
      dm[i * NC + r] = NPH == 2 ? d[r] : 0.0;
      dp[i * NC + r] = NPH == 2 ? d[r] : 0.0;
    }
  } else {
    // aux holds [min eig of left, max eig of right] per face
    double a[NC], b[NC];
    face_hll<NPH>(mat, ql, qr, fl, fr, aux[2 * i], aux[2 * i + 1], a, b);
    for (int r = 0; r < NC; ++r) {
      cons[i * NC + r] = 0.0;
      dm[i * NC + r] = a[r];
      dp[i * NC + r] = b[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Solver handle
// ---------------------------------------------------------------------------
struct hes_solver {
  int nph, flux, nc;
  long long n, pitch;
  double mats[16];
  MatDev mat[2];
  double cfl, dx;
  double relax;         // 1 / tau of the two-phase velocity relaxation (0: none)
  double relax_p;       // 1 / tau of the two-phase pressure relaxation (0: none)
  double* dpath;        // two phases: half path integrals of the faces (ncomp x pitch)
  unsigned long long* jcnt;  // two phases: jump-face counts (2 slots, by step parity)
  long long* jlist;          // two phases: jump-face list (n entries)
  double* Q[2];
  double* C[2];
  unsigned long long* slots;
  double* tring;
  int* status;
  double* dt_hist;
  long long dt_cap;
  long long step;       // steps executed so far (state index)
  double t;             // time of the current state
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  bool has_state;
};

namespace {

template <int NPH, int FLUX>
void launch_step_t(const StepArgs& a, long long nchunks, bool fast, hipStream_t s) {
  constexpr int WPB = BLOCK / 64;
  long long nblocks = (nchunks + WPB - 1) / WPB;
  if (NPH == 2) {
    // jump faces first (a light scan), then their path terms over a fixed grid
    const unsigned nbp = (unsigned)((a.n + MARK_BLK - 1) / MARK_BLK);
    hipLaunchKernelGGL(k_mark2, dim3(nbp), dim3(MARK_BLK), 0, s, a);
    // 4 faces per block; as many blocks as the kernels' occupancy keeps
    // resident (1024 SIMDs): two waves per SIMD for Rusanov, one for HLL
    const long long nbf = (a.n + 254) / 256;
    if (FLUX == FLUX_RUSANOV) {
      const unsigned nbl = (unsigned)std::min<long long>(nbf, 2 * 1024);
      if (fast) hipLaunchKernelGGL(k_path2<1>, dim3(nbl), dim3(64), 0, s, a);
      else hipLaunchKernelGGL(k_path2<0>, dim3(nbl), dim3(64), 0, s, a);
    } else {
      const unsigned nbl = (unsigned)std::min<long long>(nbf, 1024);
      if (fast) hipLaunchKernelGGL(k_path2_hll<1>, dim3(nbl), dim3(64), 0, s, a);
      else hipLaunchKernelGGL(k_path2_hll<0>, dim3(nbl), dim3(64), 0, s, a);
    }
  }
  if constexpr (NPH == 1) {
    // single phase from 2^21 cells: the pinned schedule (chunk_step)
    if (a.n >= PIN_MIN_CELLS) {
      if (fast)
        hipLaunchKernelGGL((k_step<NPH, FLUX, 1, 1>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
      else
        hipLaunchKernelGGL((k_step<NPH, FLUX, 0, 1>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
      return;
    }
  }
  if constexpr (NPH == 2) {
    if (a.relax > 0.0 || a.relax_p > 0.0) {  // with the relaxation sources
      if (fast)
        hipLaunchKernelGGL((k_step<NPH, FLUX, 1, 0, 1>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
      else
        hipLaunchKernelGGL((k_step<NPH, FLUX, 0, 0, 1>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
      return;
    }
  }
  if (fast)
    hipLaunchKernelGGL((k_step<NPH, FLUX, 1>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL((k_step<NPH, FLUX, 0>), dim3((unsigned)nblocks), dim3(BLOCK), 0, s, a);
}

bool all_fast(const hes_solver* h);

void launch_step(hes_solver* h, const StepArgs& a) {
  long long nwaves = (h->n + TILE - 1) / TILE;
  bool fast = all_fast(h);
  // Lax-Friedrichs runs the Rusanov kernels with the fixed face speed (a.lxf)
  const bool hll = h->flux == HES_FLUX_HLL;
  if (h->nph == 1 && !hll) launch_step_t<1, FLUX_RUSANOV>(a, nwaves, fast, h->stream);
  else if (h->nph == 1) launch_step_t<1, FLUX_HLL>(a, nwaves, fast, h->stream);
  else if (!hll) launch_step_t<2, FLUX_RUSANOV>(a, nwaves, fast, h->stream);
  else launch_step_t<2, FLUX_HLL>(a, nwaves, fast, h->stream);
}

bool all_fast(const hes_solver* h) {
  return h->mat[0].fast && (h->nph == 1 || h->mat[1].fast);
}

StepArgs make_args(hes_solver* h, long long s, double t_end, double* hist, int hist_idx) {
  StepArgs a;
  a.Qin = h->Q[s & 1];
  a.Qout = h->Q[(s + 1) & 1];
  a.Cin = h->C[s & 1];
  a.Cout = h->C[(s + 1) & 1];
  a.slots = h->slots;
  a.tring = h->tring;
  a.dt_hist = hist;
  a.status = h->status;
  a.n = h->n;
  a.pitch = h->pitch;
  a.cfl = h->cfl;
  a.dx = h->dx;
  a.t_end = t_end;
  a.relax = h->relax;
  a.relax_p = h->relax_p;
  a.dpath = h->dpath;
  a.jcnt = h->jcnt;
  a.jlist = h->jlist;
  a.lxf = h->flux == HES_FLUX_LXF;
  a.step = (int)(s % 6);
  a.alt = h->n <= ALT_MAX_CELLS;
  a.hist_idx = hist_idx;
  a.mat[0] = h->mat[0];
  a.mat[1] = h->mat[1];
  return a;
}

// Recompute sound speeds and the CFL max of the current state; reset rings.
int init_speeds(hes_solver* h) {
  long long s = h->step;
  HIP_TRY(hipMemsetAsync(h->slots, 0, 3 * sizeof(unsigned long long), h->stream));
  if (h->jcnt) HIP_TRY(hipMemsetAsync(h->jcnt, 0, 2 * sizeof(unsigned long long), h->stream));
  double tr[2] = {h->t, h->t};
  HIP_TRY(hipMemcpyAsync(h->tring, tr, sizeof(tr), hipMemcpyHostToDevice, h->stream));
  long long nb = (h->n + BLOCK - 1) / BLOCK;
  const double* Qs = h->Q[s & 1];
  double* Cs = h->C[s & 1];
  if (h->nph == 1)
    hipLaunchKernelGGL(k_speeds<1>, dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, Qs, Cs, h->n,
                       h->pitch, h->mat[0], h->mat[1], h->slots + (s % 3), h->status);
  else
    hipLaunchKernelGGL(k_speeds<2>, dim3((unsigned)nb), dim3(BLOCK), 0, h->stream, Qs, Cs, h->n,
                       h->pitch, h->mat[0], h->mat[1], h->slots + (s % 3), h->status);
  HIP_TRY(hipGetLastError());
  // the end cells never change: copy their state and speeds into the other buffer
  for (int r = 0; r < h->nc; ++r) {
    HIP_TRY(hipMemcpyAsync(h->Q[(s + 1) & 1] + r * h->pitch, h->Q[s & 1] + r * h->pitch,
                           sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->Q[(s + 1) & 1] + r * h->pitch + h->n - 1,
                           h->Q[s & 1] + r * h->pitch + h->n - 1, sizeof(double),
                           hipMemcpyDeviceToDevice, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  int st[2];
  HIP_TRY(copy_on(h->stream, st, h->status, sizeof(st), hipMemcpyDeviceToHost));
  if (st[0]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

}  // namespace

extern "C" {

const char* hes_last_error(void) { return g_err.c_str(); }

int hes_version(void) { return HES_ABI_VERSION; }

int hes_device_info(char* name, int len, int* cus, double* mem_gb) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  if (name && len > 0) {
    std::snprintf(name, (size_t)len, "%s", p.gcnArchName);
  }
  if (cus) *cus = p.multiProcessorCount;
  if (mem_gb) *mem_gb = (double)p.totalGlobalMem / 1e9;
  return HES_OK;
}

int hes_solver_create(hes_solver** out, int nphase, int flux, long long ncells,
                      const double* materials, double cfl, double dx) {
  if (!out) return fail(HES_ERR_ARG, "null handle pointer");
  *out = nullptr;
  if (nphase != 1 && nphase != 2) return fail(HES_ERR_ARG, "nphase must be 1 or 2");
  if (flux != HES_FLUX_RUSANOV && flux != HES_FLUX_HLL && flux != HES_FLUX_LXF)
    return fail(HES_ERR_ARG, "unknown flux");
  if (ncells < 3) return fail(HES_ERR_ARG, "need at least 3 cells");
  if (!materials) return fail(HES_ERR_ARG, "null materials");
  for (int p = 0; p < nphase; ++p)
    if (!mat_ok(materials + 8 * p)) return fail(HES_ERR_ARG, "invalid material constants");
  if (!(cfl > 0) || !(dx > 0)) return fail(HES_ERR_ARG, "cfl and dx must be positive");
  hes_solver* h = new hes_solver();
  std::memset(h, 0, sizeof(*h));
  h->nph = nphase;
  h->flux = flux;
  h->nc = nphase == 1 ? 14 : 30;
  h->n = ncells;
  // rows 64-aligned and padded by 2 KB: with a power-of-two row stride (n =
  // 2^k cells) a wave's 15 row loads fall on the same HBM channel; the pad
  // spreads them (about 2% in a memory-only copy with the step's pattern)
  h->pitch = (ncells + 63) / 64 * 64 + 256;
  for (int k = 0; k < 8 * nphase; ++k) h->mats[k] = materials[k];
  h->mat[0] = make_mat(materials);
  h->mat[1] = make_mat(nphase == 2 ? materials + 8 : materials);
  h->cfl = cfl;
  h->dx = dx;
  size_t qbytes = sizeof(double) * (size_t)h->nc * (size_t)h->pitch;
  size_t cbytes = sizeof(double) * (size_t)nphase * (size_t)h->pitch;
  hipError_t e = hipSuccess;
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    e = hipMalloc(&h->Q[b], qbytes);
    if (e == hipSuccess) e = hipMalloc(&h->C[b], cbytes);
  }
  if (e == hipSuccess && nphase == 2) e = hipMalloc(&h->dpath, 2 * qbytes);  // D-, D+ rows
  if (e == hipSuccess && nphase == 2) e = hipMalloc(&h->jcnt, 2 * sizeof(unsigned long long));
  if (e == hipSuccess && nphase == 2) e = hipMalloc(&h->jlist, sizeof(long long) * (size_t)ncells);
  if (e == hipSuccess) e = hipMalloc(&h->slots, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&h->tring, 2 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->status, 4 * sizeof(int));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = zero_on(h->stream, h->status, 4 * sizeof(int));
  if (e != hipSuccess) {
    g_err = std::string("allocation failed: ") + hipGetErrorString(e);
    hes_solver_destroy(h);
    return HES_ERR_HIP;
  }
  *out = h;
  return HES_OK;
}

int hes_solver_set_relaxation(hes_solver* h, double tau) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (h->nph != 2) return fail(HES_ERR_ARG, "velocity relaxation needs two phases");
  if (!(tau > 0.0)) return fail(HES_ERR_ARG, "tau must be positive (infinity: off)");
  h->relax = std::isinf(tau) ? 0.0 : 1.0 / tau;
  return HES_OK;
}

int hes_solver_set_pressure_relaxation(hes_solver* h, double tau) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (h->nph != 2) return fail(HES_ERR_ARG, "pressure relaxation needs two phases");
  if (!(tau > 0.0)) return fail(HES_ERR_ARG, "tau must be positive (INFINITY: off)");
  h->relax_p = std::isinf(tau) ? 0.0 : 1.0 / tau;
  return HES_OK;
}

int hes_solver_destroy(hes_solver* h) {
  if (!h) return HES_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int b = 0; b < 2; ++b) {
    if (h->Q[b]) (void)hipFree(h->Q[b]);
    if (h->C[b]) (void)hipFree(h->C[b]);
  }
  if (h->slots) (void)hipFree(h->slots);
  if (h->dpath) (void)hipFree(h->dpath);
  if (h->jcnt) (void)hipFree(h->jcnt);
  if (h->jlist) (void)hipFree(h->jlist);
  if (h->tring) (void)hipFree(h->tring);
  if (h->status) (void)hipFree(h->status);
  if (h->dt_hist) (void)hipFree(h->dt_hist);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return HES_OK;
}

static int set_state_common(hes_solver* h, const double* src, hipMemcpyKind kind, double t0) {
  size_t bytes = sizeof(double) * (size_t)h->nc * (size_t)h->n;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  hipError_t e = hipMemcpyAsync(tmp, src, bytes, kind, h->stream);
  if (e == hipSuccess) {
    h->step = 0;
    h->t = t0;
    long long nb = (h->n + 255) / 256;
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)nb), dim3(256), 0, h->stream, tmp, h->Q[0],
                       h->nc, h->n, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    g_err = std::string("set_state: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  HIP_TRY(zero_on(h->stream, h->status, 4 * sizeof(int)));
  h->has_state = true;
  return init_speeds(h);
}

int hes_solver_set_state(hes_solver* h, const double* Q, double t0) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  return set_state_common(h, Q, hipMemcpyHostToDevice, t0);
}

int hes_solver_set_piecewise(hes_solver* h, int nstates, const double* states,
                             const double* interfaces, double t0) {
  if (!h || !states || nstates < 1 || (nstates > 1 && !interfaces))
    return fail(HES_ERR_ARG, "null argument");
  for (int j = 0; j + 1 < nstates - 1; ++j)
    if (!(interfaces[j] <= interfaces[j + 1])) return fail(HES_ERR_ARG, "interfaces not sorted");
  double *ds = nullptr, *di = nullptr;
  HIP_TRY(hipMalloc(&ds, sizeof(double) * (size_t)nstates * h->nc));
  HIP_TRY(hipMalloc(&di, sizeof(double) * (size_t)(nstates > 1 ? nstates - 1 : 1)));
  hipError_t e = copy_on(h->stream, ds, states, sizeof(double) * (size_t)nstates * h->nc,
                          hipMemcpyHostToDevice);
  if (e == hipSuccess && nstates > 1)
    e = copy_on(h->stream, di, interfaces, sizeof(double) * (size_t)(nstates - 1),
                hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    h->step = 0;
    h->t = t0;
    long long nb = (h->n + 255) / 256;
    hipLaunchKernelGGL(k_piecewise, dim3((unsigned)nb), dim3(256), 0, h->stream, ds, di,
                       nstates - 1, h->nc, h->Q[0], h->n, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(ds);
  (void)hipFree(di);
  if (e != hipSuccess) {
    g_err = std::string("set_piecewise: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  HIP_TRY(zero_on(h->stream, h->status, 4 * sizeof(int)));
  h->has_state = true;
  return init_speeds(h);
}

int hes_solver_set_state_device(hes_solver* h, const double* dQ, double t0) {
  if (!h || !dQ) return fail(HES_ERR_ARG, "null argument");
  return set_state_common(h, dQ, hipMemcpyDeviceToDevice, t0);
}

int hes_solver_get_state(hes_solver* h, double* Q) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  size_t bytes = sizeof(double) * (size_t)h->nc * (size_t)h->n;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  long long nb = (h->n + 255) / 256;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)nb), dim3(256), 0, h->stream, h->Q[h->step & 1],
                     tmp, h->nc, h->n, h->pitch);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(Q, tmp, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    g_err = std::string("get_state: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return HES_OK;
}

int hes_solver_get_speeds(hes_solver* h, double* c) {
  if (!h || !c) return fail(HES_ERR_ARG, "null argument");
  if (h->nph == 1 && h->flux != HES_FLUX_HLL && h->n >= PIN_MIN_CELLS) {
    // the pinned Rusanov steps keep no sound speeds: form them for the
    // current state (the CFL maximum goes to the spare fourth slot)
    const long long nb = (h->n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_speeds<1>, dim3((unsigned)nb), dim3(BLOCK), 0, h->stream,
                       h->Q[h->step & 1], h->C[h->step & 1], h->n, h->pitch, h->mat[0], h->mat[1],
                       h->slots + 3, h->status);
    HIP_TRY(hipGetLastError());
  }
  for (int p = 0; p < h->nph; ++p)
    HIP_TRY(copy_on(h->stream, c + p * h->n, h->C[h->step & 1] + p * h->pitch,
                    sizeof(double) * h->n, hipMemcpyDeviceToHost));
  return HES_OK;
}

int hes_solver_time(hes_solver* h, double* t, long long* steps) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (t) *t = h->t;
  if (steps) *steps = h->step;
  return HES_OK;
}

int hes_solver_next_dt(hes_solver* h, double* dt) {
  if (!h || !dt) return fail(HES_ERR_ARG, "null argument");
  unsigned long long bits = 0;
  HIP_TRY(copy_on(h->stream, &bits, h->slots + (h->step % 3), sizeof(bits), hipMemcpyDeviceToHost));
  double lmax;
  std::memcpy(&lmax, &bits, sizeof(lmax));
  *dt = h->cfl * h->dx / lmax;
  return HES_OK;
}

// Advance up to max_steps steps while t < t_end.  Launches are issued in
// chunks of `chunk` steps (captured once into a hipGraph when use_graph is
// set) and the executed-step counter is read back after each chunk.
int hes_solver_run(hes_solver* h, long long max_steps, double t_end, int use_graph,
                   long long* steps_done, double* t_now, float* elapsed_ms, double* dt_out) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  if (max_steps < 0) return fail(HES_ERR_ARG, "negative step count");
  const long long chunk = 60;  // multiple of 6: ring phases repeat per chunk
  if (dt_out) {
    if (h->dt_cap < max_steps) {
      if (h->dt_hist) HIP_TRY(hipFree(h->dt_hist));
      h->dt_hist = nullptr;
      h->dt_cap = 0;
      HIP_TRY(hipMalloc(&h->dt_hist, sizeof(double) * (size_t)(max_steps > 0 ? max_steps : 1)));
      h->dt_cap = max_steps;
    }
  }
  HIP_TRY(hipMemsetAsync(h->status + 1, 0, sizeof(int), h->stream));
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  long long graph_base = -1;  // step (mod 6) the captured graph starts from
  long long launched = 0, done = 0;
  const long long s0 = h->step;
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  int rc = HES_OK;
  while (launched < max_steps) {
    long long k = max_steps - launched < chunk ? max_steps - launched : chunk;
    long long sb = s0 + launched;
    bool graph_ok = use_graph && k == chunk && !dt_out;
    if (graph_ok) {
      if (!gexec || (graph_base % 6) != (sb % 6)) {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) { rc = fail(HES_ERR_HIP, "stream capture failed"); break; }
        for (long long j = 0; j < k; ++j) launch_step(h, make_args(h, sb + j, t_end, nullptr, 0));
        e = hipStreamEndCapture(h->stream, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) { g_err = std::string("graph: ") + hipGetErrorString(e); rc = HES_ERR_HIP; break; }
        graph_base = sb;
      }
      hipError_t e = hipGraphLaunch(gexec, h->stream);
      if (e != hipSuccess) { g_err = std::string("graph launch: ") + hipGetErrorString(e); rc = HES_ERR_HIP; break; }
    } else {
      for (long long j = 0; j < k; ++j)
        launch_step(h, make_args(h, sb + j, t_end, dt_out ? h->dt_hist : nullptr, (int)(launched + j)));
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) { g_err = std::string("launch: ") + hipGetErrorString(e); rc = HES_ERR_HIP; break; }
    }
    launched += k;
    if (t_end < INFINITY || launched >= max_steps) {
      int cnt = 0;
      hipError_t e = hipMemcpyAsync(&cnt, h->status + 1, sizeof(int), hipMemcpyDeviceToHost, h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) { g_err = std::string("sync: ") + hipGetErrorString(e); rc = HES_ERR_HIP; break; }
      done = cnt;
      if (done < launched) break;  // t reached t_end inside this chunk
    }
  }
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc != HES_OK) return rc;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  int st[2];
  HIP_TRY(copy_on(h->stream, st, h->status, sizeof(st), hipMemcpyDeviceToHost));
  done = st[1];
  h->step = s0 + done;
  double tr[2];
  HIP_TRY(copy_on(h->stream, tr, h->tring, sizeof(tr), hipMemcpyDeviceToHost));
  h->t = tr[h->step & 1];
  if (steps_done) *steps_done = done;
  if (t_now) *t_now = h->t;
  if (elapsed_ms) *elapsed_ms = ms;
  if (dt_out && done > 0)
    HIP_TRY(copy_on(h->stream, dt_out, h->dt_hist, sizeof(double) * (size_t)done,
                    hipMemcpyDeviceToHost));
  if (st[0]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

// Time `reps` launches of the step kernel alone on the solver's stream (HIP
// events bracket exactly these launches).  The state advances by reps steps.
int hes_solver_time_steps(hes_solver* h, long long reps, float* elapsed_ms) {
  if (!h || !elapsed_ms) return fail(HES_ERR_ARG, "null argument");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  if (reps > 100000) return fail(HES_ERR_ARG, "at most 100000 timed steps");
  HIP_TRY(hipMemsetAsync(h->status + 1, 0, sizeof(int), h->stream));
  // capture the steps first (outside the timed region)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(HES_ERR_HIP, "stream capture failed");
  for (long long j = 0; j < reps; ++j) launch_step(h, make_args(h, h->step + j, INFINITY, nullptr, 0));
  e = hipStreamEndCapture(h->stream, &graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
  if (e == hipSuccess) e = hipGraphLaunch(gexec, h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
  if (e == hipSuccess) e = hipEventElapsedTime(elapsed_ms, h->ev0, h->ev1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g_err = std::string("timed steps: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  int st[2];
  HIP_TRY(copy_on(h->stream, st, h->status, sizeof(st), hipMemcpyDeviceToHost));
  h->step += st[1];
  double tr[2];
  HIP_TRY(copy_on(h->stream, tr, h->tring, sizeof(tr), hipMemcpyDeviceToHost));
  h->t = tr[h->step & 1];
  if (st[0]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

// Device-to-device copy bandwidth (bytes read + written per second) of a plain
// streaming copy of `bytes` bytes, best of `reps`: the measured second
// denominator of the HBM roofline.  One double per thread with a nontemporal
// load and no loop: of the copy forms tried on MI355X (grid-stride loops with
// 1..8 loads in flight per lane, 16-byte elements) this one moves the most
// bytes per second.
__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = __builtin_nontemporal_load(src + i);
}

int hes_copy_bandwidth(long long bytes, int reps, double* gbs) {
  if (!gbs || bytes < 1024 || reps < 1) return fail(HES_ERR_ARG, "bad argument");
  const long long n = bytes / 8;
  const long long nblocks = (n + 255) / 256;
  if (nblocks > 0x7fffffffLL) return fail(HES_ERR_ARG, "copy size too large");
  double *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, n * 8));
  hipError_t e = hipMalloc(&b, n * 8);
  hipEvent_t e0, e1;
  if (e == hipSuccess) e = hipMemset(a, 0, n * 8);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float best = 1e30f;
  for (int r = 0; r <= reps && e == hipSuccess; ++r) {
    e = hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nblocks), dim3(256), 0, 0, a, b, n);
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && ms < best) best = ms;  // first launch is a warm-up
  }
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) {
    g_err = std::string("copy bandwidth: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  *gbs = 2.0 * (double)n * 8.0 / (best * 1e-3) / 1e9;
  return HES_OK;
}

// ---------------------------------------------------------------------------
// Point-wise API (host arrays, synchronous)
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
  double* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(double) + 8); }
};

struct DevStatus {
  int* p = nullptr;
  ~DevStatus() { if (p) (void)hipFree(p); }
};

int finish_status(DevStatus& st) {
  int v = 0;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(&v, st.p, sizeof(int), hipMemcpyDeviceToHost));
  if (v) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

int point_call(int op, int nph, const double* mats, const double* in, double* out, long long n,
               int in_w, int out_w, const double* nvec = nullptr) {
  if (nph != 1 && nph != 2) return fail(HES_ERR_ARG, "nphase must be 1 or 2");
  if (!mats || !in || !out || n < 0) return fail(HES_ERR_ARG, "null argument");
  for (int p = 0; p < nph; ++p)
    if (!mat_ok(mats + 8 * p)) return fail(HES_ERR_ARG, "invalid material constants");
  if (n == 0) return HES_OK;
  MatDev m0 = make_mat(mats), m1 = make_mat(nph == 2 ? mats + 8 : mats);
  DevBuf di, dout;
  DevStatus st;
  HIP_TRY(di.alloc((size_t)n * in_w));
  HIP_TRY(dout.alloc((size_t)n * out_w));
  HIP_TRY(hipMalloc(&st.p, sizeof(int)));
  HIP_TRY(hipMemset(st.p, 0, sizeof(int)));
  HIP_TRY(hipMemcpy(di.p, in, sizeof(double) * (size_t)n * in_w, hipMemcpyHostToDevice));
  unsigned nb = (unsigned)((n + 63) / 64);
  double nv[3] = {1.0, 0.0, 0.0};
  if (nvec) { nv[0] = nvec[0]; nv[1] = nvec[1]; nv[2] = nvec[2]; }
  if (nph == 1)
    hipLaunchKernelGGL(k_point<1>, dim3(nb), dim3(64), 0, 0, op, di.p, dout.p, n, m0, m1, nv[0],
                       nv[1], nv[2], st.p);
  else
    hipLaunchKernelGGL(k_point<2>, dim3(nb), dim3(64), 0, 0, op, di.p, dout.p, n, m0, m1, nv[0],
                       nv[1], nv[2], st.p);
  HIP_TRY(hipGetLastError());
  int rc = finish_status(st);
  HIP_TRY(hipMemcpy(out, dout.p, sizeof(double) * (size_t)n * out_w, hipMemcpyDeviceToHost));
  return rc;
}

int eos_call(int op, const double* mat, const double* s, const double* s2, const double* v,
             int vw, const double* nv, double* out, int ow, long long n) {
  if (!mat || !out || n < 0) return fail(HES_ERR_ARG, "null argument");
  if (!mat_ok(mat)) return fail(HES_ERR_ARG, "invalid material constants");
  if (n == 0) return HES_OK;
  MatDev m = make_mat(mat);
  DevBuf ds, ds2, dv, dn, dout;
  DevStatus st;
  HIP_TRY(hipMalloc(&st.p, sizeof(int)));
  HIP_TRY(hipMemset(st.p, 0, sizeof(int)));
  if (s) { HIP_TRY(ds.alloc(n)); HIP_TRY(hipMemcpy(ds.p, s, sizeof(double) * n, hipMemcpyHostToDevice)); }
  if (s2) { HIP_TRY(ds2.alloc(n)); HIP_TRY(hipMemcpy(ds2.p, s2, sizeof(double) * n, hipMemcpyHostToDevice)); }
  if (v) { HIP_TRY(dv.alloc((size_t)n * vw)); HIP_TRY(hipMemcpy(dv.p, v, sizeof(double) * n * vw, hipMemcpyHostToDevice)); }
  if (nv) { HIP_TRY(dn.alloc((size_t)n * 3)); HIP_TRY(hipMemcpy(dn.p, nv, sizeof(double) * n * 3, hipMemcpyHostToDevice)); }
  HIP_TRY(dout.alloc((size_t)n * ow));
  unsigned nb = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(k_eos, dim3(nb), dim3(64), 0, 0, op, m, ds.p, ds2.p, dv.p, dn.p, dout.p, n, st.p);
  HIP_TRY(hipGetLastError());
  int rc = finish_status(st);
  HIP_TRY(hipMemcpy(out, dout.p, sizeof(double) * (size_t)n * ow, hipMemcpyDeviceToHost));
  return rc;
}

}  // namespace

int hes_cons2prim(int nph, const double* mats, const double* Q, double* P, long long n) {
  int w = nph == 1 ? 14 : 30;
  return point_call(OP_CONS2PRIM, nph, mats, Q, P, n, w, w);
}
int hes_prim2cons(int nph, const double* mats, const double* P, double* Q, long long n) {
  int w = nph == 1 ? 14 : 30;
  return point_call(OP_PRIM2CONS, nph, mats, P, Q, n, w, w);
}
int hes_flux(int nph, const double* mats, const double* Q, double* F, long long n) {
  int w = nph == 1 ? 14 : 30;
  return point_call(OP_FLUX, nph, mats, Q, F, n, w, w);
}
int hes_get_eigvals(int nph, const double* mats, const double* Q, const double* nvec, double* eig,
                    long long n) {
  int w = nph == 1 ? 14 : 30;
  return point_call(OP_EIGVALS, nph, mats, Q, eig, n, w, 6 * nph, nvec);
}
int hes_max_speed(int nph, const double* mats, const double* Q, double* lam, long long n) {
  int w = nph == 1 ? 14 : 30;
  return point_call(OP_MAXSPEED, nph, mats, Q, lam, n, w, 1);
}
int hes_noncons_flux(const double* mats, const double* Q, double* B, long long n) {
  return point_call(OP_NONCONS, 2, mats, Q, B, n, 30, 900);
}

int hes_energy(const double* mat, const double* S, const double* G, double* e, long long n) {
  if (!S || !G) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_ENERGY, mat, S, nullptr, G, 9, nullptr, e, 1, n);
}
int hes_entropy(const double* mat, const double* e_int, const double* G, double* S, long long n) {
  if (!e_int || !G) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_ENTROPY, mat, e_int, nullptr, G, 9, nullptr, S, 1, n);
}
int hes_temperature(const double* mat, const double* S, const double* G, double* T, long long n) {
  if (!S || !G) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_TEMPERATURE, mat, S, nullptr, G, 9, nullptr, T, 1, n);
}
int hes_stress(const double* mat, const double* ent, const double* F, double* sig, long long n) {
  if (!ent || !F) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_STRESS, mat, ent, nullptr, F, 9, nullptr, sig, 9, n);
}
int hes_stress_den(const double* mat, const double* den, const double* e_int, const double* F,
                   double* sig, long long n) {
  if (!den || !e_int || !F) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_STRESS_DEN, mat, den, e_int, F, 9, nullptr, sig, 9, n);
}
int hes_acoustic(const double* mat, const double* ent, const double* F, const double* nvec,
                 double* ac, long long n) {
  if (!ent || !F || !nvec) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_ACOUSTIC, mat, ent, nullptr, F, 9, nvec, ac, 9, n);
}
int hes_acoustic_prim(const double* mat, const double* P, const double* nvec, double* ac,
                      long long n) {
  if (!P || !nvec) return fail(HES_ERR_ARG, "null argument");
  return eos_call(EOS_ACOUSTIC_PRIM, mat, nullptr, nullptr, P, 15, nvec, ac, 9, n);
}
int hes_finger(const double* A, double* G, long long n) {
  if (!A) return fail(HES_ERR_ARG, "null argument");
  static const double dummy[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  return eos_call(EOS_FINGER, dummy, nullptr, nullptr, A, 9, nullptr, G, 9, n);
}
int hes_invariants(const double* G, double* inv, long long n) {
  if (!G) return fail(HES_ERR_ARG, "null argument");
  static const double dummy[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  return eos_call(EOS_INVARIANTS, dummy, nullptr, nullptr, G, 9, nullptr, inv, 3, n);
}

int hes_face_flux(int nph, int flux, const double* mats, const double* QL, const double* QR,
                  const double* aux, double* cons, double* dm, double* dp, long long n) {
  if (nph != 1 && nph != 2) return fail(HES_ERR_ARG, "nphase must be 1 or 2");
  if (flux == HES_FLUX_LXF) flux = HES_FLUX_RUSANOV;  // one face: the same lxf, lambda given in aux
  if (flux != HES_FLUX_RUSANOV && flux != HES_FLUX_HLL) return fail(HES_ERR_ARG, "unknown flux");
  if (!mats || !QL || !QR || !aux || !cons || !dm || !dp || n < 0)
    return fail(HES_ERR_ARG, "null argument");
  for (int p = 0; p < nph; ++p)
    if (!mat_ok(mats + 8 * p)) return fail(HES_ERR_ARG, "invalid material constants");
  if (n == 0) return HES_OK;
  int w = nph == 1 ? 14 : 30;
  int aw = flux == HES_FLUX_HLL ? 2 : 1;
  MatDev m0 = make_mat(mats), m1 = make_mat(nph == 2 ? mats + 8 : mats);
  DevBuf dl, dr, da, dc, ddm, ddp;
  HIP_TRY(dl.alloc((size_t)n * w));
  HIP_TRY(dr.alloc((size_t)n * w));
  HIP_TRY(da.alloc((size_t)n * aw));
  HIP_TRY(dc.alloc((size_t)n * w));
  HIP_TRY(ddm.alloc((size_t)n * w));
  HIP_TRY(ddp.alloc((size_t)n * w));
  HIP_TRY(hipMemcpy(dl.p, QL, sizeof(double) * n * w, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr.p, QR, sizeof(double) * n * w, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(da.p, aux, sizeof(double) * n * aw, hipMemcpyHostToDevice));
  unsigned nb = (unsigned)((n + 63) / 64);
  if (nph == 1 && flux == HES_FLUX_RUSANOV)
    hipLaunchKernelGGL((k_face<1, FLUX_RUSANOV>), dim3(nb), dim3(64), 0, 0, dl.p, dr.p, da.p, dc.p, ddm.p, ddp.p, n, m0, m1);
  else if (nph == 1)
    hipLaunchKernelGGL((k_face<1, FLUX_HLL>), dim3(nb), dim3(64), 0, 0, dl.p, dr.p, da.p, dc.p, ddm.p, ddp.p, n, m0, m1);
  else if (flux == HES_FLUX_RUSANOV)
    hipLaunchKernelGGL((k_face<2, FLUX_RUSANOV>), dim3(nb), dim3(64), 0, 0, dl.p, dr.p, da.p, dc.p, ddm.p, ddp.p, n, m0, m1);
  else
    hipLaunchKernelGGL((k_face<2, FLUX_HLL>), dim3(nb), dim3(64), 0, 0, dl.p, dr.p, da.p, dc.p, ddm.p, ddp.p, n, m0, m1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(cons, dc.p, sizeof(double) * n * w, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dm, ddm.p, sizeof(double) * n * w, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dp, ddp.p, sizeof(double) * n * w, hipMemcpyDeviceToHost));
  return HES_OK;
}

// ---- Hyperelasticity (the single-phase module, Hyperelasticity.jl:25-114):
// 13-component states [rho u (3), rho F (9, row-major), rho E], converted on
// the host to and from the library's one-phase layout [rho, rho u, rho E,
// rho vec(A)] (vec column-major); the physics runs in the point-wise kernels.
namespace {

inline int sp_index(int i, int j) { return 3 + 3 * i + j; }  // rho F[i, j] in the 13-vector

void sp_to14(const double* q, double* o) {
  o[0] = 0.0;  // rho: cons2prim forms it from det(rho F) and never reads this slot
  o[1] = q[0];
  o[2] = q[1];
  o[3] = q[2];
  o[4] = q[12];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[5 + i + 3 * j] = q[sp_index(i, j)];
}

void sp_from14(const double* o, double* q) {
  q[0] = o[1];
  q[1] = o[2];
  q[2] = o[3];
  q[12] = o[4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) q[sp_index(i, j)] = o[5 + i + 3 * j];
}

}  // namespace

int hes_sp_cons2prim(const double* mat, const double* Q, double* P, long long n) {
  if (!mat || n < 0 || (n > 0 && (!Q || !P))) return fail(HES_ERR_ARG, "null argument");
  if (n == 0) return HES_OK;
  std::vector<double> q14((size_t)14 * n), p14((size_t)14 * n);
  for (long long k = 0; k < n; ++k) sp_to14(Q + 13 * k, q14.data() + 14 * k);
  const int rc = hes_cons2prim(1, mat, q14.data(), p14.data(), n);
  if (rc != HES_OK && rc != HES_ERR_DOMAIN) return rc;
  for (long long k = 0; k < n; ++k) {
    const double* pk = p14.data() + 14 * k;
    double* o = P + 14 * k;
    const double den = pk[0];
    o[0] = den;
    for (int r = 0; r < 3; ++r) o[1 + r] = pk[1 + r];
    for (int r = 0; r < 9; ++r) o[4 + r] = pk[5 + r];  // vec(F), F = A
    // e_int = E - |u|^2 / 2 with E = rho E / den (Hyperelasticity.jl:32-33)
    o[13] = Q[13 * k + 12] / den - 0.5 * (pk[1] * pk[1] + pk[2] * pk[2] + pk[3] * pk[3]);
  }
  return rc;
}

int hes_sp_prim2cons(const double* mat, const double* P, double* Q, long long n) {
  if (!mat || n < 0 || (n > 0 && (!Q || !P))) return fail(HES_ERR_ARG, "null argument");
  if (n == 0) return HES_OK;
  std::vector<double> p14((size_t)14 * n), q14((size_t)14 * n);
  for (long long k = 0; k < n; ++k) {
    const double* pk = P + 13 * k;  // [vel (3), vec(F) (9), S]
    double* o = p14.data() + 14 * k;
    const double* F = pk + 3;
    const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
                        F[6] * (F[1] * F[5] - F[4] * F[2]);
    o[0] = mat[0] / detF;  // den = rho0 / det F (Hyperelasticity.jl:83)
    for (int r = 0; r < 3; ++r) o[1 + r] = pk[r];
    o[4] = pk[12];
    for (int r = 0; r < 9; ++r) o[5 + r] = F[r];
  }
  const int rc = hes_prim2cons(1, mat, p14.data(), q14.data(), n);
  if (rc != HES_OK) return rc;
  for (long long k = 0; k < n; ++k) sp_from14(q14.data() + 14 * k, Q + 13 * k);
  return HES_OK;
}

int hes_sp_flux(const double* mat, const double* Q, double* F, long long n) {
  if (!mat || n < 0 || (n > 0 && (!Q || !F))) return fail(HES_ERR_ARG, "null argument");
  if (n == 0) return HES_OK;
  std::vector<double> q14((size_t)14 * n), f14((size_t)14 * n);
  for (long long k = 0; k < n; ++k) sp_to14(Q + 13 * k, q14.data() + 14 * k);
  const int rc = hes_flux(1, mat, q14.data(), f14.data(), n);
  if (rc != HES_OK && rc != HES_ERR_DOMAIN) return rc;
  for (long long k = 0; k < n; ++k) sp_from14(f14.data() + 14 * k, F + 13 * k);
  return rc;
}

// ---- One phase block of HyperelasticityMPh (15 components with the volume
// fraction): the per-phase methods cons2prim / flux / get_eigvals(eos::T, Q)
// (HyperelasticityMPh.jl:106-133, 146-175, 258-266).  A phase's values do
// not depend on the other phase, so each block is evaluated as both phases
// of a two-phase state and the first block of the result is returned.
namespace {

int phase_call(int op, const double* mat, const double* Q, double* out, long long n, int ow,
               const double* nvec) {
  if (!mat || n < 0 || (n > 0 && (!Q || !out))) return fail(HES_ERR_ARG, "null argument");
  if (n == 0) return HES_OK;
  double mm[16];
  for (int r = 0; r < 8; ++r) mm[r] = mm[8 + r] = mat[r];
  std::vector<double> q30((size_t)30 * n), o2((size_t)2 * ow * n);
  for (long long k = 0; k < n; ++k)
    for (int r = 0; r < 15; ++r) q30[30 * k + r] = q30[30 * k + 15 + r] = Q[15 * k + r];
  const int rc = point_call(op, 2, mm, q30.data(), o2.data(), n, 30, 2 * ow, nvec);
  if (rc != HES_OK && rc != HES_ERR_DOMAIN) return rc;
  for (long long k = 0; k < n; ++k)
    for (int r = 0; r < ow; ++r) out[(size_t)ow * k + r] = o2[(size_t)2 * ow * k + r];
  return rc;
}

}  // namespace

int hes_phase_cons2prim(const double* mat, const double* Q, double* P, long long n) {
  return phase_call(OP_CONS2PRIM, mat, Q, P, n, 15, nullptr);
}
int hes_phase_flux(const double* mat, const double* Q, double* F, long long n) {
  return phase_call(OP_FLUX, mat, Q, F, n, 15, nullptr);
}
int hes_phase_get_eigvals(const double* mat, const double* Q, const double* nvec, double* eig,
                          long long n) {
  return phase_call(OP_EIGVALS, mat, Q, eig, n, 6, nvec);
}

}  // extern "C"
