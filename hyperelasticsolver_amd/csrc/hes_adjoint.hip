// hes_adjoint.hip -- reverse mode of a fixed-length single-phase Rusanov run:
// the gradient of a final-state misfit
//     L = sum_i (rho_i - rho*_i)^2 + (u1_i - u1*_i)^2
// with respect to the primitive initial states (u, A, S of each constant
// piece) and the material constants b0 (shear speed) and c0 (bulk speed).
//
// Forward model, step s -> s+1 (main.jl:204-225 with NumFluxes.jl:25-60):
//   lam_i  = |u1_i| + c_i                 (max |eigenvalue|, c = sqrt rho(Omega))
//   lmax   = max_i lam_i,  dtdx = cfl / lmax (dt = cfl dx / lmax)
//   G_f    = 1/2 (F_i + F_i+1) - 1/2 max(lam_i, lam_i+1) (Q_i+1 - Q_i)
//   Q_i   <- Q_i - dtdx (G_i+1/2 - G_i-1/2)      (interior cells; the end cells stay)
//
// Reverse step, with W_i = Qbar_i of state s+1 on interior cells (0 on the
// end cells) and g_f = W_right(f) - W_left(f):
//   Qbar_i += dtdx/2 (s_R g_R - s_L g_L) + [dQ of the cell VJP]
//   Fbar_i  = dtdx/2 (g_L + g_R)
//   sbar_f  = -dtdx/2 g_f . (Q_i+1 - Q_i)   -> lam of the larger side (ties: left cell)
//   lambar  -> u1bar = sign(u1) lambar,  Omegabar = lambar / (2c) v v^T  (v: top eigenvector)
//   dtdxbar = sum_i F_i . (g_L + g_R)/2 - sum_f s_f/2 (Q_i+1 - Q_i) . g_f
//   lmaxbar = -dtdxbar dtdx / lmax          -> lam of the argmax cell (smallest index)
// The per-cell VJP (flux, acoustic tensor, velocity of a cell as functions of
// its conserved state and of b0, c0) is generated code (hes_adjoint_gen.hpp).
//
// Storage: every `every`-th state is kept (with its sound speeds); a backward
// segment recomputes the states in between with the forward kernel (bit-identical
// to the first pass), then runs the reverse steps over them.  The recompute of
// the next (earlier) segment runs on a second stream into the other of two
// segment buffers while the reverse steps of the current one run.  The lmax term of
// step s needs the global sum dtdxbar_s, so it is applied when step s-1 loads
// Qbar_s: the one or two waves holding the argmax cell reduce the block
// partials of step s and add lmaxbar_s * dlam/d(Q, b0, c0) of that cell
// (computed once per segment).
// the secular-function eigenvalue iteration throughout (see sym3_extreme);
// set before any include
#define HES_EIG_SECULAR 1
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hes_host.hpp"
#include "hes_adjoint_common.hpp"
#include "../../include/hes.h"

using namespace hes;
using namespace hes_adj;

namespace {

#ifndef HES_ADJ_BLOCK
#define HES_ADJ_BLOCK 256
#endif
constexpr int ABLK = HES_ADJ_BLOCK;  // threads per block of the tiled kernels
constexpr int LBLK = 256;  // threads per block of the one-cell-per-thread kernels

// ---------------------------------------------------------------------------
// Forward kernels
// ---------------------------------------------------------------------------

// Conserved states of the constant pieces from their primitives.
__global__ void k_adj_p2c(const double* __restrict__ prims, const double* __restrict__ gin,
                          double* __restrict__ states, int nst) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nst) return;
  double in[hes_gen::P2C_IN];
#pragma unroll
  for (int r = 0; r < NP; ++r) in[r] = prims[NP * k + r];
#pragma unroll
  for (int r = 0; r < 5; ++r) in[NP + r] = gin[r];
  double out[NQ];
  hes_gen::prim2cons_fwd(in, out);
#pragma unroll
  for (int r = 0; r < NQ; ++r) states[NQ * k + r] = out[r];
}

// Piecewise-constant fill: cells [bounds[k], bounds[k+1]) take state k.
__global__ void k_adj_fill(const double* __restrict__ states, const long long* __restrict__ bounds,
                           int nst, double* __restrict__ Q, long long n, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = 0;
  while (k + 1 < nst && i >= bounds[k + 1]) ++k;
#pragma unroll
  for (int r = 0; r < NQ; ++r) Q[r * pitch + i] = states[NQ * k + r];
}

// The end cells never change: copy them from src into every state slot.
__global__ void k_adj_ends(const double* __restrict__ src, double* const* __restrict__ dst,
                           int ndst, long long n, long long pitch) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndst * 2 * NQ) return;
  int slot = t / (2 * NQ), rem = t % (2 * NQ);
  int r = rem >> 1;
  long long i = (rem & 1) ? n - 1 : 0;
  dst[slot][r * pitch + i] = src[r * pitch + i];
}

__global__ void k_adj_zero(double* __restrict__ p, long long m) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) p[t] = 0.0;
}

// Sound speeds of state 0 and its max CFL speed (main.jl:204-212).
__global__ __launch_bounds__(LBLK) void k_adj_speeds(const double* __restrict__ Q,
                                                     double* __restrict__ C, long long n,
                                                     long long pitch, MatDev m,
                                                     unsigned long long* lmax0, double* bv,
                                                     long long* bi, int* status) {
  long long i = (long long)blockIdx.x * LBLK + threadIdx.x;
  double lam = 0.0;
  int bad = 0;
  if (i < n) {
    double q[NQ], cc[1];
#pragma unroll
    for (int r = 0; r < NQ; ++r) q[r] = Q[r * pitch + i];
    lam = cell_speeds<1, 1>(&m, q, cc, bad);
    C[i] = cc[0];
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1);
  record_speeds<LBLK>(lam, i < n, i, lmax0, bv, bi);
}

struct AdjFwdArgs {
  const double* __restrict__ Qin;
  double* __restrict__ Qout;
  const double* __restrict__ Cin;
  double* __restrict__ Cout;
  unsigned long long* lmax;  // lmax[s] = max CFL speed of state s (double bits)
  double* blk_val;           // FWD_RECORD: per-block argmax of the CFL speed of state s + 1
  long long* blk_idx;
  long long* istar;          // FWD_FIND: argmax cell of every state
  int* status;
  long long n, pitch, s;
  double cfl, dx;
  MatDev mat;
};

// What a forward step records of the new state s + 1's CFL speeds:
//   FWD_MAX     the first pass: their max, into lmax[s + 1]
//   FWD_RECORD  the first pass at a checkpoint state: also the per-block argmax
//   FWD_FIND    a recompute of the reverse pass: the argmax cell (smallest
//               index) of the max that the first pass recorded, into istar[s + 1]
// The argmax of the states between checkpoints is found while the reverse
// pass recomputes them, so the first pass does no more than a plain run's max.
enum { FWD_MAX = 0, FWD_RECORD = 1, FWD_FIND = 2 };

// One time step: the solver's chunk step (so states match a plain run bit for
// bit) plus the bookkeeping of the reverse pass.
template <int MODE>
__global__ __launch_bounds__(ABLK) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_fwd(
    AdjFwdArgs a) {
  const int lane = threadIdx.x & 63;
  const long long wave = (xcd_block() * ABLK + threadIdx.x) >> 6;
  const long long i = wave * TILE + lane - 1;
  const ChunkCtx ctx = {&a.mat, a.Qin, a.Qout, a.Cin, a.Cout, a.status, a.pitch, 0, a.n, a.cfl, a.dx};
  double q[NQ], cc[1];
  load_chunk<1>(ctx, i, q, cc);
  const double lmax = lmax_of(a.lmax, a.s);
  const double lam = chunk_step<1, FLUX_RUSANOV, 1>(ctx, i, q, cc, lmax);
  const bool owned = i >= 0 && i < a.n && lane >= 1 && lane <= TILE;
  if constexpr (MODE == FWD_FIND) {
    // the recomputed speeds are those of the first pass, bit for bit
    if (owned && lam == lmax_of(a.lmax, a.s + 1))
      atomicMin((unsigned long long*)(a.istar + a.s + 1), (unsigned long long)i);
  } else if constexpr (MODE == FWD_RECORD) {
    const long long row = (a.s + 1) * (long long)gridDim.x;
    record_speeds<ABLK>(lam, owned, i, a.lmax + a.s + 1, a.blk_val + row, a.blk_idx + row);
  } else {
    // one atomic per wave (most find a larger maximum already there)
    const double m = wave_max(owned ? lam : -1.0);
    if (lane == 0 && m > 0.0) {
      unsigned long long* slot = a.lmax + a.s + 1;
      const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
      if (bits > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(slot, bits);
    }
  }
}

// Argmax cell of the CFL speed of the checkpoint states from the per-block
// records (state 0: the per-cell speeds kernel's); the other states' entries
// are set to NO_CELL, for the reverse pass's recomputes to min-reduce into.
__global__ __launch_bounds__(64) void k_adj_argmax(const double* __restrict__ blk_val,
                                                   const long long* __restrict__ blk_idx, int nb,
                                                   int nb0, long long every,
                                                   long long* __restrict__ istar) {
  const long long s = blockIdx.x;
  const int lane = threadIdx.x;
  if (s % every != 0) {
    if (lane == 0) istar[s] = NO_CELL;
    return;
  }
  const int cnt = s == 0 ? nb0 : nb;  // state 0 was recorded by the per-cell speeds kernel
  double v = -2.0;
  long long i = NO_CELL;
  for (int b = lane; b < cnt; b += 64) arg_better(v, i, blk_val[s * nb + b], blk_idx[s * nb + b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o);
    long long oi = __shfl_xor(i, o);
    arg_better(v, i, ov, oi);
  }
  if (lane == 0) istar[s] = i;
}

// Misfit of the final state and its gradient Qbar_n (HyperelasticityMPh.jl:106-121
// gives rho and u1 of a cell).
__global__ __launch_bounds__(LBLK) void k_adj_loss(const double* __restrict__ Q,
                                                   const double* __restrict__ rho_t,
                                                   const double* __restrict__ u_t, double rho0,
                                                   double* __restrict__ Qb, double* __restrict__ part,
                                                   long long n, long long pitch) {
  long long i = (long long)blockIdx.x * LBLK + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    // rho and u1 by the routine that also makes target profiles, so a cell
    // where the runs agree has a misfit and a gradient of exactly zero
    double in[hes_gen::PROF_IN], o[2];
#pragma unroll
    for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
    in[NQ] = rho0;
    hes_gen::profiles_fwd(in, o);
    const double d1 = o[0] - rho_t[i], d2 = o[1] - u_t[i];
    l = d1 * d1 + d2 * d2;
    double ob[2] = {2.0 * d1, 2.0 * d2}, ib[hes_gen::PROF_IN];
#pragma unroll
    for (int r = 0; r < hes_gen::PROF_IN; ++r) ib[r] = 0.0;
    hes_gen::profiles_vjp(in, ob, ib);
#pragma unroll
    for (int r = 0; r < NQ; ++r) Qb[r * pitch + i] = ib[r];
  }
  double s = block_sum<LBLK>(l);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// rho and u1 of every cell with the misfit's own expressions (a target taken
// from a run's profiles then matches it exactly wherever the runs agree).
__global__ __launch_bounds__(LBLK) void k_adj_profiles(const double* __restrict__ Q, double rho0,
                                                       double* __restrict__ rho,
                                                       double* __restrict__ u1, long long n,
                                                       long long pitch) {
  long long i = (long long)blockIdx.x * LBLK + threadIdx.x;
  if (i >= n) return;
  double in[hes_gen::PROF_IN], o[2];
#pragma unroll
  for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
  in[NQ] = rho0;
  hes_gen::profiles_fwd(in, o);
  rho[i] = o[0];
  u1[i] = o[1];
}

// ---------------------------------------------------------------------------
// Reverse kernels
// ---------------------------------------------------------------------------

// dlam/d(Q, b0, c0) of the argmax cell of each state s0 .. s0+cnt-1 (lane j
// handles state s0 + j; its state lives in Q0 for j = 0, else in
// seg + (j - 1) * stride).
__global__ __launch_bounds__(64) void k_adj_lam(const double* __restrict__ Q0,
                                                const double* __restrict__ seg, long long stride,
                                                long long pitch, const long long* __restrict__ istar,
                                                const double* __restrict__ gin, long long s0,
                                                int cnt, long long n, double* __restrict__ Jlam) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= cnt) return;
  const long long s = s0 + j;
  const long long i = istar[s];
  if (i < 0 || i >= n) {  // no finite CFL speed was recorded: no lmax term
#pragma unroll
    for (int r = 0; r < JL; ++r) Jlam[s * JL + r] = 0.0;
    return;
  }
  const double* Q = j == 0 ? Q0 : seg + (long long)(j - 1) * stride;
  double in[hes_gen::CELL_IN];
#pragma unroll
  for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
#pragma unroll
  for (int r = 0; r < 5; ++r) in[NQ + r] = gin[r];
  double out[hes_gen::CELL_OUT];
  hes_gen::cell_fwd(in, out);
  double ob[hes_gen::CELL_OUT], ib[hes_gen::CELL_IN];
#pragma unroll
  for (int r = 0; r < NQ; ++r) ob[r] = 0.0;
  speed_and_grad(out + NQ, ob + NQ);
  ob[20] = out[20] < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int r = 0; r < hes_gen::CELL_IN; ++r) ib[r] = 0.0;
  hes_gen::cell_vjp(in, ob, ib);
#pragma unroll
  for (int r = 0; r < JL; ++r) Jlam[s * JL + r] = ib[r];
}

struct AdjStepArgs {
  const double* __restrict__ Q;      // state s
  const double* __restrict__ Qb_in;  // Qbar of state s+1 (without its lmax term)
  double* __restrict__ Qb_out;       // Qbar of state s (without its lmax term)
  const int* zin;                    // per tile: 1 if Qb_in is zero on all its owned cells
  int* zout;                         // the same for Qb_out
  const unsigned long long* lmax;
  const long long* istar;
  const double* Jlam;
  const double* dtp_next;  // block partials of dtdxbar of step s+1
  double* dtp;             // block partials of dtdxbar of step s
  double* thb;             // running (b0bar, c0bar) per block
  long long n, pitch, s;
  double cfl, dx;
  int nb, nw, has_next;
  int sparse;              // skip tiles whose cotangent is zero on them and their halo
  double gin[5];           // b0, c0, rho0, cv, t0
};

// One reverse step.  A tile whose Qbar of state s+1 is zero on its owned
// cells and both halo cells (and which holds no argmax cell of step s+1) has
// Qbar of state s = 0 and contributes nothing to dtdxbar or (b0bar, c0bar):
// every seed of its cell VJPs is zero, and the VJP is linear in its seeds.
#ifndef HES_ADJ_WAVES_PER_EU
#define HES_ADJ_WAVES_PER_EU 1
#endif

constexpr int PROWS = 2 * NQ + 1;  // rows of a tile in LDS: Q (14), its sound speed, Qbar_{s+1} (14)
typedef __attribute__((address_space(3))) void lds_t;

// Copies the PROWS rows of tile w (cells w*TILE-1 .. w*TILE+62, clamped to
// [0, n)) from HBM into a wave's LDS buffer at byte address `lds`, without
// passing through registers: each instruction moves 4 bytes per lane, lane l
// taking half l&1 of cell 32h + l/2 of the tile, so a row takes two
// instructions and lands in doubles 64 r .. 64 r + 63 of the buffer in cell
// order.  Byte offsets within a row are 32-bit (n < 2^28 is checked at
// creation).
__device__ __forceinline__ void tile_prefetch(const AdjStepArgs& a, long long w, int lane,
                                              unsigned lds) {
  const long long i0 = w * TILE - 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    long long c = i0 + 32 * h + (lane >> 1);
    c = c < 0 ? 0 : (c >= a.n ? a.n - 1 : c);
    unsigned off = (unsigned)c * 8u + (unsigned)(lane & 1) * 4u;
    // fresh values per call: the row addresses are formed here, one at a
    // time, rather than hoisted out of the tile loop into 58 registers
    const char* q = (const char*)a.Q;
    const char* qb = (const char*)a.Qb_in;
    unsigned d = lds + 256u * h;
    asm volatile("" : "+v"(off), "+s"(q), "+s"(qb), "+s"(d));
    const unsigned long long rowb = (unsigned long long)a.pitch * 8ull;
#pragma unroll
    for (int r = 0; r < PROWS; ++r) {
      const char* src = r <= NQ ? q + r * rowb : qb + (r - NQ - 1) * rowb;
      __builtin_amdgcn_global_load_lds((const void*)(src + off), (lds_t*)(size_t)(d + 512u * r),
                                       4, 0, 0);
    }
  }
}

// Persistent waves, one per SIMD (a wave holds 256 VGPRs + 76 AGPRs, so a
// forward wave of the recompute fits beside it): wave k of the grid takes
// tiles k, k + W, k + 2W, ...  While it evaluates one tile, the next one it
// will evaluate streams into its LDS buffer, so only the first tile's loads
// are waited for.  The block partials of dtdxbar and (b0bar, c0bar) sum the
// block's tiles in a fixed order.
__global__ __launch_bounds__(ABLK) __attribute__((amdgpu_waves_per_eu(HES_ADJ_WAVES_PER_EU)))
void k_adj_step(AdjStepArgs a) {
  __shared__ double sbuf[ABLK / 64][PROWS * 64];
  const int lane = threadIdx.x & 63;
  const int wl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double* buf = sbuf[wl];
  const unsigned lds = (unsigned)(size_t)(lds_t*)buf;
  const long long W = (long long)gridDim.x * (ABLK / 64);
  const long long pitch = a.pitch;
  const long long is1 = a.has_next ? a.istar[a.s + 1] : -2;
  // The first tile from w on (in steps of W) that must be evaluated; the
  // skipped ones on the way get their zero cotangent.  The zero flags of
  // Qbar_{s+1} are read-only in this kernel: read through the constant
  // address space, they load into scalar registers, and no wait for them
  // waits for vector memory.
  typedef const __attribute__((address_space(4))) int ci32_t;
  ci32_t* zin = (ci32_t*)a.zin;
  auto next_tile = [&](long long w) -> long long {
    for (; w < a.nw; w += W) {
      const bool hit = is1 >= w * TILE - 1 && is1 <= w * TILE + TILE;
      if (!a.sparse || hit) return w;
      const bool skip = zin[w] && (w == 0 || zin[w - 1]) && (w + 1 >= a.nw || zin[w + 1]);
      if (!skip) return w;
      const long long i = w * TILE + lane - 1;
      if (!a.zout[w] && i >= 0 && i < a.n && lane >= 1 && lane <= TILE) {
#pragma unroll
        for (int r = 0; r < NQ; ++r) a.Qb_out[r * pitch + i] = 0.0;
      }
      if (lane == 0) a.zout[w] = 1;
    }
    return w;
  };
  double dts = 0.0, th0 = 0.0, th1 = 0.0;
  const double dtdx = a.cfl * rcp(lmax_of(a.lmax, a.s));  // as chunk_step forms it
  long long w = next_tile(xcd_block() * (ABLK / 64) + wl);
  if (w < a.nw) tile_prefetch(a, w, lane, lds);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the first tile is in LDS
  while (w < a.nw) {
    const long long i = w * TILE + lane - 1;
    const bool inr = i >= 0 && i < a.n;
    const bool owned = inr && lane >= 1 && lane <= TILE;
    const bool hit = is1 >= w * TILE - 1 && is1 <= w * TILE + TILE;  // wave-uniform
    double q[NQ], qb[NQ];
#pragma unroll
    for (int r = 0; r < NQ; ++r) {
      q[r] = buf[64 * r + lane];
      qb[r] = inr ? buf[64 * (NQ + 1 + r) + lane] : 0.0;
    }
    // the state's sound speeds, stored in row 14 of its slot by the forward step
    const double cs = buf[64 * NQ + lane];
    // every lane's reads are in registers before the next tile overwrites the buffer
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    const long long wn = next_tile(w + W);
    if (wn < a.nw) tile_prefetch(a, wn, lane, lds);
    // lmax term of step s+1 at its argmax cell
    if (hit) {
      double acc = 0.0;
      for (int b = lane; b < a.nb; b += 64) acc += a.dtp_next[b];
      acc = wave_sum(acc);
      const double lm1 = lmax_of(a.lmax, a.s + 1);
      const double dtdx1 = a.cfl * rcp(lm1);  // as chunk_step forms dt / dx
      const double lbar = -acc * dtdx1 * rcp(lm1);
      if (inr && i == is1) {
        const double* J = a.Jlam + (a.s + 1) * JL;
#pragma unroll
        for (int r = 0; r < NQ; ++r) qb[r] += lbar * J[r];
        if (owned) {
          th0 += lbar * J[NQ];
          th1 += lbar * J[NQ + 1];
        }
      }
    }
    const bool interior = inr && i > 0 && i < a.n - 1;
    double ib[hes_gen::CELL_IN];
    double dtt = 0.0;
    reverse_chunk(q, qb, interior, owned, cs, 0.5 * dtdx, a.gin, ib, dtt);
    dts += dtt;
    bool zero = true;
    if (owned) {
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        const double v = qb[r] + ib[r];
        a.Qb_out[r * pitch + i] = v;
        zero = zero && v == 0.0;
      }
      th0 += ib[NQ];
      th1 += ib[NQ + 1];
    }
    const int z = __all(zero);
    if (lane == 0) a.zout[w] = z;
    // The next tile is in LDS once at most the 14 stores above are in
    // flight: vector memory operations retire in issue order, and a tile
    // below nw owns at least one cell, so all 14 are issued.
    __builtin_amdgcn_s_waitcnt(0x0F7E);  // vmcnt(14)
    w = wn;
  }
  const double bd = block_sum<ABLK>(dts);
  const double b0s = block_sum<ABLK>(th0);
  const double c0s = block_sum<ABLK>(th1);
  if (threadIdx.x == 0) {
    a.dtp[blockIdx.x] = bd;
    a.thb[2 * blockIdx.x] += b0s;
    a.thb[2 * blockIdx.x + 1] += c0s;
  }
}

// Zero flags of Qbar of the final state, per tile of TILE cells.
__global__ __launch_bounds__(64) void k_adj_flags(const double* __restrict__ Qb,
                                                  int* __restrict__ z, long long n,
                                                  long long pitch, int nw) {
  const long long w = blockIdx.x;
  const long long i = w * TILE + threadIdx.x;
  bool zero = true;
  if (threadIdx.x < TILE && i < n) {
#pragma unroll
    for (int r = 0; r < NQ; ++r) zero = zero && Qb[r * pitch + i] == 0.0;
  }
  const int all = __all(zero);
  if (threadIdx.x == 0 && w < nw) z[w] = all;
}

__global__ void k_adj_izero(int* __restrict__ p, long long m) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) p[t] = 0;
}

// Per-block, per-piece sums of Qbar of state 0 (pieces are contiguous cell
// ranges, so a block meets only a few of them).
__global__ __launch_bounds__(LBLK) void k_adj_q0sum(const double* __restrict__ Qb,
                                                    const long long* __restrict__ bounds, int nst,
                                                    double* __restrict__ part, long long n,
                                                    long long pitch) {
  const long long b0 = (long long)blockIdx.x * LBLK;
  const long long i = b0 + threadIdx.x;
  const long long last = (b0 + LBLK < n ? b0 + LBLK : n) - 1;
  int klo = 0;
  while (klo + 1 < nst && b0 >= bounds[klo + 1]) ++klo;
  int khi = klo;
  while (khi + 1 < nst && last >= bounds[khi + 1]) ++khi;
  int k = 0;
  while (k + 1 < nst && i >= bounds[k + 1]) ++k;
  double* out = part + (long long)blockIdx.x * nst * NQ;
  if (threadIdx.x == 0)
    for (int kk = 0; kk < nst; ++kk)
      if (kk < klo || kk > khi)
        for (int r = 0; r < NQ; ++r) out[kk * NQ + r] = 0.0;
  for (int kk = klo; kk <= khi; ++kk) {
    for (int r = 0; r < NQ; ++r) {
      double v = (i < n && k == kk) ? Qb[r * pitch + i] : 0.0;
      double s = block_sum<LBLK>(v);
      if (threadIdx.x == 0) out[kk * NQ + r] = s;
    }
  }
}

// Final reduction: Qbar per piece (with the lmax term of step 0), the VJP of
// prim2cons per piece, and the material gradient.  One wave.
__global__ __launch_bounds__(64) void k_adj_final(
    const double* __restrict__ part, int nbl, const double* __restrict__ thb, int nb,
    const double* __restrict__ dtp0, const unsigned long long* __restrict__ lmax,
    const long long* __restrict__ istar, const double* __restrict__ Jlam,
    const long long* __restrict__ bounds, int nst, const double* __restrict__ prims,
    const double* __restrict__ gin, double cfl, double dx, double* __restrict__ sums,
    double* __restrict__ gout) {
  const int lane = threadIdx.x;
  // lmax term of step 0
  double acc = 0.0;
  for (int b = lane; b < nb; b += 64) acc += dtp0[b];
  acc = wave_sum(acc);
  const double lm0 = lmax_of(lmax, 0);
  const double dtdx0 = cfl * rcp(lm0);
  const double lbar = -acc * dtdx0 * rcp(lm0);
  const long long i0 = istar[0];
  int k0 = 0;
  while (k0 + 1 < nst && i0 >= bounds[k0 + 1]) ++k0;
  for (int kr = 0; kr < nst * NQ; ++kr) {
    double v = 0.0;
    for (int b = lane; b < nbl; b += 64) v += part[(long long)b * nst * NQ + kr];
    v = wave_sum(v);
    const int k = kr / NQ, r = kr % NQ;
    if (k == k0) v += lbar * Jlam[r];
    if (lane == 0) sums[kr] = v;
  }
  double tb0 = 0.0, tc0 = 0.0;
  for (int b = lane; b < nb; b += 64) {
    tb0 += thb[2 * b];
    tc0 += thb[2 * b + 1];
  }
  tb0 = wave_sum(tb0);
  tc0 = wave_sum(tc0);
  __syncthreads();
  double pb0 = 0.0, pc0 = 0.0;
  for (int k = lane; k < nst; k += 64) {
    double in[hes_gen::P2C_IN], ib[hes_gen::P2C_IN];
#pragma unroll
    for (int r = 0; r < NP; ++r) in[r] = prims[NP * k + r];
#pragma unroll
    for (int r = 0; r < 5; ++r) in[NP + r] = gin[r];
#pragma unroll
    for (int r = 0; r < hes_gen::P2C_IN; ++r) ib[r] = 0.0;
    hes_gen::prim2cons_vjp(in, sums + NQ * k, ib);
#pragma unroll
    for (int r = 0; r < NP; ++r) gout[NP * k + r] = ib[r];
    pb0 += ib[NP];
    pc0 += ib[NP + 1];
  }
  pb0 = wave_sum(pb0);
  pc0 = wave_sum(pc0);
  if (lane == 0) {
    gout[NP * nst] = tb0 + pb0 + lbar * Jlam[NQ];
    gout[NP * nst + 1] = tc0 + pc0 + lbar * Jlam[NQ + 1];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct hes_adjoint {
  long long n, pitch, nsteps, every, nck;
  int nb;    // blocks of the tiled kernels
  int nbr;   // blocks of the reverse step: one per CU (persistent waves)
  int nbl;   // blocks of the one-cell-per-thread kernels
  int nw;    // tiles of TILE cells
  int sparse = 1;
  int use_graph = 1;  // replay captured graphs (else launch the kernels one by one)
  double cfl, dx;
  double mat8[8];
  MatDev mat;
  int nst;
  double* ck = nullptr;    // nck slots of 15 x pitch (Q, then C)
  double* seg[2] = {nullptr, nullptr};  // two segment buffers of every - 1 slots of 15 x pitch (Q, C)
  double* fq[2] = {nullptr, nullptr};
  double* fc[2] = {nullptr, nullptr};
  double* qb[2] = {nullptr, nullptr};
  int* zf = nullptr;                   // 2 x nw zero flags of qb[0], qb[1]
  unsigned long long* lmax = nullptr;  // nsteps + 1
  double* blk_val = nullptr;           // nsteps x nb
  long long* blk_idx = nullptr;
  long long* istar = nullptr;          // nsteps
  double* Jlam = nullptr;              // nsteps x 16
  double* dtp = nullptr;               // 2 x nb
  double* thb = nullptr;               // 2 x nb
  double* lossp = nullptr;             // nbl
  double* part = nullptr;              // nbl x nst x 14
  double* sums = nullptr;              // nst x 14
  double* gout = nullptr;              // 13 nst + 2
  double* gin = nullptr;               // b0, c0, rho0, cv, t0
  double* prims = nullptr;             // 13 x nst
  double* states = nullptr;            // 14 x nst
  long long* bounds = nullptr;         // nst + 1
  double* rho_t = nullptr;
  double* u_t = nullptr;
  double** slots = nullptr;            // device list of state slots (end-cell copies)
  int nslots = 0;
  int* status = nullptr;
  hipStream_t stream = nullptr;
  // the backward pass recomputes segment c - 1 on stream2 while the reverse
  // steps of segment c run on stream (see record_backward)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_ready[2] = {nullptr, nullptr},
             ev_done[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gf = nullptr, gb = nullptr;
  bool has_init = false, has_target = false, fwd_done = false;
};

namespace {

double* ck_q(hes_adjoint* h, long long c) { return h->ck + c * 15 * h->pitch; }
double* ck_c(hes_adjoint* h, long long c) { return ck_q(h, c) + 14 * h->pitch; }

// State slot of state s in the first forward pass.
void fwd_slot(hes_adjoint* h, long long s, double** q, double** c) {
  if (s % h->every == 0 && s < h->nsteps) {
    *q = ck_q(h, s / h->every);
    *c = ck_c(h, s / h->every);
  } else {
    *q = h->fq[s & 1];
    *c = h->fc[s & 1];
  }
}

// State slot of state s inside its backward segment (s0 = first state); the
// segment's recomputed states live in segment buffer `buf`.
double* seg_slot(hes_adjoint* h, long long s, long long s0, int buf) {
  if (s == s0) return ck_q(h, s0 / h->every);
  return h->seg[buf] + (s - s0 - 1) * 15 * h->pitch;
}

AdjFwdArgs fwd_args(hes_adjoint* h, long long s, const double* qi, double* qo, const double* ci,
                    double* co) {
  AdjFwdArgs a;
  a.Qin = qi;
  a.Qout = qo;
  a.Cin = ci;
  a.Cout = co;
  a.lmax = h->lmax;
  a.blk_val = h->blk_val;
  a.blk_idx = h->blk_idx;
  a.istar = h->istar;
  a.status = h->status;
  a.n = h->n;
  a.pitch = h->pitch;
  a.s = s;
  a.cfl = h->cfl;
  a.dx = h->dx;
  a.mat = h->mat;
  return a;
}

hipError_t record_forward(hes_adjoint* h) {
  hipStream_t st = h->stream;
  const long long n = h->n;
  const unsigned nbl = (unsigned)h->nbl;
  hipLaunchKernelGGL(k_adj_p2c, dim3((h->nst + 63) / 64), dim3(64), 0, st, h->prims, h->gin,
                     h->states, h->nst);
  hipLaunchKernelGGL(k_adj_fill, dim3(nbl), dim3(LBLK), 0, st, h->states, h->bounds, h->nst,
                     ck_q(h, 0), n, h->pitch);
  hipLaunchKernelGGL(k_adj_ends, dim3((h->nslots * 2 * NQ + 63) / 64), dim3(64), 0, st,
                     ck_q(h, 0), h->slots, h->nslots, n, h->pitch);
  hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)((h->nsteps + 1 + 255) / 256)), dim3(256), 0, st,
                     (double*)h->lmax, h->nsteps + 1);
  hipLaunchKernelGGL(k_adj_speeds, dim3(nbl), dim3(LBLK), 0, st, ck_q(h, 0), ck_c(h, 0), n,
                     h->pitch, h->mat, h->lmax, h->blk_val, h->blk_idx, h->status);
  for (long long s = 0; s < h->nsteps; ++s) {
    double *qi, *ci, *qo, *co;
    fwd_slot(h, s, &qi, &ci);
    fwd_slot(h, s + 1, &qo, &co);
    // per-block argmax records of the checkpoint states 1 .. nsteps-1 (the
    // final state needs none)
    if ((s + 1) % h->every == 0 && s + 1 < h->nsteps)
      hipLaunchKernelGGL(k_adj_fwd<FWD_RECORD>, dim3((unsigned)h->nb), dim3(ABLK), 0, st,
                         fwd_args(h, s, qi, qo, ci, co));
    else
      hipLaunchKernelGGL(k_adj_fwd<FWD_MAX>, dim3((unsigned)h->nb), dim3(ABLK), 0, st,
                         fwd_args(h, s, qi, qo, ci, co));
  }
  hipLaunchKernelGGL(k_adj_argmax, dim3((unsigned)h->nsteps), dim3(64), 0, st, h->blk_val,
                     h->blk_idx, h->nb, h->nbl, h->every, h->istar);
  double *qn, *cn;
  fwd_slot(h, h->nsteps, &qn, &cn);
  hipLaunchKernelGGL(k_adj_loss, dim3(nbl), dim3(LBLK), 0, st, qn, h->rho_t, h->u_t, h->mat8[0],
                     h->qb[h->nsteps & 1], h->lossp, n, h->pitch);
  return hipGetLastError();
}

// Returns a failed HIP call's status from a recording function.
#define HIP_TRY_E(expr)            \
  do {                             \
    hipError_t e_ = (expr);        \
    if (e_ != hipSuccess) return e_; \
  } while (0)

hipError_t record_backward(hes_adjoint* h) {
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)((2 * h->nb + 255) / 256)), dim3(256), 0, st,
                     h->thb, (long long)2 * h->nb);
  // zero flags: those of Qbar_n from its values, the other buffer's unknown
  const int fin = (int)(h->nsteps & 1);
  hipLaunchKernelGGL(k_adj_flags, dim3((unsigned)h->nw), dim3(64), 0, st, h->qb[fin],
                     h->zf + (long long)fin * h->nw, h->n, h->pitch, h->nw);
  hipLaunchKernelGGL(k_adj_izero, dim3((unsigned)((h->nw + 255) / 256)), dim3(256), 0, st,
                     h->zf + (long long)(1 - fin) * h->nw, (long long)h->nw);
  // Segment c's states are recomputed into segment buffer c & 1 on stream2
  // (the k_adj_fwd steps and k_adj_lam, which read only the checkpoints and
  // the first pass's CFL records), while the reverse steps of segment c + 1
  // run on `st` from the other buffer.  Events order the two: the reverse of
  // segment c waits for its buffer to be ready, and the recompute of segment
  // c - 1 waits until the reverse of segment c + 1 (the last reader of the
  // buffer it overwrites) is done.  Both streams are in the capture: stream2
  // forks from `st` and joins it at the end.
  hipStream_t st2 = h->stream2;
  HIP_TRY_E(hipEventRecord(h->ev_fork, st));
  HIP_TRY_E(hipStreamWaitEvent(st2, h->ev_fork, 0));
  auto recompute = [&](long long c) -> hipError_t {
    const long long s0 = c * h->every;
    const long long end = std::min(s0 + h->every, h->nsteps);
    const int buf = (int)(c & 1);
    for (long long s = s0; s + 1 < end; ++s) {
      // each recomputed state keeps its sound speeds in row 14 of its slot,
      // as the checkpoints do: the reverse step reads its face speeds there
      const double* ci = seg_slot(h, s, s0, buf) + 14 * h->pitch;
      double* co = seg_slot(h, s + 1, s0, buf) + 14 * h->pitch;
      hipLaunchKernelGGL(k_adj_fwd<FWD_FIND>, dim3((unsigned)h->nb), dim3(ABLK), 0, st2,
                         fwd_args(h, s, seg_slot(h, s, s0, buf), seg_slot(h, s + 1, s0, buf), ci,
                                  co));
    }
    const int cnt = (int)(end - s0);
    hipLaunchKernelGGL(k_adj_lam, dim3((cnt + 63) / 64), dim3(64), 0, st2, ck_q(h, c),
                       h->seg[buf], 15 * h->pitch, h->pitch, h->istar, h->gin, s0, cnt, h->n,
                       h->Jlam);
    return hipEventRecord(h->ev_ready[buf], st2);
  };
  HIP_TRY_E(recompute(h->nck - 1));
  for (long long c = h->nck - 1; c >= 0; --c) {
    const long long s0 = c * h->every;
    const long long end = std::min(s0 + h->every, h->nsteps);
    const int buf = (int)(c & 1);
    if (c >= 1) {
      if (c + 1 < h->nck) HIP_TRY_E(hipStreamWaitEvent(st2, h->ev_done[(c + 1) & 1], 0));
      HIP_TRY_E(recompute(c - 1));
    }
    HIP_TRY_E(hipStreamWaitEvent(st, h->ev_ready[buf], 0));
    for (long long s = end - 1; s >= s0; --s) {
      AdjStepArgs a;
      a.Q = seg_slot(h, s, s0, buf);
      a.Qb_in = h->qb[(s + 1) & 1];
      a.Qb_out = h->qb[s & 1];
      a.zin = h->zf + ((s + 1) & 1) * h->nw;
      a.zout = h->zf + (s & 1) * h->nw;
      a.lmax = h->lmax;
      a.istar = h->istar;
      a.Jlam = h->Jlam;
      a.dtp_next = h->dtp + ((s + 1) & 1) * h->nb;
      a.dtp = h->dtp + (s & 1) * h->nb;
      a.thb = h->thb;
      a.n = h->n;
      a.pitch = h->pitch;
      a.s = s;
      a.cfl = h->cfl;
      a.dx = h->dx;
      a.nb = h->nbr;
      a.nw = h->nw;
      a.sparse = h->sparse;
      a.has_next = s + 1 < h->nsteps ? 1 : 0;
      a.gin[0] = h->mat8[4];
      a.gin[1] = h->mat8[1];
      a.gin[2] = h->mat8[0];
      a.gin[3] = h->mat8[2];
      a.gin[4] = h->mat8[3];
      hipLaunchKernelGGL(k_adj_step, dim3((unsigned)h->nbr), dim3(ABLK), 0, st, a);
    }
    HIP_TRY_E(hipEventRecord(h->ev_done[buf], st));
  }
  HIP_TRY_E(hipEventRecord(h->ev_join, st2));
  HIP_TRY_E(hipStreamWaitEvent(st, h->ev_join, 0));
  hipLaunchKernelGGL(k_adj_q0sum, dim3((unsigned)h->nbl), dim3(LBLK), 0, st, h->qb[0], h->bounds,
                     h->nst, h->part, h->n, h->pitch);
  hipLaunchKernelGGL(k_adj_final, dim3(1), dim3(64), 0, st, h->part, h->nbl, h->thb, h->nbr,
                     h->dtp, h->lmax, h->istar, h->Jlam, h->bounds, h->nst, h->prims, h->gin,
                     h->cfl, h->dx, h->sums, h->gout);
  return hipGetLastError();
}

hipError_t capture(hes_adjoint* h, bool fwd, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t er = fwd ? record_forward(h) : record_backward(h);
  e = hipStreamEndCapture(h->stream, &g);
  if (er != hipSuccess) e = er;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  return e;
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, sizeof(T) * (count > 0 ? count : 1));
}

}  // namespace

extern "C" {

int hes_adjoint_create(hes_adjoint** out, long long ncells, const double* material, double cfl,
                       double dx, long long nsteps, long long every) {
  if (!out) return fail(HES_ERR_ARG, "null handle pointer");
  *out = nullptr;
  if (ncells < 3) return fail(HES_ERR_ARG, "need at least 3 cells");
  if (nsteps < 1) return fail(HES_ERR_ARG, "need at least one step");
  if (every < 1) return fail(HES_ERR_ARG, "checkpoint interval must be positive");
  if (!material || !mat_ok(material)) return fail(HES_ERR_ARG, "invalid material constants");
  if (!(material[5] == 1.0 && material[6] == 3.0 && material[7] == 2.0))
    return fail(HES_ERR_ARG, "the adjoint supports the exponents alpha=1, beta=3, gamma=2");
  if (!(cfl > 0) || !(dx > 0)) return fail(HES_ERR_ARG, "cfl and dx must be positive");
  if (nsteps > 0x7fffffffLL) return fail(HES_ERR_ARG, "too many steps");
  if (ncells >= (1LL << 28)) return fail(HES_ERR_ARG, "at most 2^28 - 1 cells");
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    return fail(HES_ERR_HIP, "cannot read the device's compute-unit count");
  hes_adjoint* h = new hes_adjoint();
  h->n = ncells;
  h->pitch = (ncells + 63) / 64 * 64 + 256;  // rows padded off a power-of-two stride, as the solver's
  h->nsteps = nsteps;
  h->every = every < nsteps ? every : nsteps;
  h->nck = (nsteps + h->every - 1) / h->every;
  const long long nwaves = (ncells + TILE - 1) / TILE;
  h->nw = (int)nwaves;
  h->nb = (int)((nwaves + ABLK / 64 - 1) / (ABLK / 64));
  h->nbr = std::min(h->nb, cus * (256 / ABLK));
  h->nbl = (int)((ncells + LBLK - 1) / LBLK);
  h->cfl = cfl;
  h->dx = dx;
  std::memcpy(h->mat8, material, sizeof(h->mat8));
  h->mat = make_mat(material);
  h->nst = 0;
  const size_t P = (size_t)h->pitch;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  hipEvent_t* sync_events[] = {&h->ev_fork, &h->ev_join, &h->ev_ready[0], &h->ev_ready[1],
                               &h->ev_done[0], &h->ev_done[1]};
  for (hipEvent_t* ev : sync_events)
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (e == hipSuccess) e = dalloc(&h->ck, (size_t)h->nck * 15 * P);
  for (int b = 0; b < 2 && e == hipSuccess && h->every > 1; ++b)
    e = dalloc(&h->seg[b], (size_t)(h->every - 1) * 15 * P);
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    e = dalloc(&h->fq[b], 14 * P);
    if (e == hipSuccess) e = dalloc(&h->fc[b], P);
    if (e == hipSuccess) e = dalloc(&h->qb[b], 14 * P);
  }
  if (e == hipSuccess) e = dalloc(&h->zf, (size_t)2 * h->nw);
  if (e == hipSuccess) e = dalloc(&h->lmax, (size_t)nsteps + 1);
  if (e == hipSuccess) e = dalloc(&h->blk_val, (size_t)nsteps * h->nb);
  if (e == hipSuccess) e = dalloc(&h->blk_idx, (size_t)nsteps * h->nb);
  if (e == hipSuccess) e = dalloc(&h->istar, (size_t)nsteps);
  if (e == hipSuccess) e = dalloc(&h->Jlam, (size_t)nsteps * JL);
  if (e == hipSuccess) e = dalloc(&h->dtp, (size_t)2 * h->nb);
  if (e == hipSuccess) e = dalloc(&h->thb, (size_t)2 * h->nb);
  if (e == hipSuccess) e = dalloc(&h->lossp, (size_t)h->nbl);
  if (e == hipSuccess) e = dalloc(&h->gin, 5);
  if (e == hipSuccess) e = dalloc(&h->rho_t, P);
  if (e == hipSuccess) e = dalloc(&h->u_t, P);
  if (e == hipSuccess) e = dalloc(&h->status, 4);
  if (e == hipSuccess) e = zero_on(h->stream, h->status, 4 * sizeof(int));
  if (e == hipSuccess) {
    // gin = (b0, c0, rho0, cv, t0): the inputs of the generated routines
    const double g[5] = {material[4], material[1], material[0], material[2], material[3]};
    e = copy_on(h->stream, h->gin, g, sizeof(g), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) {
    // every state slot whose end cells must be filled
    std::vector<double*> sl;
    for (long long c = 1; c < h->nck; ++c) sl.push_back(ck_q(h, c));
    for (int b = 0; b < 2; ++b)
      for (long long j = 0; j + 1 < h->every; ++j) sl.push_back(h->seg[b] + j * 15 * P);
    sl.push_back(h->fq[0]);
    sl.push_back(h->fq[1]);
    h->nslots = (int)sl.size();
    e = dalloc(&h->slots, sl.size());
    if (e == hipSuccess)
      e = copy_on(h->stream, h->slots, sl.data(), sizeof(double*) * sl.size(),
                  hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    g_err = std::string("adjoint allocation failed: ") + hipGetErrorString(e);
    hes_adjoint_destroy(h);
    return HES_ERR_HIP;
  }
  *out = h;
  return HES_OK;
}

int hes_adjoint_destroy(hes_adjoint* h) {
  if (!h) return HES_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  if (h->gf) (void)hipGraphExecDestroy(h->gf);
  if (h->gb) (void)hipGraphExecDestroy(h->gb);
  void* bufs[] = {h->ck, h->seg[0], h->seg[1], h->fq[0], h->fq[1], h->fc[0], h->fc[1], h->qb[0], h->qb[1], h->zf,
                  h->lmax, h->blk_val, h->blk_idx, h->istar, h->Jlam, h->dtp, h->thb, h->lossp,
                  h->part, h->sums, h->gout, h->gin, h->prims, h->states, h->bounds, h->rho_t,
                  h->u_t, h->slots, h->status};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  hipEvent_t evs[] = {h->ev0, h->ev1, h->ev_fork, h->ev_join, h->ev_ready[0], h->ev_ready[1],
                      h->ev_done[0], h->ev_done[1]};
  for (hipEvent_t ev : evs)
    if (ev) (void)hipEventDestroy(ev);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return HES_OK;
}

int hes_adjoint_set_initial(hes_adjoint* h, int nstates, const double* prims,
                            const double* interfaces) {
  if (!h || !prims || nstates < 1 || (nstates > 1 && !interfaces))
    return fail(HES_ERR_ARG, "null argument");
  for (int j = 0; j + 1 < nstates - 1; ++j)
    if (!(interfaces[j] <= interfaces[j + 1])) return fail(HES_ERR_ARG, "interfaces not sorted");
  for (int k = 0; k < NP * nstates; ++k)
    if (!std::isfinite(prims[k])) return fail(HES_ERR_ARG, "non-finite primitive state");
  // cell i (centre x = (i + 1/2) / n) takes piece #{j : x > s_j}, as the solver's fill
  std::vector<long long> bounds(nstates + 1);
  bounds[0] = 0;
  bounds[nstates] = h->n;
  for (int k = 1; k < nstates; ++k) {
    long long lo = 0, hi = h->n;  // first i with x_i > s_{k-1}
    while (lo < hi) {
      long long mid = (lo + hi) / 2;
      double x = ((double)mid + 0.5) / (double)h->n;
      if (x > interfaces[k - 1]) hi = mid;
      else lo = mid + 1;
    }
    bounds[k] = lo;
  }
  if (nstates != h->nst) {
    // the old piece buffers go: until the new ones are in place the handle
    // has no initial state (a failed allocation leaves it so)
    h->has_init = false;
    h->fwd_done = false;
    (void)hipStreamSynchronize(h->stream);
    if (h->gf) { (void)hipGraphExecDestroy(h->gf); h->gf = nullptr; }
    if (h->gb) { (void)hipGraphExecDestroy(h->gb); h->gb = nullptr; }
    void* old[] = {h->prims, h->states, h->bounds, h->part, h->sums, h->gout};
    for (void* p : old)
      if (p) (void)hipFree(p);
    h->prims = h->states = h->part = h->sums = h->gout = nullptr;
    h->bounds = nullptr;
    h->nst = 0;
    hipError_t e = dalloc(&h->prims, (size_t)NP * nstates);
    if (e == hipSuccess) e = dalloc(&h->states, (size_t)NQ * nstates);
    if (e == hipSuccess) e = dalloc(&h->bounds, (size_t)nstates + 1);
    if (e == hipSuccess) e = dalloc(&h->part, (size_t)h->nbl * nstates * NQ);
    if (e == hipSuccess) e = dalloc(&h->sums, (size_t)nstates * NQ);
    if (e == hipSuccess) e = dalloc(&h->gout, (size_t)NP * nstates + 2);
    if (e != hipSuccess) {
      g_err = std::string("adjoint allocation failed: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
    h->nst = nstates;
  }
  HIP_TRY(copy_on(h->stream, h->prims, prims, sizeof(double) * NP * nstates,
                  hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->bounds, bounds.data(), sizeof(long long) * (nstates + 1),
                  hipMemcpyHostToDevice));
  h->has_init = true;
  h->fwd_done = false;
  return HES_OK;
}

int hes_adjoint_set_target(hes_adjoint* h, const double* rho, const double* u1) {
  if (!h || !rho || !u1) return fail(HES_ERR_ARG, "null argument");
  HIP_TRY(copy_on(h->stream, h->rho_t, rho, sizeof(double) * h->n, hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->u_t, u1, sizeof(double) * h->n, hipMemcpyHostToDevice));
  h->has_target = true;
  h->fwd_done = false;
  return HES_OK;
}

int hes_adjoint_forward(hes_adjoint* h, double* loss, float* elapsed_ms) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->has_init || !h->has_target) return fail(HES_ERR_ARG, "initial state and target must be set");
  HIP_TRY(zero_on(h->stream, h->status, 4 * sizeof(int)));
  if (h->use_graph && !h->gf) {
    hipError_t e = capture(h, true, &h->gf);
    if (e != hipSuccess) {
      g_err = std::string("forward graph: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (h->use_graph) HIP_TRY(hipGraphLaunch(h->gf, h->stream));
  else HIP_TRY(record_forward(h));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  int st[2];
  HIP_TRY(copy_on(h->stream, st, h->status, sizeof(st), hipMemcpyDeviceToHost));
  if (st[0]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  std::vector<double> lp(h->nbl);
  HIP_TRY(copy_on(h->stream, lp.data(), h->lossp, sizeof(double) * h->nbl, hipMemcpyDeviceToHost));
  double l = 0.0;
  for (double v : lp) l += v;
  if (loss) *loss = l;
  h->fwd_done = true;
  return HES_OK;
}

int hes_adjoint_backward(hes_adjoint* h, double* grad_states, double* grad_material,
                         float* elapsed_ms) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  if (h->use_graph && !h->gb) {
    hipError_t e = capture(h, false, &h->gb);
    if (e != hipSuccess) {
      g_err = std::string("backward graph: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (h->use_graph) HIP_TRY(hipGraphLaunch(h->gb, h->stream));
  else HIP_TRY(record_backward(h));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  std::vector<double> g((size_t)NP * h->nst + 2);
  HIP_TRY(copy_on(h->stream, g.data(), h->gout, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
  if (grad_states) std::memcpy(grad_states, g.data(), sizeof(double) * NP * h->nst);
  if (grad_material) {
    grad_material[0] = g[NP * h->nst];
    grad_material[1] = g[NP * h->nst + 1];
  }
  for (double v : g)
    if (!std::isfinite(v)) return fail(HES_ERR_DOMAIN, "non-finite gradient");
  return HES_OK;
}

int hes_adjoint_set_sparse(hes_adjoint* h, int on) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if ((on != 0) != (h->sparse != 0)) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->gb) {
      (void)hipGraphExecDestroy(h->gb);
      h->gb = nullptr;
    }
    h->sparse = on != 0;
  }
  return HES_OK;
}

int hes_adjoint_set_launch(hes_adjoint* h, int use_graph) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  h->use_graph = use_graph != 0;
  return HES_OK;
}

int hes_adjoint_get_state(hes_adjoint* h, double* Q) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  double *qn, *cn;
  fwd_slot(h, h->nsteps, &qn, &cn);
  std::vector<double> buf((size_t)14 * h->pitch);
  HIP_TRY(copy_on(h->stream, buf.data(), qn, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
  for (long long i = 0; i < h->n; ++i)
    for (int r = 0; r < 14; ++r) Q[i * 14 + r] = buf[r * h->pitch + i];
  return HES_OK;
}

int hes_adjoint_get_profiles(hes_adjoint* h, double* rho, double* u1) {
  if (!h || !rho || !u1) return fail(HES_ERR_ARG, "null argument");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  double *qn, *cn;
  fwd_slot(h, h->nsteps, &qn, &cn);
  double *dr = nullptr, *du = nullptr;
  HIP_TRY(dalloc(&dr, (size_t)h->n));
  hipError_t e = dalloc(&du, (size_t)h->n);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_adj_profiles, dim3((unsigned)h->nbl), dim3(LBLK), 0, h->stream, qn,
                       h->mat8[0], dr, du, h->n, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = copy_on(h->stream, rho, dr, sizeof(double) * h->n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = copy_on(h->stream, u1, du, sizeof(double) * h->n, hipMemcpyDeviceToHost);
  (void)hipFree(dr);
  if (du) (void)hipFree(du);
  if (e != hipSuccess) {
    g_err = std::string("get_profiles: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return HES_OK;
}

int hes_adjoint_info(hes_adjoint* h, long long* device_bytes, long long* checkpoints) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  const long long P = h->pitch;
  long long b = h->nck * 15 * P + 2 * (h->every - 1) * 15 * P + 2 * (14 + 1 + 14) * P;
  b = b * 8 + (h->nsteps + 1) * 8 + h->nsteps * (long long)h->nb * 16 + h->nsteps * (8 + JL * 8);
  if (device_bytes) *device_bytes = b;
  if (checkpoints) *checkpoints = h->nck;
  return HES_OK;
}

}  // extern "C"
