// hes_badjoint.hip -- reverse mode of a batch of independent fixed-length
// single-phase Rusanov runs: for every problem p of the batch, the gradient of
// its own final-state misfit
//     L_p = sum_{i in p} (rho_i - rho*_i)^2 + (u1_i - u1*_i)^2
// with respect to the primitive initial states of its constant pieces and its
// material constants b0 and c0.
//
// Each problem runs the forward model of hes_adjoint.hip (main.jl:204-225 with
// NumFluxes.jl:25-60) with its own material, dx and CFL step dt_p = cfl dx_p /
// max_{i in p} lam_i; all problems take the same number of steps.  The cells
// of all problems are packed into one structure-of-arrays state (problem p
// owns cells [off_p, off_p+1)); one wave works on one tile of TILE cells of
// one problem, as in the batched solver, so a launch covers the whole batch.
//
// Per problem and step the forward pass records the CFL maximum (an atomic
// max per wave) and, per tile, the argmax (value, smallest index) of the new
// state's speeds.  The reverse step is the one-problem reverse step
// (reverse_chunk) on each tile.  Every reduction of a problem's dtdxbar and
// (b0bar, c0bar) runs over its tiles in tile order, so results do not depend
// on scheduling.  Storage, recompute and the overlap of the recompute with the
// reverse steps follow hes_adjoint.hip: every `every`-th state is kept, and a
// backward segment's states are recomputed on a second stream into one of two
// segment buffers while the reverse steps of the previous segment run.
// the secular-function eigenvalue iteration throughout (see sym3_extreme);
// set before any include
#define HES_EIG_SECULAR 1
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hes_host.hpp"
#include "hes_adjoint_common.hpp"
#include "../../include/hes.h"

using namespace hes;
using namespace hes_adj;

namespace {

constexpr int LBLK = 256;  // threads per block of the one-item-per-thread kernels

struct BTile {  // one wave's cells: [first, first + TILE) of problem p = [lo, hi)
  long long first, lo, hi;
  int p, pad;
};

// The tile of this block (one wave per block); tiles of one problem are
// contiguous, and the XCD-aware order gives each XCD a contiguous range.
__device__ __forceinline__ int btile() { return __builtin_amdgcn_readfirstlane((int)xcd_block()); }

// ---------------------------------------------------------------------------
// Forward kernels
// ---------------------------------------------------------------------------

// Conserved states of the constant pieces from their primitives (piece k
// belongs to problem piece_prob[k]; gin holds (b0, c0, rho0, cv, t0) per problem).
__global__ void k_badj_p2c(const double* __restrict__ prims, const double* __restrict__ gin,
                           const int* __restrict__ piece_prob, double* __restrict__ states,
                           int npieces) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npieces) return;
  const double* g = gin + 5 * piece_prob[k];
  double in[hes_gen::P2C_IN];
#pragma unroll
  for (int r = 0; r < NP; ++r) in[r] = prims[(long long)NP * k + r];
#pragma unroll
  for (int r = 0; r < 5; ++r) in[NP + r] = g[r];
  double out[NQ];
  hes_gen::prim2cons_fwd(in, out);
#pragma unroll
  for (int r = 0; r < NQ; ++r) states[(long long)NQ * k + r] = out[r];
}

// Index of the piece holding packed cell i: the last k with bounds[k] <= i.
__device__ __forceinline__ int piece_of(const long long* bounds, int npieces, long long i) {
  int lo = 0, hi = npieces - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bounds[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Piecewise-constant fill of the packed state: cells [bounds[k], bounds[k+1]) take state k.
__global__ void k_badj_fill(const double* __restrict__ states, const long long* __restrict__ bounds,
                            int npieces, double* __restrict__ Q, long long n, long long pitch) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = piece_of(bounds, npieces, i);
#pragma unroll
  for (int r = 0; r < NQ; ++r) Q[r * pitch + i] = states[(long long)NQ * k + r];
}

// The end cells of every problem never change: copy them from src into every state slot.
__global__ void k_badj_ends(const double* __restrict__ src, double* const* __restrict__ dst,
                            int ndst, const long long* __restrict__ offs, int P, long long pitch) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per_slot = (long long)P * 2 * NQ;
  if (t >= ndst * per_slot) return;
  const int slot = (int)(t / per_slot);
  const int rem = (int)(t % per_slot);
  const int p = rem / (2 * NQ), k = rem % (2 * NQ);
  const int r = k >> 1;
  const long long i = (k & 1) ? offs[p + 1] - 1 : offs[p];
  dst[slot][r * pitch + i] = src[r * pitch + i];
}

__global__ void k_badj_zero(double* __restrict__ p, long long m) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) p[t] = 0.0;
}

__global__ void k_badj_izero(int* __restrict__ p, long long m) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) p[t] = 0;
}

// Per-wave CFL record of a problem's new state: its maximum into lmax_p (an
// atomic max of the double bits; the speeds are positive) and, if bv is
// given, the tile's argmax (value, smallest index).
__device__ __forceinline__ void record_tile(double lam, bool owned, long long i,
                                            unsigned long long* lmax_p, double* bv, long long* bi,
                                            int tile) {
  double v = owned ? lam : -1.0;
  long long idx = owned ? i : NO_CELL;
  wave_argmax(v, idx);
  if ((threadIdx.x & 63) == 0) {
    if (v > 0.0) atomicMax(lmax_p, (unsigned long long)__double_as_longlong(v));
    if (bv) {
      bv[tile] = v;
      bi[tile] = idx;
    }
  }
}

// Sound speeds of state 0 and the CFL record of each problem (main.jl:204-212).
__global__ __launch_bounds__(64) void k_badj_speeds(const BTile* __restrict__ tiles,
                                                    const double* __restrict__ Q,
                                                    double* __restrict__ C, long long pitch,
                                                    const MatDev* __restrict__ mats,
                                                    unsigned long long* lmax0, double* bv,
                                                    long long* bi, int* status) {
  const int tile = btile();
  const BTile td = tiles[tile];
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane - 1;
  const bool owned = lane >= 1 && lane <= TILE && i < td.hi;
  double lam = 0.0;
  int bad = 0;
  if (owned) {
    double q[NQ], cc[1];
#pragma unroll
    for (int r = 0; r < NQ; ++r) q[r] = Q[r * pitch + i];
    lam = cell_speeds<1, 1>(mats + td.p, q, cc, bad);
    C[i] = cc[0];
  }
  if (__any(bad) && lane == 0) atomicOr(status, 1);
  record_tile(lam, owned, i, lmax0 + td.p, bv, bi, tile);
}

struct BAdjFwdArgs {
  const BTile* __restrict__ tiles;
  const MatDev* __restrict__ mats;
  const double* __restrict__ dx;  // per problem
  const double* __restrict__ Qin;
  double* __restrict__ Qout;
  const double* __restrict__ Cin;
  double* __restrict__ Cout;
  unsigned long long* lmax;  // lmax[s * P + p] = max CFL speed of problem p in state s
  double* blk_val;           // optional: per-tile argmax of the CFL speeds of state s + 1
  long long* blk_idx;
  int* status;
  long long pitch, s;
  double cfl;
  int P, ntiles;
  int write_lmax;            // max-reduce the new state's speeds into lmax[(s + 1) P + p]
};

// One time step of every problem: the solver's chunk step (so states match a
// plain batched run bit for bit) plus the bookkeeping of the reverse pass.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_badj_fwd(
    BAdjFwdArgs a) {
  const int tile = btile();
  const BTile td = a.tiles[tile];
  const int p = td.p;
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane - 1;
  const ChunkCtx ctx = {a.mats + p, a.Qin, a.Qout, a.Cin,  a.Cout, a.status,
                        a.pitch,    td.lo,  td.hi,  a.cfl, a.dx[p]};
  double q[NQ], cc[1];
  load_chunk<1>(ctx, i, q, cc);
  const double lmax = lmax_of(a.lmax, a.s * a.P + p);
  const double lam = chunk_step<1, FLUX_RUSANOV, 1>(ctx, i, q, cc, lmax);
  if (a.write_lmax) {
    const bool owned = lane >= 1 && lane <= TILE && i < td.hi;
    const long long row = (a.s + 1) * (long long)a.ntiles;
    record_tile(lam, owned, i, a.lmax + (a.s + 1) * a.P + p, a.blk_val ? a.blk_val + row : nullptr,
                a.blk_idx ? a.blk_idx + row : nullptr, tile);
  }
}

// Argmax cell of the CFL speed of every (state, problem) from the per-tile
// records (one wave per pair; ties to the smallest index).
__global__ __launch_bounds__(256) void k_badj_argmax(const double* __restrict__ blk_val,
                                                     const long long* __restrict__ blk_idx,
                                                     const int* __restrict__ tfirst, int ntiles,
                                                     int P, long long npairs,
                                                     long long* __restrict__ istar) {
  const long long pair = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pair >= npairs) return;  // the whole wave
  const long long s = pair / P;
  const int p = (int)(pair % P);
  const int lane = threadIdx.x & 63;
  double v = -2.0;
  long long i = NO_CELL;
  const double* rv = blk_val + s * ntiles;
  const long long* ri = blk_idx + s * ntiles;
  for (int b = tfirst[p] + lane; b < tfirst[p + 1]; b += 64) arg_better(v, i, rv[b], ri[b]);
  wave_argmax(v, i);
  if (lane == 0) istar[pair] = i;
}

// Misfit of each tile's cells in the final state and its gradient Qbar_n
// (rho and u1 as HyperelasticityMPh.jl:106-121 gives them).
__global__ __launch_bounds__(64) void k_badj_loss(const BTile* __restrict__ tiles,
                                                  const double* __restrict__ Q,
                                                  const double* __restrict__ rho_t,
                                                  const double* __restrict__ u_t,
                                                  const double* __restrict__ gin,
                                                  double* __restrict__ Qb,
                                                  double* __restrict__ lossp, long long pitch) {
  const int tile = btile();
  const BTile td = tiles[tile];
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane - 1;
  double l = 0.0;
  if (lane >= 1 && lane <= TILE && i < td.hi) {
    // rho and u1 by the routine that also makes target profiles, so a cell
    // where the runs agree has a misfit and a gradient of exactly zero
    double in[hes_gen::PROF_IN], o[2];
#pragma unroll
    for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
    in[NQ] = gin[5 * td.p + 2];
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
  l = wave_sum(l);
  if (lane == 0) lossp[tile] = l;
}

// rho and u1 of every cell with the misfit's own expressions.
__global__ __launch_bounds__(64) void k_badj_profiles(const BTile* __restrict__ tiles,
                                                      const double* __restrict__ Q,
                                                      const double* __restrict__ gin,
                                                      double* __restrict__ rho,
                                                      double* __restrict__ u1, long long pitch) {
  const BTile td = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane - 1;
  if (!(lane >= 1 && lane <= TILE && i < td.hi)) return;
  double in[hes_gen::PROF_IN], o[2];
#pragma unroll
  for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
  in[NQ] = gin[5 * td.p + 2];
  hes_gen::profiles_fwd(in, o);
  rho[i] = o[0];
  u1[i] = o[1];
}

// ---------------------------------------------------------------------------
// Reverse kernels
// ---------------------------------------------------------------------------

// dlam/d(Q, b0, c0) of the argmax cell of every problem in states s0 ..
// s0+cnt-1 (thread t: state s0 + t / P, problem t % P; the state lives in Q0
// for the first, else in seg + (j - 1) * stride).
__global__ __launch_bounds__(64) void k_badj_lam(const double* __restrict__ Q0,
                                                 const double* __restrict__ seg, long long stride,
                                                 long long pitch,
                                                 const long long* __restrict__ istar,
                                                 const double* __restrict__ gin, long long s0,
                                                 int cnt, int P, long long n,
                                                 double* __restrict__ Jlam) {
  const long long t = (long long)blockIdx.x * 64 + threadIdx.x;
  if (t >= (long long)cnt * P) return;
  const int j = (int)(t / P), p = (int)(t % P);
  const long long row = (s0 + j) * P + p;
  const long long i = istar[row];
  if (i < 0 || i >= n) {  // no finite CFL speed was recorded: no lmax term
#pragma unroll
    for (int r = 0; r < JL; ++r) Jlam[row * JL + r] = 0.0;
    return;
  }
  const double* Q = j == 0 ? Q0 : seg + (long long)(j - 1) * stride;
  double in[hes_gen::CELL_IN];
#pragma unroll
  for (int r = 0; r < NQ; ++r) in[r] = Q[r * pitch + i];
#pragma unroll
  for (int r = 0; r < 5; ++r) in[NQ + r] = gin[5 * p + r];
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
  for (int r = 0; r < JL; ++r) Jlam[row * JL + r] = ib[r];
}

struct BAdjStepArgs {
  const BTile* __restrict__ tiles;
  const int* __restrict__ tfirst;    // first tile of each problem (P + 1)
  const double* __restrict__ Q;      // state s
  const double* __restrict__ Qb_in;  // Qbar of state s+1 (without its lmax term)
  double* __restrict__ Qb_out;       // Qbar of state s (without its lmax term)
  const int* zin;                    // per tile: 1 if Qb_in is zero on all its owned cells
  int* zout;                         // the same for Qb_out
  const unsigned long long* lmax;
  const long long* istar;
  const double* Jlam;
  const double* dtp_next;            // per-tile partials of dtdxbar of step s+1
  double* dtp;                       // per-tile partials of dtdxbar of step s
  double* thb;                       // running (b0bar, c0bar) per tile
  const double* __restrict__ dx;     // per problem
  const double* __restrict__ gin;    // (b0, c0, rho0, cv, t0) per problem
  long long pitch, s;
  double cfl;
  int P, has_next;
  int sparse;                        // skip tiles whose cotangent is zero on them and their halo
  int ntiles;
};

constexpr int PROWS = 2 * NQ + 1;  // rows of a tile in LDS: Q (14), its sound speed, Qbar_{s+1} (14)
typedef __attribute__((address_space(3))) void lds_t;

// Copies the PROWS rows of tile td (cells first-1 .. first+62, clamped to its
// problem's [lo, hi)) from HBM into the wave's LDS buffer at byte address
// `lds`, without passing through registers: each instruction moves 4 bytes
// per lane, lane l taking half l&1 of cell 32h + l/2 of the tile, so a row
// takes two instructions and lands in doubles 64 r .. 64 r + 63 of the buffer
// in cell order.  Byte offsets within a row are 32-bit (fewer than 2^28
// cells in all is checked at creation).
__device__ __forceinline__ void btile_prefetch(const BAdjStepArgs& a, const BTile& td, int lane,
                                               unsigned lds) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    long long c = td.first - 1 + 32 * h + (lane >> 1);
    c = c < td.lo ? td.lo : (c >= td.hi ? td.hi - 1 : c);
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

// One reverse step of every problem.  A tile whose Qbar of state s+1 is zero
// on its owned cells and on the neighbouring tiles of its problem (and which
// holds no argmax cell of step s+1) has Qbar of state s = 0 and contributes
// nothing to dtdxbar or (b0bar, c0bar).
//
// Persistent waves, one per SIMD: wave k of the grid takes tiles k, k + W,
// k + 2W, ...  While it evaluates one tile, the next one it will evaluate
// streams into its LDS buffer with that tile's CFL maximum, argmax cell,
// running (b0bar, c0bar) and material inputs, so only the first tile's loads
// are waited for.  Every tile's arithmetic and partials are those of a wave
// that evaluates it alone.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1))) void k_badj_step(
    BAdjStepArgs a) {
  // the tile's rows, then its context: CFL maximum (dwords 0-1), argmax
  // cell of step s+1 (2-3), running (b0bar, c0bar) (4-7), material inputs
  // (8-17)
  __shared__ double buf[PROWS * 64 + 32];
  const unsigned lds = (unsigned)(size_t)(lds_t*)buf;
  const double* ctx = buf + PROWS * 64;
  const int lane = threadIdx.x & 63;
  const int W = (int)gridDim.x;
  const long long pitch = a.pitch;
  // The tile records, the CFL records, the argmax cells and the zero flags
  // of Qbar_{s+1} are read-only in this kernel: read through the constant
  // address space, they load into scalar registers, and no wait for them
  // waits for vector memory.
  typedef const __attribute__((address_space(4))) BTile ctile_t;
  typedef const __attribute__((address_space(4))) long long ci64_t;
  typedef const __attribute__((address_space(4))) int ci32_t;
  auto record = [&](int t) -> BTile {
    ctile_t* tr = (ctile_t*)a.tiles + t;
    BTile r;
    r.first = tr->first;
    r.lo = tr->lo;
    r.hi = tr->hi;
    r.p = tr->p;
    return r;
  };
  auto is1_of = [&](int p) -> long long {
    return a.has_next ? ((ci64_t*)a.istar)[(a.s + 1) * a.P + p] : -2;
  };
  // The first tile from t on (in steps of W) that must be evaluated; the
  // skipped ones on the way get their zero cotangent and partial.
  auto next_tile = [&](int t) -> int {
    for (; t < a.ntiles; t += W) {
      if (!a.sparse) return t;
      const BTile td = record(t);
      const long long is1 = is1_of(td.p);
      if (is1 >= td.first - 1 && is1 <= td.first + TILE) return t;
      ci32_t* zin = (ci32_t*)a.zin;
      const bool skip = zin[t] && (td.first == td.lo || zin[t - 1]) &&
                        (td.first + TILE >= td.hi || zin[t + 1]);
      if (!skip) return t;
      const long long i = td.first + lane - 1;
      if (!a.zout[t] && i >= td.lo && i < td.hi && lane >= 1 && lane <= TILE) {
#pragma unroll
        for (int r = 0; r < NQ; ++r) a.Qb_out[r * pitch + i] = 0.0;
      }
      if (lane == 0) {
        a.zout[t] = 1;
        a.dtp[t] = 0.0;
      }
    }
    return t;
  };
  // The next tile's record (one scalar round trip: its rows' addresses
  // depend on it), then its context and its rows into LDS.
  BTile td;
  auto fetch = [&](int t) {
    td = record(t);
    const char* src = (const char*)(a.lmax + a.s * a.P + td.p);
    if (lane >= 2 && lane < 4 && a.has_next)
      src = (const char*)(a.istar + (a.s + 1) * a.P + td.p) + 4 * (lane - 2);
    else if (lane >= 4 && lane < 8)
      src = (const char*)(a.thb + 2 * (long long)t) + 4 * (lane - 4);
    else if (lane >= 8 && lane < 18)
      src = (const char*)(a.gin + 5 * (long long)td.p) + 4 * (lane - 8);
    else if (lane < 2)
      src += 4 * lane;
    unsigned dc = lds + PROWS * 512u;
    asm volatile("" : "+s"(dc));  // the explicit waits below order the context's reads
    __builtin_amdgcn_global_load_lds((const void*)src, (lds_t*)(size_t)dc, 4, 0, 0);
    btile_prefetch(a, td, lane, lds);
  };
  int tile = __builtin_amdgcn_readfirstlane(next_tile(btile()));
  if (tile < a.ntiles) fetch(tile);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the first tile is in LDS
  while (tile < a.ntiles) {
    const int p = td.p;
    const long long lo = td.lo, hi = td.hi, first = td.first;
    // an opaque offset: the context's reads are ordered by the explicit
    // waits, and the compiler adds none of its own for the transfer into it
    unsigned co = 0;
    asm volatile("" : "+s"(co));
    const double lmc = ctx[co], tb0c = ctx[co + 2], tb1c = ctx[co + 3];
    const long long is1c = a.has_next ? __double_as_longlong(ctx[co + 1]) : -2;
    double ginc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) ginc[k] = ctx[co + 4 + k];
    const long long i = first + lane - 1;
    const bool inr = i >= lo && i < hi;
    const bool owned = inr && lane >= 1 && lane <= TILE;
    const bool hit = is1c >= first - 1 && is1c <= first + TILE;  // wave-uniform
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
    const int tn = __builtin_amdgcn_readfirstlane(next_tile(tile + W));
    if (tn < a.ntiles) fetch(tn);
    double dts = 0.0, th0 = 0.0, th1 = 0.0;
    // lmax term of step s+1 at its argmax cell
    if (hit) {
      double acc = 0.0;
      for (int b = a.tfirst[p] + lane; b < a.tfirst[p + 1]; b += 64) acc += a.dtp_next[b];
      acc = wave_sum(acc);
      const double lm1 = lmax_of(a.lmax, (a.s + 1) * a.P + p);
      const double dtdx1 = a.cfl * rcp(lm1);  // as chunk_step forms dt / dx
      const double lbar = -acc * dtdx1 * rcp(lm1);
      if (inr && i == is1c) {
        const double* J = a.Jlam + ((a.s + 1) * a.P + p) * JL;
#pragma unroll
        for (int r = 0; r < NQ; ++r) qb[r] += lbar * J[r];
        if (owned) {
          th0 += lbar * J[NQ];
          th1 += lbar * J[NQ + 1];
        }
      }
    }
    const bool interior = inr && i > lo && i < hi - 1;
    const double dtdx = a.cfl * rcp(lmc);  // as chunk_step forms it
    double ib[hes_gen::CELL_IN];
    reverse_chunk(q, qb, interior, owned, cs, 0.5 * dtdx, ginc, ib, dts);
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
    dts = wave_sum(dts);
    th0 = wave_sum(th0);
    th1 = wave_sum(th1);
    if (lane == 0) {
      a.zout[tile] = z;
      a.dtp[tile] = dts;
      a.thb[2 * tile] = tb0c + th0;
      a.thb[2 * tile + 1] = tb1c + th1;
    }
    // The next tile is in LDS once at most the 14 stores of Qbar above are
    // in flight: vector memory operations retire in issue order, and every
    // tile owns at least one cell, so all 14 are issued.
    __builtin_amdgcn_s_waitcnt(0x0F7E);  // vmcnt(14)
    tile = tn;
  }
}

// Zero flags of Qbar of the final state, per tile.
__global__ __launch_bounds__(64) void k_badj_flags(const BTile* __restrict__ tiles,
                                                   const double* __restrict__ Qb,
                                                   int* __restrict__ z, long long pitch) {
  const BTile td = tiles[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane - 1;
  bool zero = true;
  if (lane >= 1 && lane <= TILE && i < td.hi) {
#pragma unroll
    for (int r = 0; r < NQ; ++r) zero = zero && Qb[r * pitch + i] == 0.0;
  }
  const int all = __all(zero);
  if (lane == 0) z[blockIdx.x] = all;
}

// Sum of Qbar of state 0 over each piece's cells (one wave per piece, lanes
// striding the piece in a fixed order).
__global__ __launch_bounds__(64) void k_badj_piece_sum(const double* __restrict__ Qb,
                                                       const long long* __restrict__ bounds,
                                                       double* __restrict__ sums, long long pitch) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const long long b = bounds[k], e = bounds[k + 1];
  for (int r = 0; r < NQ; ++r) {
    double v = 0.0;
    for (long long i = b + lane; i < e; i += 64) v += Qb[r * pitch + i];
    v = wave_sum(v);
    if (lane == 0) sums[(long long)NQ * k + r] = v;
  }
}

// Per problem (one wave each): the lmax term of step 0, the VJP of prim2cons
// per piece, and the material gradient.
__global__ __launch_bounds__(64) void k_badj_final(
    const double* __restrict__ sums, const double* __restrict__ thb,
    const double* __restrict__ dtp0, const int* __restrict__ tfirst,
    const unsigned long long* __restrict__ lmax, const long long* __restrict__ istar,
    const double* __restrict__ Jlam, const long long* __restrict__ bounds,
    const int* __restrict__ poff, const double* __restrict__ prims,
    const double* __restrict__ gin, const double* __restrict__ dx, double cfl,
    double* __restrict__ gstates, double* __restrict__ gmat) {
  const int p = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const double* g = gin + 5 * p;
  double acc = 0.0, tb0 = 0.0, tc0 = 0.0;
  for (int b = tfirst[p] + lane; b < tfirst[p + 1]; b += 64) {
    acc += dtp0[b];
    tb0 += thb[2 * b];
    tc0 += thb[2 * b + 1];
  }
  acc = wave_sum(acc);
  tb0 = wave_sum(tb0);
  tc0 = wave_sum(tc0);
  const double lm0 = lmax_of(lmax, p);
  const double dtdx0 = cfl * rcp(lm0);
  const double lbar = -acc * dtdx0 * rcp(lm0);
  const double* J0 = Jlam + (long long)p * JL;  // state 0, problem p
  const long long i0 = istar[p];
  double pb0 = 0.0, pc0 = 0.0;
  for (int k = poff[p] + lane; k < poff[p + 1]; k += 64) {
    const bool has = i0 >= bounds[k] && i0 < bounds[k + 1];
    double sb[NQ];
#pragma unroll
    for (int r = 0; r < NQ; ++r) sb[r] = sums[(long long)NQ * k + r] + (has ? lbar * J0[r] : 0.0);
    double in[hes_gen::P2C_IN], ib[hes_gen::P2C_IN];
#pragma unroll
    for (int r = 0; r < NP; ++r) in[r] = prims[(long long)NP * k + r];
#pragma unroll
    for (int r = 0; r < 5; ++r) in[NP + r] = g[r];
#pragma unroll
    for (int r = 0; r < hes_gen::P2C_IN; ++r) ib[r] = 0.0;
    hes_gen::prim2cons_vjp(in, sb, ib);
#pragma unroll
    for (int r = 0; r < NP; ++r) gstates[(long long)NP * k + r] = ib[r];
    pb0 += ib[NP];
    pc0 += ib[NP + 1];
  }
  pb0 = wave_sum(pb0);
  pc0 = wave_sum(pc0);
  if (lane == 0) {
    gmat[2 * p] = tb0 + pb0 + lbar * J0[NQ];
    gmat[2 * p + 1] = tc0 + pc0 + lbar * J0[NQ + 1];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct hes_badjoint {
  int P = 0, ntiles = 0, npieces = 0;
  int nbr = 0;  // waves of the reverse step (persistent)
  long long N = 0, pitch = 0, nsteps = 0, every = 0, nck = 0;
  int sparse = 1;
  int use_graph = 1;
  double cfl = 0.6;
  std::vector<long long> offs_h;
  std::vector<int> tfirst_h, tprob_h, poff_h;
  std::vector<double> mats_h;  // 8 x P
  BTile* tiles = nullptr;
  int* tfirst = nullptr;
  long long* offs = nullptr;
  MatDev* mats = nullptr;
  double* dx = nullptr;
  double* gin = nullptr;               // 5 x P: b0, c0, rho0, cv, t0
  double* ck = nullptr;                // nck slots of 15 x pitch (Q, then C)
  double* seg[2] = {nullptr, nullptr};
  double* fq[2] = {nullptr, nullptr};
  double* fc[2] = {nullptr, nullptr};
  double* qb[2] = {nullptr, nullptr};
  int* zf = nullptr;                   // 2 x ntiles
  unsigned long long* lmax = nullptr;  // (nsteps + 1) x P
  double* blk_val = nullptr;           // nsteps x ntiles
  long long* blk_idx = nullptr;
  long long* istar = nullptr;          // nsteps x P
  double* Jlam = nullptr;              // nsteps x P x 16
  double* dtp = nullptr;               // 2 x ntiles
  double* thb = nullptr;               // 2 x ntiles
  double* lossp = nullptr;             // ntiles
  double* rho_t = nullptr;
  double* u_t = nullptr;
  double** slots = nullptr;
  int nslots = 0;
  int* status = nullptr;
  // pieces (set_initial)
  int* poff = nullptr;                 // P + 1
  int* piece_prob = nullptr;
  long long* bounds = nullptr;         // npieces + 1, packed cell coordinates
  double* prims = nullptr;             // 13 x npieces
  double* states = nullptr;            // 14 x npieces
  double* sums = nullptr;              // 14 x npieces
  double* gstates = nullptr;           // 13 x npieces
  double* gmat = nullptr;              // 2 x P
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_ready[2] = {nullptr, nullptr},
             ev_done[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t gf = nullptr, gb = nullptr;
  bool has_init = false, has_target = false, fwd_done = false;
};

namespace {

double* bck_q(hes_badjoint* h, long long c) { return h->ck + c * 15 * h->pitch; }
double* bck_c(hes_badjoint* h, long long c) { return bck_q(h, c) + 14 * h->pitch; }

void bfwd_slot(hes_badjoint* h, long long s, double** q, double** c) {
  if (s % h->every == 0 && s < h->nsteps) {
    *q = bck_q(h, s / h->every);
    *c = bck_c(h, s / h->every);
  } else {
    *q = h->fq[s & 1];
    *c = h->fc[s & 1];
  }
}

double* bseg_slot(hes_badjoint* h, long long s, long long s0, int buf) {
  if (s == s0) return bck_q(h, s0 / h->every);
  return h->seg[buf] + (s - s0 - 1) * 15 * h->pitch;
}

BAdjFwdArgs bfwd_args(hes_badjoint* h, long long s, const double* qi, double* qo, const double* ci,
                      double* co, bool first_pass) {
  BAdjFwdArgs a;
  a.tiles = h->tiles;
  a.mats = h->mats;
  a.dx = h->dx;
  a.Qin = qi;
  a.Qout = qo;
  a.Cin = ci;
  a.Cout = co;
  a.lmax = h->lmax;
  const bool rec = first_pass && s + 1 < h->nsteps;  // the final state needs no argmax
  a.blk_val = rec ? h->blk_val : nullptr;
  a.blk_idx = rec ? h->blk_idx : nullptr;
  a.status = h->status;
  a.pitch = h->pitch;
  a.s = s;
  a.cfl = h->cfl;
  a.P = h->P;
  a.ntiles = h->ntiles;
  a.write_lmax = first_pass ? 1 : 0;
  return a;
}

unsigned blocks_of(long long m, int per) { return (unsigned)((m + per - 1) / per); }

hipError_t brecord_forward(hes_badjoint* h) {
  hipStream_t st = h->stream;
  const unsigned nt = (unsigned)h->ntiles;
  hipLaunchKernelGGL(k_badj_p2c, dim3(blocks_of(h->npieces, 64)), dim3(64), 0, st, h->prims,
                     h->gin, h->piece_prob, h->states, h->npieces);
  hipLaunchKernelGGL(k_badj_fill, dim3(blocks_of(h->N, LBLK)), dim3(LBLK), 0, st, h->states,
                     h->bounds, h->npieces, bck_q(h, 0), h->N, h->pitch);
  hipLaunchKernelGGL(k_badj_ends, dim3(blocks_of((long long)h->nslots * h->P * 2 * NQ, LBLK)),
                     dim3(LBLK), 0, st, bck_q(h, 0), h->slots, h->nslots, h->offs, h->P, h->pitch);
  hipLaunchKernelGGL(k_badj_zero, dim3(blocks_of((h->nsteps + 1) * h->P, LBLK)), dim3(LBLK), 0, st,
                     (double*)h->lmax, (h->nsteps + 1) * h->P);
  hipLaunchKernelGGL(k_badj_speeds, dim3(nt), dim3(64), 0, st, h->tiles, bck_q(h, 0), bck_c(h, 0),
                     h->pitch, h->mats, h->lmax, h->blk_val, h->blk_idx, h->status);
  for (long long s = 0; s < h->nsteps; ++s) {
    double *qi, *ci, *qo, *co;
    bfwd_slot(h, s, &qi, &ci);
    bfwd_slot(h, s + 1, &qo, &co);
    hipLaunchKernelGGL(k_badj_fwd, dim3(nt), dim3(64), 0, st, bfwd_args(h, s, qi, qo, ci, co, true));
  }
  const long long npairs = h->nsteps * h->P;
  hipLaunchKernelGGL(k_badj_argmax, dim3(blocks_of(npairs, 4)), dim3(256), 0, st, h->blk_val,
                     h->blk_idx, h->tfirst, h->ntiles, h->P, npairs, h->istar);
  double *qn, *cn;
  bfwd_slot(h, h->nsteps, &qn, &cn);
  hipLaunchKernelGGL(k_badj_loss, dim3(nt), dim3(64), 0, st, h->tiles, qn, h->rho_t, h->u_t, h->gin,
                     h->qb[h->nsteps & 1], h->lossp, h->pitch);
  return hipGetLastError();
}

#define HIP_TRY_E(expr)              \
  do {                               \
    hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

hipError_t brecord_backward(hes_badjoint* h) {
  hipStream_t st = h->stream;
  const unsigned nt = (unsigned)h->ntiles;
  hipLaunchKernelGGL(k_badj_zero, dim3(blocks_of(2LL * h->ntiles, LBLK)), dim3(LBLK), 0, st, h->thb,
                     2LL * h->ntiles);
  const int fin = (int)(h->nsteps & 1);
  hipLaunchKernelGGL(k_badj_flags, dim3(nt), dim3(64), 0, st, h->tiles, h->qb[fin],
                     h->zf + (long long)fin * h->ntiles, h->pitch);
  hipLaunchKernelGGL(k_badj_izero, dim3(blocks_of(h->ntiles, LBLK)), dim3(LBLK), 0, st,
                     h->zf + (long long)(1 - fin) * h->ntiles, (long long)h->ntiles);
  // segment c is recomputed on stream2 into segment buffer c & 1 while the
  // reverse steps of segment c + 1 run on `st` (events order the two, as in
  // the one-problem adjoint)
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
      const double* ci = bseg_slot(h, s, s0, buf) + 14 * h->pitch;
      double* co = bseg_slot(h, s + 1, s0, buf) + 14 * h->pitch;
      hipLaunchKernelGGL(k_badj_fwd, dim3(nt), dim3(64), 0, st2,
                         bfwd_args(h, s, bseg_slot(h, s, s0, buf), bseg_slot(h, s + 1, s0, buf), ci,
                                   co, false));
    }
    const int cnt = (int)(end - s0);
    hipLaunchKernelGGL(k_badj_lam, dim3(blocks_of((long long)cnt * h->P, 64)), dim3(64), 0, st2,
                       bck_q(h, c), h->seg[buf], 15 * h->pitch, h->pitch, h->istar, h->gin, s0, cnt,
                       h->P, h->N, h->Jlam);
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
      BAdjStepArgs a;
      a.tiles = h->tiles;
      a.tfirst = h->tfirst;
      a.Q = bseg_slot(h, s, s0, buf);
      a.Qb_in = h->qb[(s + 1) & 1];
      a.Qb_out = h->qb[s & 1];
      a.zin = h->zf + ((s + 1) & 1) * h->ntiles;
      a.zout = h->zf + (s & 1) * h->ntiles;
      a.lmax = h->lmax;
      a.istar = h->istar;
      a.Jlam = h->Jlam;
      a.dtp_next = h->dtp + ((s + 1) & 1) * h->ntiles;
      a.dtp = h->dtp + (s & 1) * h->ntiles;
      a.thb = h->thb;
      a.dx = h->dx;
      a.gin = h->gin;
      a.pitch = h->pitch;
      a.s = s;
      a.cfl = h->cfl;
      a.P = h->P;
      a.has_next = s + 1 < h->nsteps ? 1 : 0;
      a.sparse = h->sparse;
      a.ntiles = h->ntiles;
      hipLaunchKernelGGL(k_badj_step, dim3(h->nbr), dim3(64), 0, st, a);
    }
    HIP_TRY_E(hipEventRecord(h->ev_done[buf], st));
  }
  HIP_TRY_E(hipEventRecord(h->ev_join, st2));
  HIP_TRY_E(hipStreamWaitEvent(st, h->ev_join, 0));
  hipLaunchKernelGGL(k_badj_piece_sum, dim3((unsigned)h->npieces), dim3(64), 0, st, h->qb[0],
                     h->bounds, h->sums, h->pitch);
  hipLaunchKernelGGL(k_badj_final, dim3((unsigned)h->P), dim3(64), 0, st, h->sums, h->thb, h->dtp,
                     h->tfirst, h->lmax, h->istar, h->Jlam, h->bounds, h->poff, h->prims, h->gin,
                     h->dx, h->cfl, h->gstates, h->gmat);
  return hipGetLastError();
}

hipError_t bcapture(hes_badjoint* h, bool fwd, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t er = fwd ? brecord_forward(h) : brecord_backward(h);
  e = hipStreamEndCapture(h->stream, &g);
  if (er != hipSuccess) e = er;
  if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  return e;
}

template <class T>
hipError_t balloc(T** p, size_t count) {
  return hipMalloc((void**)p, sizeof(T) * (count > 0 ? count : 1));
}

void drop_graphs(hes_badjoint* h) {
  if (h->gf) { (void)hipGraphExecDestroy(h->gf); h->gf = nullptr; }
  if (h->gb) { (void)hipGraphExecDestroy(h->gb); h->gb = nullptr; }
}

}  // namespace

extern "C" {

int hes_badjoint_create(hes_badjoint** out, int nproblems, const long long* ncells,
                        const double* materials, double cfl, const double* dx, long long nsteps,
                        long long every) {
  if (!out) return fail(HES_ERR_ARG, "null handle pointer");
  *out = nullptr;
  if (nproblems < 1 || !ncells || !materials) return fail(HES_ERR_ARG, "null argument");
  if (nsteps < 1) return fail(HES_ERR_ARG, "need at least one step");
  if (every < 1) return fail(HES_ERR_ARG, "checkpoint interval must be positive");
  if (!(cfl > 0)) return fail(HES_ERR_ARG, "cfl must be positive");
  if (nsteps > 0x7fffffffLL || nsteps * (long long)nproblems > 0x7fffffffLL)
    return fail(HES_ERR_ARG, "too many steps x problems");
  std::vector<double> vdx(nproblems);
  std::vector<MatDev> mats(nproblems);
  std::vector<double> gin(5 * (size_t)nproblems);
  for (int p = 0; p < nproblems; ++p) {
    if (ncells[p] < 3) return fail(HES_ERR_ARG, "every problem needs at least 3 cells");
    const double* m = materials + 8 * (size_t)p;
    if (!mat_ok(m)) return fail(HES_ERR_ARG, "invalid material constants");
    if (!(m[5] == 1.0 && m[6] == 3.0 && m[7] == 2.0))
      return fail(HES_ERR_ARG, "the adjoint supports the exponents alpha=1, beta=3, gamma=2");
    vdx[p] = dx ? dx[p] : 1.0 / (double)ncells[p];
    if (!(vdx[p] > 0)) return fail(HES_ERR_ARG, "dx must be positive");
    mats[p] = make_mat(m);
    const double g[5] = {m[4], m[1], m[0], m[2], m[3]};  // b0, c0, rho0, cv, t0
    std::memcpy(&gin[5 * (size_t)p], g, sizeof(g));
  }
  hes_badjoint* h = new hes_badjoint();
  h->P = nproblems;
  h->cfl = cfl;
  h->nsteps = nsteps;
  h->every = every < nsteps ? every : nsteps;
  h->nck = (nsteps + h->every - 1) / h->every;
  h->mats_h.assign(materials, materials + 8 * (size_t)nproblems);
  h->offs_h.resize(nproblems + 1);
  h->tfirst_h.resize(nproblems + 1);
  h->offs_h[0] = 0;
  std::vector<BTile> tdesc;
  for (int p = 0; p < nproblems; ++p) {
    h->offs_h[p + 1] = h->offs_h[p] + ncells[p];
    h->tfirst_h[p] = (int)tdesc.size();
    for (long long c = h->offs_h[p]; c < h->offs_h[p + 1]; c += TILE) {
      tdesc.push_back({c, h->offs_h[p], h->offs_h[p + 1], p, 0});
      h->tprob_h.push_back(p);
    }
    if (tdesc.size() > 0x7fffffffULL) { delete h; return fail(HES_ERR_ARG, "too many cells"); }
  }
  h->tfirst_h[nproblems] = (int)tdesc.size();
  h->ntiles = (int)tdesc.size();
  h->N = h->offs_h[nproblems];
  if (h->N >= (1LL << 28)) { delete h; return fail(HES_ERR_ARG, "at most 2^28 - 1 cells in all"); }
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1) {
      delete h;
      return fail(HES_ERR_HIP, "cannot read the device's compute-unit count");
    }
    h->nbr = std::min(h->ntiles, 4 * cus);  // one reverse wave per SIMD
  }
  h->pitch = (h->N + 63) / 64 * 64 + 256;  // rows padded off a power-of-two stride, as the solver's
  const size_t Pt = (size_t)h->pitch;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  hipEvent_t* sync_events[] = {&h->ev_fork, &h->ev_join, &h->ev_ready[0], &h->ev_ready[1],
                               &h->ev_done[0], &h->ev_done[1]};
  for (hipEvent_t* ev : sync_events)
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  auto up = [&](auto** d, const void* src, size_t count) {
    using T = std::remove_reference_t<decltype(**d)>;
    if (e == hipSuccess) e = balloc(d, count);
    if (e == hipSuccess && src) e = copy_on(h->stream, *d, src, sizeof(T) * count, hipMemcpyHostToDevice);
  };
  up(&h->tiles, tdesc.data(), tdesc.size());
  up(&h->tfirst, h->tfirst_h.data(), h->tfirst_h.size());
  up(&h->offs, h->offs_h.data(), h->offs_h.size());
  up(&h->mats, mats.data(), mats.size());
  up(&h->dx, vdx.data(), vdx.size());
  up(&h->gin, gin.data(), gin.size());
  up(&h->ck, nullptr, (size_t)h->nck * 15 * Pt);
  for (int b = 0; b < 2 && h->every > 1; ++b) up(&h->seg[b], nullptr, (size_t)(h->every - 1) * 15 * Pt);
  for (int b = 0; b < 2; ++b) {
    up(&h->fq[b], nullptr, 14 * Pt);
    up(&h->fc[b], nullptr, Pt);
    up(&h->qb[b], nullptr, 14 * Pt);
  }
  up(&h->zf, nullptr, 2 * (size_t)h->ntiles);
  up(&h->lmax, nullptr, (size_t)(nsteps + 1) * nproblems);
  up(&h->blk_val, nullptr, (size_t)nsteps * h->ntiles);
  up(&h->blk_idx, nullptr, (size_t)nsteps * h->ntiles);
  up(&h->istar, nullptr, (size_t)nsteps * nproblems);
  up(&h->Jlam, nullptr, (size_t)nsteps * nproblems * JL);
  up(&h->dtp, nullptr, 2 * (size_t)h->ntiles);
  up(&h->thb, nullptr, 2 * (size_t)h->ntiles);
  up(&h->lossp, nullptr, (size_t)h->ntiles);
  up(&h->rho_t, nullptr, Pt);
  up(&h->u_t, nullptr, Pt);
  up(&h->gmat, nullptr, 2 * (size_t)nproblems);
  up(&h->status, nullptr, 4);
  if (e == hipSuccess) e = zero_on(h->stream, h->status, 4 * sizeof(int));
  if (e == hipSuccess) {
    std::vector<double*> sl;
    for (long long c = 1; c < h->nck; ++c) sl.push_back(bck_q(h, c));
    for (int b = 0; b < 2; ++b)
      for (long long j = 0; j + 1 < h->every; ++j) sl.push_back(h->seg[b] + j * 15 * Pt);
    sl.push_back(h->fq[0]);
    sl.push_back(h->fq[1]);
    h->nslots = (int)sl.size();
    up(&h->slots, sl.data(), sl.size());
  }
  if (e != hipSuccess) {
    g_err = std::string("batched adjoint allocation failed: ") + hipGetErrorString(e);
    hes_badjoint_destroy(h);
    return HES_ERR_HIP;
  }
  *out = h;
  return HES_OK;
}

int hes_badjoint_destroy(hes_badjoint* h) {
  if (!h) return HES_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  drop_graphs(h);
  void* bufs[] = {h->tiles, h->tfirst, h->offs, h->mats, h->dx, h->gin, h->ck, h->seg[0],
                  h->seg[1], h->fq[0], h->fq[1], h->fc[0], h->fc[1], h->qb[0], h->qb[1], h->zf,
                  h->lmax, h->blk_val, h->blk_idx, h->istar, h->Jlam, h->dtp, h->thb, h->lossp,
                  h->rho_t, h->u_t, h->slots, h->status, h->poff, h->piece_prob, h->bounds,
                  h->prims, h->states, h->sums, h->gstates, h->gmat};
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

int hes_badjoint_set_initial(hes_badjoint* h, const int* nstates, const double* prims,
                             const double* interfaces) {
  if (!h || !nstates || !prims) return fail(HES_ERR_ARG, "null argument");
  std::vector<int> poff(h->P + 1);
  poff[0] = 0;
  long long nifc = 0;
  for (int p = 0; p < h->P; ++p) {
    if (nstates[p] < 1) return fail(HES_ERR_ARG, "every problem needs at least one piece");
    if ((long long)poff[p] + nstates[p] > 0x7fffffffLL) return fail(HES_ERR_ARG, "too many pieces");
    poff[p + 1] = poff[p] + nstates[p];
    nifc += nstates[p] - 1;
  }
  const int npieces = poff[h->P];
  if (nifc > 0 && !interfaces) return fail(HES_ERR_ARG, "null interfaces");
  for (long long k = 0; k < (long long)NP * npieces; ++k)
    if (!std::isfinite(prims[k])) return fail(HES_ERR_ARG, "non-finite primitive state");
  // problem p's piece j covers its cells i (centre x = (i + 1/2) / n_p) with
  // #{interfaces of p below x} = j, as the one-problem fill does
  std::vector<long long> bounds(npieces + 1);
  std::vector<int> pprob(npieces);
  long long ip = 0;
  for (int p = 0; p < h->P; ++p) {
    const long long n = h->offs_h[p + 1] - h->offs_h[p];
    const double* ifc = interfaces ? interfaces + ip : nullptr;
    for (int j = 0; j + 1 < nstates[p] - 1; ++j)
      if (!(ifc[j] <= ifc[j + 1])) return fail(HES_ERR_ARG, "interfaces not sorted");
    bounds[poff[p]] = h->offs_h[p];
    for (int j = 1; j < nstates[p]; ++j) {
      long long lo = 0, hi = n;  // first i with x_i > s_{j-1}
      while (lo < hi) {
        const long long mid = (lo + hi) / 2;
        const double x = ((double)mid + 0.5) / (double)n;
        if (x > ifc[j - 1]) hi = mid;
        else lo = mid + 1;
      }
      bounds[poff[p] + j] = h->offs_h[p] + lo;
    }
    for (int j = 0; j < nstates[p]; ++j) pprob[poff[p] + j] = p;
    ip += nstates[p] - 1;
  }
  bounds[npieces] = h->N;
  (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  // until the new pieces are in place the handle has no initial state (a
  // failed allocation or copy leaves it so)
  h->has_init = false;
  h->fwd_done = false;
  if (npieces != h->npieces) {
    void* old[] = {h->poff, h->piece_prob, h->bounds, h->prims, h->states, h->sums, h->gstates};
    for (void* q : old)
      if (q) (void)hipFree(q);
    h->poff = h->piece_prob = nullptr;
    h->bounds = nullptr;
    h->prims = h->states = h->sums = h->gstates = nullptr;
    h->npieces = 0;
    hipError_t e = balloc(&h->poff, (size_t)h->P + 1);
    if (e == hipSuccess) e = balloc(&h->piece_prob, (size_t)npieces);
    if (e == hipSuccess) e = balloc(&h->bounds, (size_t)npieces + 1);
    if (e == hipSuccess) e = balloc(&h->prims, (size_t)NP * npieces);
    if (e == hipSuccess) e = balloc(&h->states, (size_t)NQ * npieces);
    if (e == hipSuccess) e = balloc(&h->sums, (size_t)NQ * npieces);
    if (e == hipSuccess) e = balloc(&h->gstates, (size_t)NP * npieces);
    if (e != hipSuccess) {
      g_err = std::string("batched adjoint allocation failed: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
    h->npieces = npieces;
  }
  h->poff_h = poff;
  HIP_TRY(copy_on(h->stream, h->poff, poff.data(), sizeof(int) * poff.size(), hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->piece_prob, pprob.data(), sizeof(int) * pprob.size(),
                  hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->bounds, bounds.data(), sizeof(long long) * bounds.size(),
                  hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->prims, prims, sizeof(double) * NP * npieces, hipMemcpyHostToDevice));
  h->has_init = true;
  h->fwd_done = false;
  return HES_OK;
}

int hes_badjoint_set_target(hes_badjoint* h, const double* rho, const double* u1) {
  if (!h || !rho || !u1) return fail(HES_ERR_ARG, "null argument");
  HIP_TRY(copy_on(h->stream, h->rho_t, rho, sizeof(double) * h->N, hipMemcpyHostToDevice));
  HIP_TRY(copy_on(h->stream, h->u_t, u1, sizeof(double) * h->N, hipMemcpyHostToDevice));
  h->has_target = true;
  h->fwd_done = false;
  return HES_OK;
}

int hes_badjoint_forward(hes_badjoint* h, double* losses, double* total, float* elapsed_ms) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->has_init || !h->has_target) return fail(HES_ERR_ARG, "initial states and targets must be set");
  HIP_TRY(zero_on(h->stream, h->status, 4 * sizeof(int)));
  if (h->use_graph && !h->gf) {
    hipError_t e = bcapture(h, true, &h->gf);
    if (e != hipSuccess) {
      g_err = std::string("forward graph: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (h->use_graph) HIP_TRY(hipGraphLaunch(h->gf, h->stream));
  else HIP_TRY(brecord_forward(h));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  int st[2];
  HIP_TRY(copy_on(h->stream, st, h->status, sizeof(st), hipMemcpyDeviceToHost));
  if (st[0]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  std::vector<double> lp(h->ntiles);
  HIP_TRY(copy_on(h->stream, lp.data(), h->lossp, sizeof(double) * lp.size(), hipMemcpyDeviceToHost));
  std::vector<double> per(h->P, 0.0);
  for (int t = 0; t < h->ntiles; ++t) per[h->tprob_h[t]] += lp[t];
  double tot = 0.0;
  for (double v : per) tot += v;
  if (losses) std::memcpy(losses, per.data(), sizeof(double) * h->P);
  if (total) *total = tot;
  h->fwd_done = true;
  return HES_OK;
}

int hes_badjoint_backward(hes_badjoint* h, double* grad_states, double* grad_material,
                          float* elapsed_ms) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  if (h->use_graph && !h->gb) {
    hipError_t e = bcapture(h, false, &h->gb);
    if (e != hipSuccess) {
      g_err = std::string("backward graph: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  if (h->use_graph) HIP_TRY(hipGraphLaunch(h->gb, h->stream));
  else HIP_TRY(brecord_backward(h));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  std::vector<double> gs((size_t)NP * h->npieces), gm(2 * (size_t)h->P);
  HIP_TRY(copy_on(h->stream, gs.data(), h->gstates, sizeof(double) * gs.size(), hipMemcpyDeviceToHost));
  HIP_TRY(copy_on(h->stream, gm.data(), h->gmat, sizeof(double) * gm.size(), hipMemcpyDeviceToHost));
  if (grad_states) std::memcpy(grad_states, gs.data(), sizeof(double) * gs.size());
  if (grad_material) std::memcpy(grad_material, gm.data(), sizeof(double) * gm.size());
  for (double v : gs)
    if (!std::isfinite(v)) return fail(HES_ERR_DOMAIN, "non-finite gradient");
  for (double v : gm)
    if (!std::isfinite(v)) return fail(HES_ERR_DOMAIN, "non-finite gradient");
  return HES_OK;
}

int hes_badjoint_set_sparse(hes_badjoint* h, int on) {
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

int hes_badjoint_set_launch(hes_badjoint* h, int use_graph) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  h->use_graph = use_graph != 0;
  return HES_OK;
}

int hes_badjoint_get_state(hes_badjoint* h, double* Q) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  double *qn, *cn;
  bfwd_slot(h, h->nsteps, &qn, &cn);
  std::vector<double> buf((size_t)14 * h->pitch);
  HIP_TRY(copy_on(h->stream, buf.data(), qn, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
  for (long long i = 0; i < h->N; ++i)
    for (int r = 0; r < 14; ++r) Q[i * 14 + r] = buf[r * h->pitch + i];
  return HES_OK;
}

int hes_badjoint_get_profiles(hes_badjoint* h, double* rho, double* u1) {
  if (!h || !rho || !u1) return fail(HES_ERR_ARG, "null argument");
  if (!h->fwd_done) return fail(HES_ERR_ARG, "run the forward pass first");
  double *qn, *cn;
  bfwd_slot(h, h->nsteps, &qn, &cn);
  double *dr = nullptr, *du = nullptr;
  HIP_TRY(balloc(&dr, (size_t)h->N));
  hipError_t e = balloc(&du, (size_t)h->N);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_badj_profiles, dim3((unsigned)h->ntiles), dim3(64), 0, h->stream, h->tiles,
                       qn, h->gin, dr, du, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = copy_on(h->stream, rho, dr, sizeof(double) * h->N, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = copy_on(h->stream, u1, du, sizeof(double) * h->N, hipMemcpyDeviceToHost);
  (void)hipFree(dr);
  if (du) (void)hipFree(du);
  if (e != hipSuccess) {
    g_err = std::string("get_profiles: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return HES_OK;
}

int hes_badjoint_info(hes_badjoint* h, long long* total_cells, long long* device_bytes,
                      long long* checkpoints) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  const long long Pt = h->pitch, T = h->ntiles, P = h->P;
  long long b = h->nck * 15 * Pt + 2 * (h->every - 1) * 15 * Pt + 2 * (14 + 1 + 14) * Pt + 2 * Pt;
  b = b * 8 + (h->nsteps + 1) * P * 8 + h->nsteps * T * 16 + h->nsteps * P * (8 + JL * 8) +
      T * (2 * 4 + 2 * 8 + 2 * 8 + 8);
  if (total_cells) *total_cells = h->N;
  if (device_bytes) *device_bytes = b;
  if (checkpoints) *checkpoints = h->nck;
  return HES_OK;
}

}  // extern "C"
