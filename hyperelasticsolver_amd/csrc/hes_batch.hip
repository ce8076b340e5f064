// hes_batch.hip -- many independent 1D problems of uneven length advanced
// together: one launch per time step covers every problem.
//
// All cells of all problems are packed into one structure-of-arrays state
// (problem p owns cells [offsets[p], offsets[p+1])).  Work is cut into tiles of
// 62 owned cells that never straddle two problems, so within a wave the
// material constants, dx, CFL and dt are uniform and come from scalar loads.
// Each problem keeps its own CFL ring (max lambda of state s in
// slots[3p + s % 3]) and time ring, i.e. its own dt (main.jl:204-214 per
// problem) and its own frozen end cells (main.jl:219-220).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "hes_host.hpp"
#include "hes_step.hpp"
#include "../../include/hes.h"

using namespace hes;

// One tile: its first owned cell and its problem's cell range, in one record
// so a wave's state loads depend on a single (scalar) load.
struct TileDesc {
  long long first, lo, hi;
  int p, pad;
};

__device__ __forceinline__ TileDesc tile_record(const TileDesc* tiles, int t) {
  typedef const __attribute__((address_space(4))) TileDesc ctile_t;
  ctile_t* r = (ctile_t*)tiles + t;
  TileDesc d;
  d.first = r->first;
  d.lo = r->lo;
  d.hi = r->hi;
  d.p = r->p;
  d.pad = 0;
  return d;
}

struct BatchArgs {
  const double* __restrict__ Qin;
  double* __restrict__ Qout;
  const double* __restrict__ Cin;
  double* __restrict__ Cout;
  const TileDesc* __restrict__ tiles;
  const long long* __restrict__ offsets;     // P + 1
  const MatDev* __restrict__ mats;           // nphase per problem
  const double* __restrict__ cfl;
  const double* __restrict__ dx;
  const double* __restrict__ t_end;
  unsigned long long* slots;                 // 3 per problem
  double* tring;                             // 2 per problem
  int* psteps;                               // executed steps per problem
  // per problem: the launch index after its last step, i.e. its state is in
  // Q[plaunch & 1], C[plaunch & 1] and its CFL maximum in slots[3p + plaunch % 3]
  long long* plaunch;
  long long launch;                          // this launch's step index
  int* status;                               // [0] error bits
  long long pitch;
  int tile_begin, tile_end;
  int step;  // ring phase (step index modulo 6)
  int alt;   // odd steps walk the waves from the far end (launches of <= ALT_MAX_CELLS)
  int xcd;   // deal consecutive blocks to one XCD
  int lxf;   // Rusanov kernels: Lax-Friedrichs face speed dx / dt (main.jl:50)
  // tiles per wave: the grid has nwaves = ceil(tiles / tpw) waves and wave w
  // steps through the adjacent tiles w tpw .. w tpw + tpw - 1 (each lane takes
  // tpw cells in turn; a tile's halo is then the tile just done)
  int tpw, nwaves;
  // two phases: the half path integrals of the launch's jump faces
  // (k_bpath2, k_bpath2_hll; face i+1/2 at [r * pitch + i]) and the jump-face
  // list they come from (k_bmark2): face index and problem per entry.  The
  // list has JSEG segments of jcap entries, each with its own counter (the
  // scan's blocks append to segment block % JSEG, so no one counter takes
  // every block's atomic); segment k of slot s counts in
  // jcnt[(s JSEG + k) JSTRIDE].  The slot alternates per launch; each scan
  // clears the other one for the next launch.
  const double* dpath;
  unsigned long long* jcnt;
  long long* jlist;
  int* jprob;
  long long jcap;
  int jslot;
};

#ifndef HES_JSEG
#define HES_JSEG 32
#endif
constexpr int JSEG = HES_JSEG;   // jump-list segments (at most 64)
constexpr int JSTRIDE = 16;      // counters 128 bytes apart
constexpr size_t JCNT_BYTES = 2 * JSEG * JSTRIDE * sizeof(unsigned long long);

// This wave's index in the grid (wave-uniform).
template <int BLK>
__device__ __forceinline__ int batch_wave(const BatchArgs& a) {
  long long b = a.xcd ? xcd_block() : (long long)blockIdx.x;
  return __builtin_amdgcn_readfirstlane((int)((b * BLK + threadIdx.x) >> 6));
}

// Max of a positive CFL speed into a problem's slot.  Fire-and-forget: a read
// of the slot first would keep every wave alive for one more memory round trip,
// and the slot of a problem of up to BIG_PROBLEM cells sees at most about a
// thousand waves per step; larger problems read the slot first (in the step).
constexpr long long BIG_PROBLEM = 1LL << 16;
__device__ __forceinline__ void atomic_max_pos(unsigned long long* slot, double m) {
  if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
}

// One step of one tile (wave-uniform `tile`).
template <int NPH, int FLUX, int FAST, int PIN>
__device__ __forceinline__ void batch_tile_step(const BatchArgs& a, int tile) {
  constexpr int NC = NComp<NPH>::value;
  const int lane = threadIdx.x & 63;
  const TileDesc td = tile_record(a.tiles, tile);
  const int p = td.p;
  const long long lo = td.lo, hi = td.hi, first = td.first;
  const bool lead = first == lo;  // this tile does the problem's bookkeeping
  // Two phases: the problem's materials in scalar registers (a local copy,
  // read through the constant address space).  One phase: read in place,
  // which measured faster there.
  MatDev mloc[NPH];
  const MatDev* m = a.mats + NPH * p;
  if constexpr (NPH == 2) {
    typedef const __attribute__((address_space(4))) MatDev cmat_t;
#pragma unroll
    for (int k = 0; k < NPH; ++k) {
      cmat_t* g = (cmat_t*)(m + k);
      MatDev& d = mloc[k];
      d.rho0 = g->rho0; d.irho0 = g->irho0; d.cv = g->cv; d.t0 = g->t0; d.b0sq = g->b0sq;
      d.k0 = g->k0; d.alpha = g->alpha; d.beta = g->beta; d.gamma = g->gamma;
      d.hk0a2 = g->hk0a2; d.k0_2a = g->k0_2a; d.k0_a = g->k0_a; d.cvt0 = g->cvt0;
      d.icvt0 = g->icvt0; d.hb0sq = g->hb0sq; d.half_a = g->half_a; d.half_b = g->half_b;
      d.half_g = g->half_g; d.ac_t = g->ac_t; d.ac_s = g->ac_s; d.ac_e = g->ac_e;
      d.ac_w = g->ac_w; d.fast = g->fast; d.pad = 0;
    }
    m = mloc;
  }
  const ChunkCtx ctx = {m,  a.Qin, a.Qout, a.Cin, a.Cout, a.status, a.pitch, lo, hi, cload(a.cfl + p),
                        cload(a.dx + p), 0.0, NPH == 2 ? a.dpath : nullptr, a.lxf};
  const long long i = first + lane - 1;
  double q[NC], cc[NPH];
  load_chunk<NPH>(ctx, i, q, cc);  // in flight while the problem's time is read
  // this launch writes the other time and CFL slots of the problem
  const double tcur = cload(a.tring + 2 * p + (a.step & 1));
  if (!(tcur < cload(a.t_end + p))) {
    // the problem has reached its end time: it does not step, and its state
    // stays in the buffer of its last step (plaunch), from where
    // batch_carry moves it when the host reads it or changes end times
    if (lead && lane == 0) a.tring[2 * p + ((a.step + 1) & 1)] = tcur;
    return;
  }
  const double lmax = __longlong_as_double((long long)cload(a.slots + 3 * p + (a.step % 3)));
  // two phases: the path terms of the jump faces come from k_bpath2 / k_bpath2_hll
  // one phase, Rusanov: every launch shape forms the input sound speeds from
  // the state and stores none (see chunk_step's NOC), so switching shapes or
  // launch modes never reads a stale stored speed
  double lam = chunk_step<NPH, FLUX, FAST, NPH == 2, PIN != 0, false,
                          NPH == 1 && FLUX == FLUX_RUSANOV>(ctx, i, q, cc, lmax);
  lam = wave_max(lam);
  if (lane == 0) {
    unsigned long long* slot = a.slots + 3 * p + ((a.step + 1) % 3);
    if (hi - lo > BIG_PROBLEM) {
      // thousands of waves share this slot: skip the atomic when a larger
      // maximum is already there (one atomic per wave would serialise on the
      // slot: a 2^24-cell problem took 3.0 instead of 0.8 ms per step)
      if (lam > 0.0) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(lam);
        if (bits > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
          atomicMax(slot, bits);
      }
    } else {
      atomic_max_pos(slot, lam);
    }
    if (lead) {
      const double dt = ctx.cfl * ctx.dx / lmax;
      a.tring[2 * p + ((a.step + 1) & 1)] = tcur + dt;
      a.slots[3 * p + ((a.step + 2) % 3)] = 0ull;
      a.psteps[p] += 1;
      a.plaunch[p] = a.launch + 1;  // the problem's state now lives in buffer (launch + 1) & 1
    }
  }
}

// LOOP = 0: one tile per wave (tpw = 1); LOOP = 1: each wave steps through
// tpw tiles.  The loop is a separate instantiation because keeping its
// bookkeeping live costs registers (the pinned HLL step would drop from four
// waves per SIMD to three).
template <int NPH, int FLUX, int FAST, int BLK, int PIN = 0, int LOOP = 0>
__global__ __launch_bounds__(BLK)
__attribute__((amdgpu_waves_per_eu(3))) void k_batch_step(BatchArgs a) {
  int w = batch_wave<BLK>(a);
  // the same set of waves either way (see ALT_MAX_CELLS)
  if ((a.step & 1) && a.alt) w = (int)(gridDim.x * (BLK / 64)) - 1 - w;
  if constexpr (!LOOP) {
    const int tile = a.tile_begin + w;
    if (tile >= a.tile_end) return;  // the whole wave (no block-level sync here)
    batch_tile_step<NPH, FLUX, FAST, PIN>(a, tile);
  } else {
#pragma unroll 1
    for (int k = 0; k < a.tpw; ++k) {
      const int tile = a.tile_begin + w * a.tpw + k;
      if (tile >= a.tile_end) return;
      batch_tile_step<NPH, FLUX, FAST, PIN>(a, tile);
    }
  }
}

// Two phases: the faces with a volume-fraction jump in the launch's tiles
// (face i+1/2 for each owned cell i < hi - 1 of a problem that steps),
// appended with their problems to the jump list (segment block % JSEG of
// slot jslot); block 0 clears the other slot for the next launch.  Each wave
// scans MARK_TPW consecutive tiles, issuing all their loads before it tests
// any, and a block appends its waves' faces with one atomic.
#ifndef HES_BMARK_BLK
#define HES_BMARK_BLK 1024
#endif
#ifndef HES_BMARK_TPW
#define HES_BMARK_TPW 4
#endif
constexpr int MARK_BLK = HES_BMARK_BLK;
constexpr int MARK_TPW = HES_BMARK_TPW;

// Offset of this lane's first entry in a list that every wave of a block of
// NT threads appends to, the lane contributing c entries: one atomic per
// block on *count, the waves' shares in block order, the lanes' in lane
// order.  Every thread of the block must call it.
template <int NT>
__device__ __forceinline__ unsigned long long block_append_n(unsigned c, unsigned long long* count) {
  __shared__ unsigned long long share[NT / 64];
  __shared__ unsigned long long base;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) share[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long total = 0ull;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
      const unsigned long long n = share[k];
      share[k] = total;
      total += n;
    }
    base = total ? atomicAdd(count, total) : 0ull;
  }
  __syncthreads();
  return base + share[w] + (unsigned long long)(incl - c);
}

__global__ __launch_bounds__(MARK_BLK) void k_bmark2(BatchArgs a) {
  if (blockIdx.x == 0 && threadIdx.x < JSEG)
    a.jcnt[((a.jslot ^ 1) * JSEG + threadIdx.x) * JSTRIDE] = 0ull;
  const int t0 = __builtin_amdgcn_readfirstlane(
      a.tile_begin + (int)((blockIdx.x * MARK_BLK + threadIdx.x) >> 6) * MARK_TPW);
  const int lane = threadIdx.x & 63;
  double v[MARK_TPW][4];
  long long idx[MARK_TPW];
  int prob[MARK_TPW];
  bool ok[MARK_TPW];
#pragma unroll
  for (int k = 0; k < MARK_TPW; ++k) {
    const int tile = t0 + k;
    ok[k] = false;
    idx[k] = 0;
    prob[k] = 0;
    if (tile < a.tile_end) {
      const TileDesc td = tile_record(a.tiles, tile);
      const long long i = td.first + lane;
      const long long ic = i < td.hi - 1 ? i : td.hi - 2;  // a face of the problem on every lane
      v[k][0] = a.Qin[ic];
      v[k][1] = a.Qin[ic + 1];
      v[k][2] = a.Qin[15 * a.pitch + ic];
      v[k][3] = a.Qin[15 * a.pitch + ic + 1];
      // the problem steps (its times are written by the step kernel only)
      const bool live =
          cload(a.tring + 2 * td.p + (a.step & 1)) < cload(a.t_end + td.p);
      ok[k] = live && lane < TILE && i < td.hi - 1;
      idx[k] = i;
      prob[k] = td.p;
    }
  }
  bool jump[MARK_TPW];
  unsigned c = 0;
#pragma unroll
  for (int k = 0; k < MARK_TPW; ++k) {
    // face_jump2's test
    jump[k] = ok[k] && !(((v[k][1] - v[k][0]) == 0.0) & ((v[k][3] - v[k][2]) == 0.0));
    c += jump[k] ? 1u : 0u;
  }
  const int seg = (int)(blockIdx.x % JSEG);
  unsigned long long pos = block_append_n<MARK_BLK>(c, a.jcnt + (a.jslot * JSEG + seg) * JSTRIDE);
#pragma unroll
  for (int k = 0; k < MARK_TPW; ++k) {
    if (jump[k]) {
      if (pos < (unsigned long long)a.jcap) {
        const long long e = seg * a.jcap + (long long)pos;
        a.jlist[e] = idx[k];
        a.jprob[e] = prob[k];
      }
      ++pos;
    }
  }
}

// The jump-list segments as one sequence: pre[k] (k < JSEG) receives the
// inclusive prefix of the segment counts (each capped at jcap), pre[k] for
// k >= JSEG the total, which is returned.  Every lane of the wave runs it.
__device__ __forceinline__ unsigned jump_prefix(const BatchArgs& a, unsigned* pre) {
  const int lane = threadIdx.x & 63;
  unsigned c = 0;
  if (lane < JSEG) {
    const unsigned long long n = a.jcnt[(a.jslot * JSEG + lane) * JSTRIDE];
    c = (unsigned)min(n, (unsigned long long)a.jcap);
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned u = __shfl_up(c, o);
    if (lane >= o) c += u;
  }
  pre[lane] = c;
  return __shfl(c, 63);
}

// Entry of the j-th listed face (j < total): a binary search of the prefix
// table.  The table is read from LDS, not from other lanes, because lanes
// whose loop has ended are inactive.
__device__ __forceinline__ long long jump_entry(const BatchArgs& a, const unsigned* pre,
                                                unsigned j) {
  int k = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    if (step >= JSEG) continue;
    if (pre[k + step - 1] <= j) k += step;
  }
  const unsigned base = k ? pre[k - 1] : 0u;
  return k * a.jcap + (long long)(j - base);
}

// The path terms of the listed jump faces with each face's own materials,
// HES_PATH_LANES lanes per face for Rusanov (as the one-problem k_path2) and
// HES_BPATH_HLL_LANES for HLL.  A batch lists many jump faces (every
// problem's interfaces and their spread), so the HLL form takes two lanes per
// face, 32 faces per wave, against the one-problem k_path2_hll's four: each
// lane then evaluates more quadrature nodes in series and spills more (184
// against 84 B per lane), but the grid covers the list in half the rounds
// (512 problems, steps 0 .. 20: 307.5-308.1 -> 300.0-305.8 us per step, same
// box).  One wave per block.
#ifndef HES_BPATH_HLL_LANES
#define HES_BPATH_HLL_LANES 2
#endif
template <int FAST>
__global__ __launch_bounds__(64) void k_bpath2(BatchArgs a) {
  __shared__ unsigned pre[64];
  const unsigned cnt = jump_prefix(a, pre);
  constexpr int GL = HES_PATH_LANES;
  const unsigned stride = gridDim.x * (64 / GL);
  for (unsigned j = (blockIdx.x * 64 + threadIdx.x) / GL; j < cnt; j += stride) {
    const long long e = jump_entry(a, pre, j);
    face_path2<FAST>(a.mats + 2 * a.jprob[e], a.Qin, (double*)a.dpath, a.jlist[e], a.pitch,
                     threadIdx.x & (GL - 1));
  }
}

template <int FAST>
__global__ __launch_bounds__(64) void k_bpath2_hll(BatchArgs a) {
  __shared__ unsigned pre[64];
  const unsigned cnt = jump_prefix(a, pre);
  constexpr int GL = HES_BPATH_HLL_LANES;
  const unsigned stride = gridDim.x * (64 / GL);
  for (unsigned j = (blockIdx.x * 64 + threadIdx.x) / GL; j < cnt; j += stride) {
    const long long e = jump_entry(a, pre, j);
    face_path2_hll<FAST, GL>(a.mats + 2 * a.jprob[e], a.Qin, a.Cin, (double*)a.dpath,
                             a.jlist[e], a.pitch, threadIdx.x & (GL - 1));
  }
}

// Sound speeds and per-problem CFL maximum of the current state.
template <int NPH>
__global__ __launch_bounds__(64) void k_batch_speeds(BatchArgs a) {
  constexpr int NC = NComp<NPH>::value;
  const int lane = threadIdx.x & 63;
  const int tile = __builtin_amdgcn_readfirstlane(a.tile_begin + (int)blockIdx.x);
  if (tile >= a.tile_end) return;
  const TileDesc td = a.tiles[tile];
  const int p = td.p;
  const long long lo = td.lo, hi = td.hi;
  const long long i = td.first + lane - 1;
  const MatDev* m = a.mats + NPH * p;
  double lam = 0.0;
  int bad = 0;
  if (lane >= 1 && lane <= TILE && i < hi) {
    double q[NC], cc[NPH];
#pragma unroll
    for (int r = 0; r < NC; ++r) q[r] = a.Qin[r * a.pitch + i];
    lam = cell_speeds<NPH, -1>(m, q, cc, bad);
#pragma unroll
    for (int k = 0; k < NPH; ++k) a.Cout[k * a.pitch + i] = cc[k];
  }
  (void)lo;
  if (__any(bad) && lane == 0) atomicOr(a.status, 1);
  lam = wave_max(lam);
  if (lane == 0) atomic_max_pos(a.slots + 3 * p + (a.step % 3), lam);
}

// Piecewise-constant initial state per problem: problem p has nstates[p]
// states (states + nc * soff[p]) and nstates[p]-1 sorted interfaces
// (ifc + soff[p] - p); cell j of the problem (centre (j+1/2)/n_p) takes state
// #{k : centre > s_k}.
__global__ void k_batch_piecewise(const long long* __restrict__ offsets, int P,
                                  const int* __restrict__ soff, const double* __restrict__ states,
                                  const double* __restrict__ ifc, int nc, double* __restrict__ Q,
                                  long long N, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo = 0, hi = P;  // largest p with offsets[p] <= i
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid;
    else hi = mid;
  }
  const int p = lo;
  const long long n = offsets[p + 1] - offsets[p];
  const double x = ((double)(i - offsets[p]) + 0.5) / (double)n;
  const int k0 = soff[p], nk = soff[p + 1] - soff[p];
  const double* s = ifc + (k0 - p);
  int a = 0, b = nk - 1;  // first k with !(x > s_k)
  while (a < b) {
    int mid = (a + b) >> 1;
    if (x > s[mid]) a = mid + 1;
    else b = mid;
  }
  const double* src = states + (long long)(k0 + a) * nc;
  for (int r = 0; r < nc; ++r) Q[r * pitch + i] = src[r];
}

// The end cells of every problem never change: copy them to the other buffer.
__global__ void k_batch_copy_ends(const long long* __restrict__ offsets, int P, int nc,
                                  const double* __restrict__ src, double* __restrict__ dst,
                                  long long pitch) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  long long a = offsets[p], b = offsets[p + 1] - 1;
  for (int r = 0; r < nc; ++r) {
    dst[r * pitch + a] = src[r * pitch + a];
    dst[r * pitch + b] = src[r * pitch + b];
  }
}

// Problems that reached their end time stop stepping while the batch goes on,
// so each one's state is left in the buffers of its last step (plaunch).
// Before the host reads the states or changes end times, k_batch_carry copies
// the owned cells and sound speeds of every such problem whose buffer differs
// from that of the current launch index `launch` (one wave per tile: Qout and
// Cout of the arguments are the other buffers, Qin and Cin the current ones),
// and k_batch_carry_fin then moves its CFL maximum into the current slot and
// records it as current.  Two launches, so that no tile reads plaunch after
// its problem's entry changed.
template <int NPH>
__global__ __launch_bounds__(64) void k_batch_carry(BatchArgs a) {
  constexpr int NC = NComp<NPH>::value;
  const int tile = __builtin_amdgcn_readfirstlane(a.tile_begin + (int)blockIdx.x);
  if (tile >= a.tile_end) return;
  const TileDesc td = a.tiles[tile];
  if (!((a.plaunch[td.p] ^ a.launch) & 1)) return;  // already in the current buffer
  const int lane = threadIdx.x & 63;
  const long long i = td.first + lane;
  if (lane >= TILE || i >= td.hi) return;
  double* Q = const_cast<double*>(a.Qin);
  double* C = const_cast<double*>(a.Cin);
#pragma unroll
  for (int r = 0; r < NC; ++r) Q[r * a.pitch + i] = a.Qout[r * a.pitch + i];
#pragma unroll
  for (int k = 0; k < NPH; ++k) C[k * a.pitch + i] = a.Cout[k * a.pitch + i];
}

__global__ void k_batch_carry_fin(unsigned long long* slots, long long* plaunch, int P,
                                  long long launch) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const long long L = plaunch[p];
  if (L == launch) return;
  // the problem's last step max-reduced the CFL speeds of its state into
  // slot L % 3, and nothing has written its slots since
  const unsigned long long v = slots[3 * p + L % 3];
  slots[3 * p + 0] = 0ull;
  slots[3 * p + 1] = 0ull;
  slots[3 * p + 2] = 0ull;
  slots[3 * p + launch % 3] = v;
  plaunch[p] = launch;
}

__global__ void k_soa_to_aos_b(const double* __restrict__ in, double* __restrict__ out, int nc,
                               long long n, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < nc; ++r) out[i * nc + r] = in[r * pitch + i];
}
__global__ void k_aos_to_soa_b(const double* __restrict__ in, double* __restrict__ out, int nc,
                               long long n, long long pitch) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < nc; ++r) out[r * pitch + i] = in[i * nc + r];
}

struct hes_batch {
  int nph, flux, nc, P, ntiles;
  long long N, pitch;
  bool fast;
  int block, xcd, tpw;  // launch shape (autotuned)
  std::vector<long long> offsets_h;
  std::vector<int> tile_first_h;  // first tile of each problem (P + 1)
  double *Q[2], *C[2];
  TileDesc* tiles;
  long long* offsets;
  MatDev* mats;
  double *cfl, *dx, *t_end, *tring;
  unsigned long long* slots;
  int *psteps, *status;
  long long* plaunch;  // per problem: the launch index after its last step
  // two phases: path-term buffer and jump-face list (see BatchArgs)
  double* dpath;
  unsigned long long* jcnt;
  long long* jlist;
  int* jprob;
  long long jcap;  // entries per list segment
  int jpar;  // count slot of the next launch
  long long step;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  bool has_state;
};

namespace {

BatchArgs batch_args(hes_batch* h, long long s, int tb, int te) {
  BatchArgs a;
  a.Qin = h->Q[s & 1];
  a.Qout = h->Q[(s + 1) & 1];
  a.Cin = h->C[s & 1];
  a.Cout = h->C[(s + 1) & 1];
  a.tiles = h->tiles;
  a.offsets = h->offsets;
  a.mats = h->mats;
  a.cfl = h->cfl;
  a.dx = h->dx;
  a.t_end = h->t_end;
  a.slots = h->slots;
  a.tring = h->tring;
  a.psteps = h->psteps;
  a.plaunch = h->plaunch;
  a.launch = s;
  a.status = h->status;
  a.pitch = h->pitch;
  a.tile_begin = tb;
  a.tile_end = te;
  a.step = (int)(s % 6);
  a.alt = (long long)(te - tb) * TILE <= ALT_MAX_CELLS;
  a.xcd = h->xcd;
  a.lxf = h->flux == HES_FLUX_LXF;
  a.tpw = h->tpw;
  a.nwaves = 0;  // set per launch
  a.dpath = h->dpath;
  a.jcnt = h->jcnt;
  a.jlist = h->jlist;
  a.jprob = h->jprob;
  a.jcap = h->jcap;
  a.jslot = h->jpar;
  return a;
}

template <int NPH, int FLUX, int FAST, int PIN, int LOOP>
void launch_shape(const BatchArgs& a, int block, hipStream_t st) {
  const unsigned nw = (unsigned)a.nwaves;
  if (block == 64)
    hipLaunchKernelGGL((k_batch_step<NPH, FLUX, FAST, 64, PIN, LOOP>), dim3(nw), dim3(64), 0, st, a);
  else
    hipLaunchKernelGGL((k_batch_step<NPH, FLUX, FAST, 256, PIN, LOOP>), dim3((nw + 3) / 4), dim3(256),
                       0, st, a);
}

template <int NPH, int FLUX, int FAST>
void launch_batch_t(const BatchArgs& a0, int block, hipStream_t st) {
  const int tiles = a0.tile_end - a0.tile_begin;
  if (tiles <= 0) return;
  BatchArgs a = a0;
  a.nwaves = (tiles + a.tpw - 1) / a.tpw;  // waves of the grid
  const bool loop = a.tpw > 1;
  if constexpr (NPH == 1) {
    // launches over 2^21 cells and more (a whole batch; not the one-problem
    // launches): the pinned schedule of chunk_step, four waves per SIMD
    if ((long long)tiles * TILE >= (1LL << 21)) {
      loop ? launch_shape<NPH, FLUX, FAST, 1, 1>(a, block, st)
           : launch_shape<NPH, FLUX, FAST, 1, 0>(a, block, st);
      return;
    }
  }
  loop ? launch_shape<NPH, FLUX, FAST, 0, 1>(a, block, st) : launch_shape<NPH, FLUX, FAST, 0, 0>(a, block, st);
}

void launch_batch(hes_batch* h, const BatchArgs& a) {
  // Lax-Friedrichs runs the Rusanov kernels with the fixed face speed (a.lxf)
  if (h->nph == 1) {
    if (h->flux != HES_FLUX_HLL)
      h->fast ? launch_batch_t<1, FLUX_RUSANOV, 1>(a, h->block, h->stream)
              : launch_batch_t<1, FLUX_RUSANOV, 0>(a, h->block, h->stream);
    else
      h->fast ? launch_batch_t<1, FLUX_HLL, 1>(a, h->block, h->stream)
              : launch_batch_t<1, FLUX_HLL, 0>(a, h->block, h->stream);
  } else {
    const int tiles = a.tile_end - a.tile_begin;
    if (tiles <= 0) return;
    // jump faces first (a light scan), then their path terms over a fixed
    // grid (as many waves as stay resident: two per SIMD for Rusanov, one
    // for HLL), then the step
    hipStream_t st = h->stream;
    constexpr int TPB = MARK_BLK / 64 * MARK_TPW;  // tiles per scan block
    hipLaunchKernelGGL(k_bmark2, dim3((unsigned)((tiles + TPB - 1) / TPB)), dim3(MARK_BLK), 0, st,
                       a);
    const long long nbf = ((long long)tiles * TILE + 254) / 256;
    const unsigned nbl = (unsigned)std::min<long long>(
        nbf, h->flux != HES_FLUX_HLL ? 2 * 1024 : 1024);
    if (h->flux != HES_FLUX_HLL) {
      if (h->fast) hipLaunchKernelGGL(k_bpath2<1>, dim3(nbl), dim3(64), 0, st, a);
      else hipLaunchKernelGGL(k_bpath2<0>, dim3(nbl), dim3(64), 0, st, a);
      h->fast ? launch_batch_t<2, FLUX_RUSANOV, 1>(a, h->block, st)
              : launch_batch_t<2, FLUX_RUSANOV, 0>(a, h->block, st);
    } else {
      if (h->fast) hipLaunchKernelGGL(k_bpath2_hll<1>, dim3(nbl), dim3(64), 0, st, a);
      else hipLaunchKernelGGL(k_bpath2_hll<0>, dim3(nbl), dim3(64), 0, st, a);
      h->fast ? launch_batch_t<2, FLUX_HLL, 1>(a, h->block, st)
              : launch_batch_t<2, FLUX_HLL, 0>(a, h->block, st);
    }
    h->jpar ^= 1;  // this launch's scan cleared the other slot
  }
}

// Every problem's state, sound speeds and CFL maximum into the buffers and
// slot of the current step index (see k_batch_carry).
int batch_carry(hes_batch* h) {
  BatchArgs a = batch_args(h, h->step, 0, h->ntiles);
  if (h->nph == 1)
    hipLaunchKernelGGL(k_batch_carry<1>, dim3((unsigned)h->ntiles), dim3(64), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_batch_carry<2>, dim3((unsigned)h->ntiles), dim3(64), 0, h->stream, a);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_batch_carry_fin, dim3((unsigned)((h->P + 127) / 128)), dim3(128), 0,
                     h->stream, h->slots, h->plaunch, h->P, h->step);
  HIP_TRY(hipGetLastError());
  return HES_OK;
}

int batch_init_speeds(hes_batch* h) {
  long long s = h->step;
  HIP_TRY(hipMemsetAsync(h->slots, 0, sizeof(unsigned long long) * 3 * h->P, h->stream));
  BatchArgs a = batch_args(h, s, 0, h->ntiles);
  a.Cout = h->C[s & 1];
  if (h->nph == 1)
    hipLaunchKernelGGL(k_batch_speeds<1>, dim3((unsigned)h->ntiles), dim3(64), 0, h->stream, a);
  else
    hipLaunchKernelGGL(k_batch_speeds<2>, dim3((unsigned)h->ntiles), dim3(64), 0, h->stream, a);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_batch_copy_ends, dim3((unsigned)((h->P + 127) / 128)), dim3(128), 0, h->stream,
                     h->offsets, h->P, h->nc, h->Q[s & 1], h->Q[(s + 1) & 1], h->pitch);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  int st = 0;
  HIP_TRY(copy_on(h->stream, &st, h->status, sizeof(int), hipMemcpyDeviceToHost));
  if (st) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

int batch_after_set(hes_batch* h, const double* t0) {
  std::vector<double> tr(2 * h->P);
  for (int p = 0; p < h->P; ++p) tr[2 * p] = tr[2 * p + 1] = t0 ? t0[p] : 0.0;
  HIP_TRY(copy_on(h->stream, h->tring, tr.data(), sizeof(double) * 2 * h->P, hipMemcpyHostToDevice));
  HIP_TRY(zero_on(h->stream, h->psteps, sizeof(int) * h->P));
  HIP_TRY(zero_on(h->stream, h->plaunch, sizeof(long long) * h->P));
  HIP_TRY(zero_on(h->stream, h->status, sizeof(int) * 4));
  if (h->jcnt) HIP_TRY(zero_on(h->stream, h->jcnt, JCNT_BYTES));
  h->jpar = 0;
  h->step = 0;
  h->has_state = true;
  return batch_init_speeds(h);
}

}  // namespace

extern "C" {

int hes_batch_destroy(hes_batch* h) {
  if (!h) return HES_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ptrs[] = {h->Q[0], h->Q[1], h->C[0], h->C[1], h->tiles, h->offsets, h->mats, h->cfl,
                  h->dx, h->t_end, h->tring, h->slots, h->psteps, h->status, h->dpath, h->jcnt,
                  h->jlist, h->jprob, h->plaunch};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return HES_OK;
}

int hes_batch_create(hes_batch** out, int nphase, int flux, int nproblems, const long long* ncells,
                     const double* materials, const double* cfl, const double* dx) {
  if (!out) return fail(HES_ERR_ARG, "null handle pointer");
  *out = nullptr;
  if (nphase != 1 && nphase != 2) return fail(HES_ERR_ARG, "nphase must be 1 or 2");
  if (flux != HES_FLUX_RUSANOV && flux != HES_FLUX_HLL && flux != HES_FLUX_LXF)
    return fail(HES_ERR_ARG, "unknown flux");
  if (nproblems < 1 || !ncells || !materials) return fail(HES_ERR_ARG, "null argument");
  hes_batch* h = new hes_batch();
  h->nph = nphase;
  h->flux = flux;
  h->nc = nphase == 1 ? 14 : 30;
  h->P = nproblems;
  h->block = 256;
  h->xcd = 1;
  h->tpw = 1;
  h->offsets_h.resize(nproblems + 1);
  h->tile_first_h.resize(nproblems + 1);
  h->offsets_h[0] = 0;
  std::vector<long long> tstart;
  std::vector<int> tprob;
  std::vector<MatDev> mats((size_t)nproblems * nphase);
  std::vector<double> vcfl(nproblems), vdx(nproblems), vte(nproblems, INFINITY);
  h->fast = true;
  for (int p = 0; p < nproblems; ++p) {
    if (ncells[p] < 3) { delete h; return fail(HES_ERR_ARG, "every problem needs at least 3 cells"); }
    for (int k = 0; k < nphase; ++k) {
      const double* mp = materials + 8 * ((size_t)nphase * p + k);
      if (!mat_ok(mp)) { delete h; return fail(HES_ERR_ARG, "invalid material constants"); }
      mats[(size_t)nphase * p + k] = make_mat(mp);
      h->fast = h->fast && mats[(size_t)nphase * p + k].fast;
    }
    vcfl[p] = cfl ? cfl[p] : 0.6;
    vdx[p] = dx ? dx[p] : 1.0 / (double)ncells[p];
    if (!(vcfl[p] > 0) || !(vdx[p] > 0)) { delete h; return fail(HES_ERR_ARG, "cfl and dx must be positive"); }
    h->offsets_h[p + 1] = h->offsets_h[p] + ncells[p];
    h->tile_first_h[p] = (int)tstart.size();
    for (long long c = h->offsets_h[p]; c < h->offsets_h[p + 1]; c += TILE) {
      tstart.push_back(c);
      tprob.push_back(p);
    }
  }
  h->tile_first_h[nproblems] = (int)tstart.size();
  h->ntiles = (int)tstart.size();
  h->N = h->offsets_h[nproblems];
  h->pitch = (h->N + 63) / 64 * 64 + 256;  // rows padded off a power-of-two stride, as the solver's
  size_t qb = sizeof(double) * (size_t)h->nc * h->pitch, cb = sizeof(double) * (size_t)nphase * h->pitch;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    e = hipMalloc(&h->Q[b], qb);
    if (e == hipSuccess) e = hipMalloc(&h->C[b], cb);
  }
  auto up = [&](void** d, const void* src, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(d, bytes);
    if (e == hipSuccess && src) e = copy_on(h->stream, *d, src, bytes, hipMemcpyHostToDevice);
  };
  std::vector<TileDesc> tdesc(tstart.size());
  for (size_t t = 0; t < tstart.size(); ++t)
    tdesc[t] = {tstart[t], h->offsets_h[tprob[t]], h->offsets_h[tprob[t] + 1], tprob[t], 0};
  up((void**)&h->tiles, tdesc.data(), sizeof(TileDesc) * tdesc.size());
  up((void**)&h->offsets, h->offsets_h.data(), sizeof(long long) * (nproblems + 1));
  up((void**)&h->mats, mats.data(), sizeof(MatDev) * mats.size());
  up((void**)&h->cfl, vcfl.data(), sizeof(double) * nproblems);
  up((void**)&h->dx, vdx.data(), sizeof(double) * nproblems);
  up((void**)&h->t_end, vte.data(), sizeof(double) * nproblems);
  up((void**)&h->tring, nullptr, sizeof(double) * 2 * nproblems);
  up((void**)&h->slots, nullptr, sizeof(unsigned long long) * 3 * nproblems);
  if (nphase == 2) {
    up((void**)&h->dpath, nullptr, 2 * qb);  // D-, D+ rows
    // a scan block lists at most TPB tiles' faces, and segment k takes the
    // blocks k, k + JSEG, ...
    constexpr int TPB = MARK_BLK / 64 * MARK_TPW;
    const long long nblk = (h->ntiles + TPB - 1) / TPB;
    h->jcap = (nblk + JSEG - 1) / JSEG * TPB * TILE;
    up((void**)&h->jcnt, nullptr, JCNT_BYTES);
    up((void**)&h->jlist, nullptr, sizeof(long long) * (size_t)(JSEG * h->jcap));
    up((void**)&h->jprob, nullptr, sizeof(int) * (size_t)(JSEG * h->jcap));
    if (e == hipSuccess) e = zero_on(h->stream, h->jcnt, JCNT_BYTES);
  }
  up((void**)&h->psteps, nullptr, sizeof(int) * nproblems);
  up((void**)&h->plaunch, nullptr, sizeof(long long) * nproblems);
  up((void**)&h->status, nullptr, sizeof(int) * 4);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e != hipSuccess) {
    g_err = std::string("batch allocation failed: ") + hipGetErrorString(e);
    hes_batch_destroy(h);
    return HES_ERR_HIP;
  }
  *out = h;
  return HES_OK;
}

int hes_batch_info(hes_batch* h, long long* total_cells, int* ntiles, int* block, int* xcd,
                   int* tiles_per_wave) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (total_cells) *total_cells = h->N;
  if (ntiles) *ntiles = h->ntiles;
  if (block) *block = h->block;
  if (xcd) *xcd = h->xcd;
  if (tiles_per_wave) *tiles_per_wave = h->tpw;
  return HES_OK;
}

int hes_batch_set_t_end(hes_batch* h, const double* t_end) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  // a problem that stopped resumes from the current buffers
  if (h->has_state) {
    const int rc = batch_carry(h);
    if (rc != HES_OK) return rc;
  }
  std::vector<double> v(h->P, INFINITY);
  if (t_end)
    for (int p = 0; p < h->P; ++p) v[p] = t_end[p];
  HIP_TRY(copy_on(h->stream, h->t_end, v.data(), sizeof(double) * h->P, hipMemcpyHostToDevice));
  return HES_OK;
}

int hes_batch_set_state(hes_batch* h, const double* Q, const double* t0) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  size_t bytes = sizeof(double) * (size_t)h->nc * h->N;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  hipError_t e = copy_on(h->stream, tmp, Q, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_aos_to_soa_b, dim3((unsigned)((h->N + 255) / 256)), dim3(256), 0, h->stream,
                       tmp, h->Q[0], h->nc, h->N, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    g_err = std::string("batch set_state: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return batch_after_set(h, t0);
}

int hes_batch_set_piecewise(hes_batch* h, const int* nstates, const double* states,
                            const double* interfaces, const double* t0) {
  if (!h || !nstates || !states) return fail(HES_ERR_ARG, "null argument");
  std::vector<int> soff(h->P + 1, 0);
  for (int p = 0; p < h->P; ++p) {
    if (nstates[p] < 1) return fail(HES_ERR_ARG, "every problem needs at least one state");
    soff[p + 1] = soff[p] + nstates[p];
  }
  int S = soff[h->P], I = S - h->P;
  if (I > 0 && !interfaces) return fail(HES_ERR_ARG, "null interfaces");
  for (int p = 0; p < h->P; ++p)
    for (int k = soff[p] - p; k + 1 < soff[p + 1] - p - 1; ++k)
      if (!(interfaces[k] <= interfaces[k + 1])) return fail(HES_ERR_ARG, "interfaces not sorted");
  int *dso = nullptr;
  double *dst = nullptr, *dif = nullptr;
  HIP_TRY(hipMalloc(&dso, sizeof(int) * (h->P + 1)));
  hipError_t e = hipMalloc(&dst, sizeof(double) * (size_t)S * h->nc);
  if (e == hipSuccess) e = hipMalloc(&dif, sizeof(double) * (size_t)(I > 0 ? I : 1));
  if (e == hipSuccess)
    e = copy_on(h->stream, dso, soff.data(), sizeof(int) * (h->P + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = copy_on(h->stream, dst, states, sizeof(double) * (size_t)S * h->nc, hipMemcpyHostToDevice);
  if (e == hipSuccess && I > 0)
    e = copy_on(h->stream, dif, interfaces, sizeof(double) * I, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_batch_piecewise, dim3((unsigned)((h->N + 255) / 256)), dim3(256), 0,
                       h->stream, h->offsets, h->P, dso, dst, dif, h->nc, h->Q[0], h->N, h->pitch);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(dso);
  (void)hipFree(dst);
  (void)hipFree(dif);
  if (e != hipSuccess) {
    g_err = std::string("batch set_piecewise: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return batch_after_set(h, t0);
}

int hes_batch_get_state(hes_batch* h, double* Q) {
  if (!h || !Q) return fail(HES_ERR_ARG, "null argument");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  const int rc = batch_carry(h);  // problems that stopped early into the current buffer
  if (rc != HES_OK) return rc;
  size_t bytes = sizeof(double) * (size_t)h->nc * h->N;
  double* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, bytes));
  hipLaunchKernelGGL(k_soa_to_aos_b, dim3((unsigned)((h->N + 255) / 256)), dim3(256), 0, h->stream,
                     h->Q[h->step & 1], tmp, h->nc, h->N, h->pitch);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(Q, tmp, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    g_err = std::string("batch get_state: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  return HES_OK;
}

int hes_batch_get_times(hes_batch* h, double* t, long long* steps) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  std::vector<double> tr(2 * h->P);
  std::vector<int> ps(h->P);
  HIP_TRY(copy_on(h->stream, tr.data(), h->tring, sizeof(double) * 2 * h->P, hipMemcpyDeviceToHost));
  HIP_TRY(copy_on(h->stream, ps.data(), h->psteps, sizeof(int) * h->P, hipMemcpyDeviceToHost));
  for (int p = 0; p < h->P; ++p) {
    if (t) t[p] = tr[2 * p + (h->step & 1)];
    if (steps) steps[p] = ps[p];
  }
  return HES_OK;
}

int hes_batch_set_launch(hes_batch* h, int block, int xcd, int tiles_per_wave) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (block != 64 && block != 256) return fail(HES_ERR_ARG, "block must be 64 or 256");
  if (tiles_per_wave < 1 || tiles_per_wave > 64) return fail(HES_ERR_ARG, "tiles per wave must be 1..64");
  h->block = block;
  h->xcd = xcd ? 1 : 0;
  h->tpw = tiles_per_wave;
  return HES_OK;
}

// mode 0: one batched launch per step; 1: the whole step loop captured in one
// hipGraph and launched once; 2: one launch per problem per step.
int hes_batch_run(hes_batch* h, long long steps, int mode, float* elapsed_ms) {
  if (!h) return fail(HES_ERR_ARG, "null handle");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  if (steps < 0 || mode < 0 || mode > 2) return fail(HES_ERR_ARG, "bad argument");
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  if (mode == 1 && steps > 0) {
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (long long j = 0; j < steps; ++j)
      launch_batch(h, batch_args(h, h->step + j, 0, h->ntiles));
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      g_err = std::string("batch graph: ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  hipError_t e = hipSuccess;
  if (mode == 1 && steps > 0) {
    e = hipGraphLaunch(gexec, h->stream);
  } else {
    for (long long j = 0; j < steps && e == hipSuccess; ++j) {
      if (mode == 0) {
        launch_batch(h, batch_args(h, h->step + j, 0, h->ntiles));
      } else {
        for (int p = 0; p < h->P; ++p)
          launch_batch(h, batch_args(h, h->step + j, h->tile_first_h[p], h->tile_first_h[p + 1]));
      }
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g_err = std::string("batch run: ") + hipGetErrorString(e);
    return HES_ERR_HIP;
  }
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  h->step += steps;
  int st = 0;
  HIP_TRY(copy_on(h->stream, &st, h->status, sizeof(int), hipMemcpyDeviceToHost));
  if (st) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  return HES_OK;
}

// Try every launch shape for `steps` steps on a copy of the current state and
// keep the fastest.  The state, times and CFL rings are restored afterwards.
int hes_batch_autotune(hes_batch* h, int steps, int* best_block, int* best_xcd, int* best_tpw,
                       float* best_ms) {
  if (!h || steps < 1) return fail(HES_ERR_ARG, "bad argument");
  if (!h->has_state) return fail(HES_ERR_ARG, "no state set");
  long long s0 = h->step;
  // every problem's state into the current buffers first: the state saved
  // and restored below is then the whole state
  const int rcc = batch_carry(h);
  if (rcc != HES_OK) return rcc;
  size_t qb = sizeof(double) * (size_t)h->nc * h->pitch, cb = sizeof(double) * (size_t)h->nph * h->pitch;
  double *q0 = nullptr, *c0 = nullptr, *t0 = nullptr;
  unsigned long long* sl0 = nullptr;
  int* ps0 = nullptr;
  long long* pl0 = nullptr;
  const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
  // every buffer is freed on every path: the first failure skips the rest
  hipError_t e = hipMalloc(&q0, qb);
  if (e == hipSuccess) e = hipMalloc(&c0, cb);
  if (e == hipSuccess) e = hipMalloc(&t0, sizeof(double) * 2 * h->P);
  if (e == hipSuccess) e = hipMalloc(&sl0, sizeof(unsigned long long) * 3 * h->P);
  if (e == hipSuccess) e = hipMalloc(&ps0, sizeof(int) * h->P);
  if (e == hipSuccess) e = hipMalloc(&pl0, sizeof(long long) * h->P);
  if (e == hipSuccess) e = copy_on(h->stream, q0, h->Q[s0 & 1], qb, dd);
  if (e == hipSuccess) e = copy_on(h->stream, c0, h->C[s0 & 1], cb, dd);
  if (e == hipSuccess) e = copy_on(h->stream, t0, h->tring, sizeof(double) * 2 * h->P, dd);
  if (e == hipSuccess) e = copy_on(h->stream, sl0, h->slots, sizeof(unsigned long long) * 3 * h->P, dd);
  if (e == hipSuccess) e = copy_on(h->stream, ps0, h->psteps, sizeof(int) * h->P, dd);
  if (e == hipSuccess) e = copy_on(h->stream, pl0, h->plaunch, sizeof(long long) * h->P, dd);
  const int blocks[2] = {64, 256};
  const int tpws[3] = {1, 2, 4};
  float best = 1e30f;
  int bb = h->block, bx = h->xcd, bt = h->tpw, rc = HES_OK;
  if (e != hipSuccess) {
    g_err = std::string("autotune save failed: ") + hipGetErrorString(e);
    rc = HES_ERR_HIP;
  }
  for (int c = 0; c < 12 && rc == HES_OK; ++c) {
    const int bi = c / 6, x = (c / 3) % 2, ti = c % 3;
    h->block = blocks[bi];
    h->xcd = x;
    h->tpw = tpws[ti];
    float ms = 0.f;
    rc = hes_batch_run(h, 1, 0, &ms);  // warm-up
    if (rc == HES_OK) rc = hes_batch_run(h, steps, 0, &ms);
    if (rc == HES_OK && ms < best) { best = ms; bb = blocks[bi]; bx = x; bt = tpws[ti]; }
    // restore the starting state for the next candidate
    h->step = s0;
    if (copy_on(h->stream, h->Q[s0 & 1], q0, qb, dd) != hipSuccess ||
        copy_on(h->stream, h->C[s0 & 1], c0, cb, dd) != hipSuccess ||
        copy_on(h->stream, h->tring, t0, sizeof(double) * 2 * h->P, dd) != hipSuccess ||
        copy_on(h->stream, h->slots, sl0, sizeof(unsigned long long) * 3 * h->P, dd) != hipSuccess ||
        copy_on(h->stream, h->psteps, ps0, sizeof(int) * h->P, dd) != hipSuccess ||
        copy_on(h->stream, h->plaunch, pl0, sizeof(long long) * h->P, dd) != hipSuccess)
      rc = fail(HES_ERR_HIP, "autotune restore failed");
  }
  (void)hipFree(pl0);
  (void)hipFree(q0);
  (void)hipFree(c0);
  (void)hipFree(t0);
  (void)hipFree(sl0);
  (void)hipFree(ps0);
  h->block = bb;
  h->xcd = bx;
  h->tpw = bt;
  if (best_block) *best_block = bb;
  if (best_xcd) *best_xcd = bx;
  if (best_tpw) *best_tpw = bt;
  if (best_ms) *best_ms = best / (float)steps;
  return rc;
}

}  // extern "C"
