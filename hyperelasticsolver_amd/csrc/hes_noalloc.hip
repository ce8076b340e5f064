// hes_noalloc.hip -- an entry point that allocates no device memory, for
// callers that cannot call hipMalloc (tight serving loops, harnesses that
// forbid allocation): hes_max_speed on device arrays, together with the CFL
// maximum of the states (main.jl:204-212).
//
// Its kernels need device memory beyond the caller's inputs and outputs: one
// maximum and one status word per block, and the reduced pair.  That memory is
// the file-scope device array `scratch`, of a fixed size, which every GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress gives for the current device (each
// GPU has its own copy at its own address), never through the array's name.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hes_host.hpp"
#include "hes_faces.hpp"
#include "../../include/hes.h"

using namespace hes;

namespace {

constexpr int NA_BLOCK = 256;
constexpr long long NA_MAX_STATES = HES_NOALLOC_MAX_STATES;
constexpr long long NA_MAX_BLOCKS = NA_MAX_STATES / NA_BLOCK;
static_assert(NA_MAX_STATES % NA_BLOCK == 0, "whole blocks");

// Layout of the scratch array (64-bit words):
//   [0]                           maximum over all states (double bits)
//   [1]                           status: nonzero if any state was invalid
//   [2, 2 + NA_MAX_BLOCKS)        maximum of each block's states
//   [2 + NA_MAX_BLOCKS, 2 + 2 NA_MAX_BLOCKS)  status of each block
// A call writes the entries of exactly its own blocks before the reduction
// reads them, so nothing has to start at zero and a smaller call after a
// larger one never reads the larger one's entries.
constexpr long long NA_WORDS = 2 + 2 * NA_MAX_BLOCKS;

}  // namespace

static __device__ __align__(256) unsigned long long scratch[NA_WORDS];

namespace {

// Per state: lambda_i = max_p(|u1_p| + c_p), the speed hes_max_speed returns
// (maximum(abs.(get_eigvals(eos, Q_i, [1,0,0]))), main.jl:209-210); per
// block: the maximum and whether any state was invalid, into the block's own
// scratch entries.
template <int NPH>
__global__ __launch_bounds__(NA_BLOCK) void k_na_speeds(const double* __restrict__ Q,
                                                        double* __restrict__ lam_out, long long n,
                                                        MatDev m0, MatDev m1,
                                                        unsigned long long* __restrict__ sc) {
  constexpr int NC = NComp<NPH>::value;
  __shared__ double red[NA_BLOCK / 64];
  __shared__ int redbad[NA_BLOCK / 64];
  const long long i = (long long)blockIdx.x * NA_BLOCK + threadIdx.x;
  double lam = 0.0;
  int bad = 0;
  if (i < n) {
    const MatDev mat[2] = {m0, m1};
    double q[NC], cc[NPH];
#pragma unroll
    for (int r = 0; r < NC; ++r) q[r] = Q[i * NC + r];
    lam = cell_speeds<NPH>(mat, q, cc, bad);
    if (lam_out) lam_out[i] = lam;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lam = fmax(lam, __shfl_xor(lam, o));
  const int anybad = __any(bad) ? 1 : 0;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = lam;
    redbad[w] = anybad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    int b = redbad[0];
#pragma unroll
    for (int k = 1; k < NA_BLOCK / 64; ++k) {
      m = fmax(m, red[k]);
      b |= redbad[k];
    }
    sc[2 + blockIdx.x] = (unsigned long long)__double_as_longlong(m);
    sc[2 + NA_MAX_BLOCKS + blockIdx.x] = (unsigned long long)b;
  }
}

// One block: the maximum and the status over the first nb blocks' entries.
__global__ __launch_bounds__(NA_BLOCK) void k_na_reduce(unsigned long long* __restrict__ sc,
                                                        long long nb) {
  __shared__ double red[NA_BLOCK / 64];
  __shared__ unsigned long long redbad[NA_BLOCK / 64];
  double m = 0.0;
  unsigned long long b = 0ull;
  for (long long k = threadIdx.x; k < nb; k += NA_BLOCK) {
    m = fmax(m, __longlong_as_double((long long)sc[2 + k]));
    b |= sc[2 + NA_MAX_BLOCKS + k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  const bool anyb = __any(b != 0ull);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = m;
    redbad[w] = anyb ? 1ull : 0ull;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 1; k < NA_BLOCK / 64; ++k) {
      red[0] = fmax(red[0], red[k]);
      redbad[0] |= redbad[k];
    }
    sc[0] = (unsigned long long)__double_as_longlong(red[0]);
    sc[1] = redbad[0];
  }
}

// The scratch array's address on each device, taken once per device with
// that device current (the address differs per GPU).
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
unsigned long long* g_na_ptr[NA_MAX_DEVICES] = {};

int scratch_of_current_device(unsigned long long** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(HES_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return HES_ERR_HIP;
    }
    g_na_ptr[dev] = (unsigned long long*)p;
  }
  *out = g_na_ptr[dev];
  return HES_OK;
}

}  // namespace

extern "C" {

int hes_max_speed_noalloc(int nph, const double* mats, const double* Q_dev, double* lambda_dev,
                          double* lmax, long long n) {
  if (nph != 1 && nph != 2) return fail(HES_ERR_ARG, "nphase must be 1 or 2");
  if (!mats || !lmax || (n > 0 && !Q_dev) || n < 0) return fail(HES_ERR_ARG, "null argument");
  if (n > NA_MAX_STATES)
    return fail(HES_ERR_ARG, "more states than HES_NOALLOC_MAX_STATES (2^26)");
  for (int p = 0; p < nph; ++p)
    if (!mat_ok(mats + 8 * p)) return fail(HES_ERR_ARG, "invalid material constants");
  unsigned long long* sc = nullptr;
  const int rc = scratch_of_current_device(&sc);
  if (rc != HES_OK) return rc;
  *lmax = 0.0;
  if (n == 0) return HES_OK;
  const MatDev m0 = make_mat(mats), m1 = make_mat(nph == 2 ? mats + 8 : mats);
  const long long nb = (n + NA_BLOCK - 1) / NA_BLOCK;
  if (nph == 1)
    hipLaunchKernelGGL(k_na_speeds<1>, dim3((unsigned)nb), dim3(NA_BLOCK), 0, 0, Q_dev, lambda_dev,
                       n, m0, m1, sc);
  else
    hipLaunchKernelGGL(k_na_speeds<2>, dim3((unsigned)nb), dim3(NA_BLOCK), 0, 0, Q_dev, lambda_dev,
                       n, m0, m1, sc);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_na_reduce, dim3(1), dim3(NA_BLOCK), 0, 0, sc, nb);
  HIP_TRY(hipGetLastError());
  unsigned long long res[2] = {0ull, 0ull};
  // a synchronous copy on the null stream: ordered after both kernels
  HIP_TRY(hipMemcpy(res, sc, sizeof(res), hipMemcpyDeviceToHost));
  if (res[1]) return fail(HES_ERR_DOMAIN, "non-positive det(rho A) or non-finite state");
  *lmax = __builtin_bit_cast(double, res[0]);
  return HES_OK;
}

}  // extern "C"
