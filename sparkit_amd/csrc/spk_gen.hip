// spk_gen.hip -- device generator of the power-law SPD matrix (a Zipf(1.5)
// row-length law, symmetrised, diagonally dominant; no reference
// counterpart).  Definition:
//   streams   sL = seed ^ 0x4C454E, sC = seed ^ 0x434F4C, sV = seed ^ 0x56414C
//   u01(z)    = (z >> 11) * 2^-53, mix = splitmix64 element k of a stream
//   lengths   L_i = min{k in [3,4096] : cdf[k] >= u01(mix(sL, i))},
//             cdf[k] = S_k / S_4096, S_k = sum_{j=3..k} pow(j, -1.5)
//   columns   draw t = 0..L_i-2: c = mix(sC, i*4096 + t) mod (n-1),
//             c += (c >= i); duplicates collapse
//   one-sided a(i,j) = -u01(mix(sV, i*n + j))
//   values    A_ij = 0.5 * ([j drawn by i] a(i,j) + [i drawn by j] a(j,i))
//             on the union pattern of draws, transposes and diagonal
//   diagonal  A_ii = (sum over j != i, ascending j, of |A_ij|) + 1
//             (sparkit's dominance rule, matgen.hpp:657-658)
// Build: per-row lengths -> scan -> one key per draw and per transposed draw
// ((row*n + col) << 1 | direction) -> radix sort -> segment heads -> scan ->
// compacted unique pairs with their direction flags -> CSR -> diagonal.
// The CDF table is computed on the host (libm pow) and uploaded.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "spk_csr.hpp"

namespace spk {
namespace {

constexpr int kPlMin = 3, kPlMax = 4096, kPlN = kPlMax - kPlMin + 1;
constexpr uint64_t kStreamL = 0x4C454EULL, kStreamC = 0x434F4CULL,
                   kStreamV = 0x56414CULL;

__device__ __forceinline__ uint64_t mix_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t z) {
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void pl_lengths_kernel(int64_t n, uint64_t seed,
                                  const double* __restrict__ cdf,
                                  int64_t* __restrict__ draws) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double u = u01(mix_at(seed ^ kStreamL, (uint64_t)i));
  int lo = 0, hi = kPlN - 1;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (cdf[mid] >= u) hi = mid; else lo = mid + 1;
  }
  draws[i] = (int64_t)(lo + kPlMin) - 1;
}

__global__ void pl_keys_kernel(int64_t n, uint64_t seed,
                               const int64_t* __restrict__ draws,
                               const int64_t* __restrict__ offs,
                               uint64_t* __restrict__ keys) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t un = (uint64_t)n, ui = (uint64_t)i;
  int64_t pos = 2 * offs[i] + i;
  for (int64_t t = 0; t < draws[i]; ++t) {
    uint64_t c = mix_at(seed ^ kStreamC, ui * kPlMax + (uint64_t)t) % (un - 1);
    if (c >= ui) c += 1;
    keys[pos++] = ((ui * un + c) << 1);
    keys[pos++] = ((c * un + ui) << 1) | 1u;
  }
  keys[pos] = (ui * un + ui) << 1;
}

__global__ void pl_heads_kernel(int64_t K, const uint64_t* __restrict__ keys,
                                int32_t* __restrict__ head) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= K) return;
  head[k] = (k == 0 || (keys[k] >> 1) != (keys[k - 1] >> 1)) ? 1 : 0;
}

__global__ void pl_compact_kernel(int64_t K, const uint64_t* __restrict__ keys,
                                  const int32_t* __restrict__ uidx,
                                  uint64_t* __restrict__ pairs,
                                  uint8_t* __restrict__ has0,
                                  uint8_t* __restrict__ has1) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= K) return;
  const uint64_t key = keys[k], pair = key >> 1;
  const int32_t u = uidx[k] - 1;
  if (k == 0 || (keys[k - 1] >> 1) != pair) {  // first of its segment
    pairs[u] = pair;
    has0[u] = (key & 1u) == 0;
  }
  if (k == K - 1 || (keys[k + 1] >> 1) != pair)  // last of its segment
    has1[u] = (key & 1u) == 1;
}

__global__ void pl_fill_kernel(int64_t U, int64_t n, uint64_t seed,
                               const uint64_t* __restrict__ pairs,
                               const uint8_t* __restrict__ has0,
                               const uint8_t* __restrict__ has1,
                               int32_t* __restrict__ rp,
                               int32_t* __restrict__ ci,
                               double* __restrict__ v) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= U) return;
  const uint64_t un = (uint64_t)n, pair = pairs[q];
  const uint64_t i = pair / un, j = pair % un;
  ci[q] = (int32_t)j;
  if (q == 0 || pairs[q - 1] / un != i) rp[i] = (int32_t)q;
  if (i == j) {
    v[q] = 0.0;  // diagonal, written by pl_diag_kernel
  } else {
    double a = has0[q] ? -u01(mix_at(seed ^ kStreamV, i * un + j)) : 0.0;
    double b = has1[q] ? -u01(mix_at(seed ^ kStreamV, j * un + i)) : 0.0;
    v[q] = 0.5 * (a + b);
  }
}

__global__ void pl_diag_kernel(int64_t n, const int32_t* __restrict__ rp,
                               const int32_t* __restrict__ ci,
                               double* __restrict__ v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  int dpos = rp[i];
  for (int p = rp[i]; p < rp[i + 1]; ++p) {
    if (ci[p] == i) dpos = p;
    else s += fabs(v[p]);
  }
  v[dpos] = s + 1.0;
}

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

spk_csr_s* build_powerlaw(int64_t n, uint64_t seed) {
  SPK_REQUIRE(n > 1 && n < (1LL << 31), SPK_INVALID_ARGUMENT, "bad size");
  hipStream_t s = stream();
  // Zipf(1.5) CDF on the host (sequential sums, libm pow)
  std::vector<double> hcdf(kPlN);
  double acc = 0.0;
  for (int k = kPlMin; k <= kPlMax; ++k) {
    acc += std::pow((double)k, -1.5);
    hcdf[k - kPlMin] = acc;
  }
  for (double& c : hcdf) c /= acc;
  DevBuf<double> cdf(kPlN);
  SPK_HIP_TRY(hipMemcpyAsync(cdf.p, hcdf.data(), kPlN * 8, hipMemcpyHostToDevice, s));

  DevBuf<int64_t> draws((size_t)n), offs((size_t)n + 1);
  hipLaunchKernelGGL(pl_lengths_kernel, dim3(nb(n)), dim3(256), 0, s, n, seed,
                     cdf.p, draws.p);
  size_t tb = 0;
  SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, draws.p, offs.p, n, s));
  DevBuf<unsigned char> tmp(tb);
  SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, draws.p, offs.p, n, s));
  int64_t last_off = 0, last_draw = 0;
  SPK_HIP_TRY(hipMemcpyAsync(&last_off, offs.p + n - 1, 8, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipMemcpyAsync(&last_draw, draws.p + n - 1, 8, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  const int64_t D = last_off + last_draw;
  const int64_t K = 2 * D + n;
  SPK_REQUIRE(K < (1LL << 31), SPK_INVALID_ARGUMENT, "too many entries");

  DevBuf<uint64_t> keys_in((size_t)K), keys((size_t)K);
  hipLaunchKernelGGL(pl_keys_kernel, dim3(nb(n)), dim3(256), 0, s, n, seed,
                     draws.p, offs.p, keys_in.p);
  SPK_HIP_TRY(hipGetLastError());
  draws.release();
  offs.release();
  int bits = 1;  // bits of the largest pair value, below n*n
  while (bits < 62 && ((uint64_t)1 << bits) < (uint64_t)n * (uint64_t)n) ++bits;
  const int end_bit = bits + 1;  // + the direction bit
  tb = 0;
  SPK_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys_in.p, keys.p,
                                                (int)K, 0, end_bit, s));
  tmp.reserve(tb);
  SPK_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys_in.p, keys.p,
                                                (int)K, 0, end_bit, s));
  keys_in.release();

  DevBuf<int32_t> head((size_t)K), uidx((size_t)K);
  hipLaunchKernelGGL(pl_heads_kernel, dim3(nb(K)), dim3(256), 0, s, K, keys.p,
                     head.p);
  tb = 0;
  SPK_HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tb, head.p, uidx.p, (int)K, s));
  tmp.reserve(tb);
  SPK_HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, head.p, uidx.p, (int)K, s));
  head.release();
  int32_t U32 = 0;
  SPK_HIP_TRY(hipMemcpyAsync(&U32, uidx.p + K - 1, 4, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  const int64_t U = U32;
  DevBuf<uint64_t> pairs((size_t)U);
  DevBuf<uint8_t> has0((size_t)U), has1((size_t)U);
  SPK_HIP_TRY(hipMemsetAsync(has1.p, 0, (size_t)U, s));
  hipLaunchKernelGGL(pl_compact_kernel, dim3(nb(K)), dim3(256), 0, s, K, keys.p,
                     uidx.p, pairs.p, has0.p, has1.p);
  SPK_HIP_TRY(hipGetLastError());
  keys.release();
  uidx.release();

  auto* A = new spk_csr_s();
  try {
    A->allocate(n, n, U);
    int32_t un32 = (int32_t)U;
    SPK_HIP_TRY(hipMemcpyAsync(A->rp.p + n, &un32, 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(pl_fill_kernel, dim3(nb(U)), dim3(256), 0, s, U, n, seed,
                       pairs.p, has0.p, has1.p, A->rp.p, A->ci.p, A->v64.p);
    hipLaunchKernelGGL(pl_diag_kernel, dim3(nb(n)), dim3(256), 0, s, n, A->rp.p,
                       A->ci.p, A->v64.p);
    SPK_HIP_TRY(hipGetLastError());
    finish_store(*A, false);
  } catch (...) {
    delete A;
    throw;
  }
  return A;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_gen_powerlaw_spd(int64_t n, uint64_t seed,
                                           spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out, SPK_INVALID_ARGUMENT, "null output");
    *out = build_powerlaw(n, seed);
  });
}
