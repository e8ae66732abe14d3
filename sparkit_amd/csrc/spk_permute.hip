// spk_permute.hip -- row, column and symmetric permutations of a store on
// the device (permutation.cpp:15-35, permutation.hpp:41-148; perm[old] =
// new): rperm puts old row inv[r] at row r; cperm renames column c to
// perm[c] and re-sorts each row by the new columns; dperm does both
// (P A P^T).  Values travel with their entries, so the result is the
// reference's bit for bit.
//
//   check      perm is a bijection (range check + one hit per target)
//   rows       new row lengths -> prefix sum -> one wavefront copies a row
//   columns    renamed in the copy, then a segmented radix sort (key: new
//              column, payload: value) over the rows
#include <hipcub/hipcub.hpp>

#include "spk_csr.hpp"

namespace spk {
namespace {

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

__global__ void perm_count_kernel(int64_t n, const int64_t* __restrict__ perm,
                                  int32_t* __restrict__ hits, int* err) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t p = perm[k];
  if (p < 0 || p >= n) atomicOr(err, 1);
  else atomicAdd(hits + p, 1);
}

__global__ void perm_check_kernel(int64_t n, const int32_t* __restrict__ hits,
                                  int* err) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n && hits[k] != 1) atomicOr(err, 1);
}

__global__ void perm_inverse_kernel(int64_t n, const int64_t* __restrict__ perm,
                                    int32_t* __restrict__ inv) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) inv[perm[k]] = (int32_t)k;
}

__global__ void perm_len_kernel(int64_t n, const int32_t* __restrict__ rp,
                                const int32_t* __restrict__ inv,
                                int32_t* __restrict__ len) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t o = inv ? inv[r] : r;
  len[r] = rp[o + 1] - rp[o];
}

// one wavefront per new row
__global__ void perm_copy_kernel(int64_t n, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 const double* __restrict__ v,
                                 const int32_t* __restrict__ inv,
                                 const int64_t* __restrict__ cperm,
                                 const int32_t* __restrict__ nrp,
                                 int32_t* __restrict__ nci,
                                 double* __restrict__ nv) {
  const int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  const int64_t o = inv ? inv[r] : r;
  const int a = rp[o], b = rp[o + 1], d = nrp[r] - a;
  for (int k = a + lane; k < b; k += 64) {
    const int c = ci[k];
    nci[k + d] = cperm ? (int32_t)cperm[c] : c;
    nv[k + d] = v[k];
  }
}

spk_csr_s* permute(spk_csr_s& A, const int64_t* perm, int mode) {
  const bool rows = mode & 1, cols = mode & 2;
  SPK_REQUIRE(mode >= 1 && mode <= 3, SPK_INVALID_ARGUMENT, "bad mode");
  SPK_REQUIRE(!(rows && cols) || A.nrows == A.ncols, SPK_INVALID_ARGUMENT,
              "a symmetric permutation needs a square matrix");
  hipStream_t s = stream();
  const int64_t n = A.nrows, np = rows ? A.nrows : A.ncols;
  {  // validity (permutation.cpp:15-26)
    DevBuf<int32_t> hits((size_t)np);
    DevBuf<int> err(1);
    SPK_HIP_TRY(hipMemsetAsync(hits.p, 0, (size_t)np * 4, s));
    SPK_HIP_TRY(hipMemsetAsync(err.p, 0, 4, s));
    hipLaunchKernelGGL(perm_count_kernel, dim3(nb(np)), dim3(256), 0, s, np, perm,
                       hits.p, err.p);
    hipLaunchKernelGGL(perm_check_kernel, dim3(nb(np)), dim3(256), 0, s, np,
                       (const int32_t*)hits.p, err.p);
    SPK_HIP_TRY(hipGetLastError());
    int h = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&h, err.p, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    SPK_REQUIRE(!h, SPK_INVALID_ARGUMENT, "not a permutation");
  }
  DevBuf<int32_t> inv;
  if (rows) {
    inv.alloc((size_t)n);
    hipLaunchKernelGGL(perm_inverse_kernel, dim3(nb(n)), dim3(256), 0, s, n, perm,
                       inv.p);
  }
  DevBuf<int32_t> len((size_t)n);
  hipLaunchKernelGGL(perm_len_kernel, dim3(nb(n)), dim3(256), 0, s, n,
                     (const int32_t*)A.rp.p, rows ? (const int32_t*)inv.p : nullptr,
                     len.p);
  SPK_HIP_TRY(hipGetLastError());
  std::vector<int32_t> hl((size_t)n), hrp((size_t)n + 1);
  SPK_HIP_TRY(hipMemcpyAsync(hl.data(), len.p, n * 4, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  hrp[0] = 0;
  for (int64_t r = 0; r < n; ++r) hrp[(size_t)r + 1] = hrp[(size_t)r] + hl[(size_t)r];
  auto* B = new spk_csr_s();
  try {
    B->allocate(n, A.ncols, A.nnz);
    SPK_HIP_TRY(hipMemcpyAsync(B->rp.p, hrp.data(), (n + 1) * 4,
                               hipMemcpyHostToDevice, s));
    DevBuf<int32_t> tci;
    DevBuf<double> tv;
    int32_t* dci = B->ci.p;
    double* dv = B->v64.p;
    if (cols && A.nnz > 0) {  // copy into scratch, sort into the store
      tci.alloc((size_t)A.nnz);
      tv.alloc((size_t)A.nnz);
      dci = tci.p;
      dv = tv.p;
    }
    if (A.nnz > 0)
      hipLaunchKernelGGL(perm_copy_kernel, dim3(nb(n * 64)), dim3(256), 0, s, n,
                         (const int32_t*)A.rp.p, (const int32_t*)A.ci.p,
                         (const double*)A.v64.p,
                         rows ? (const int32_t*)inv.p : nullptr,
                         cols ? perm : nullptr, (const int32_t*)B->rp.p, dci, dv);
    SPK_HIP_TRY(hipGetLastError());
    if (cols && A.nnz > 0) {
      int bits = 1;
      while (bits < 31 && ((int64_t)1 << bits) < A.ncols) ++bits;
      size_t tb = 0;
      SPK_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(
          nullptr, tb, (const int32_t*)tci.p, B->ci.p, (const double*)tv.p, B->v64.p,
          (int)A.nnz, (int)n, (const int32_t*)B->rp.p, (const int32_t*)B->rp.p + 1,
          0, bits, s));
      DevBuf<unsigned char> tmp(tb);
      SPK_HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(
          tmp.p, tb, (const int32_t*)tci.p, B->ci.p, (const double*)tv.p, B->v64.p,
          (int)A.nnz, (int)n, (const int32_t*)B->rp.p, (const int32_t*)B->rp.p + 1,
          0, bits, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));  // scratch freed at scope end
    }
    finish_store(*B, false);
  } catch (...) {
    delete B;
    throw;
  }
  return B;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_csr_permute(spk_csr A, const int64_t* d_perm, int mode,
                                      spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(A && d_perm && out, SPK_INVALID_ARGUMENT, "null argument");
    *out = permute(*A, d_perm, mode);
  });
}
