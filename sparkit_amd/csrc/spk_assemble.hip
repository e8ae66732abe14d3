// spk_assemble.hip -- CSR assembly on the device from (row, col[, value])
// lists: the pattern build of Compressed_row_sparsity
// (Compressed_row_sparsity_Impl.cpp:10-40: sort the (row, col) pairs,
// collapse duplicates, per-row counts, prefix sum) and make_matrix's value
// placement (matgen.hpp:25-52: on duplicate positions the last entry in
// input order wins).
//
//   keys  key_k = row_k * ncols + col_k (64 bit), payload k
//   sort  stable radix sort of (key, k): duplicates stay in input order
//   runs  the last element of each run of equal keys is the surviving
//         entry (last wins); its rank among run ends is its position
//   rows  rp[r] = first position whose key >= r * ncols (binary search)
// Integer work only: the pattern is the reference's bit for bit.
#include <hipcub/hipcub.hpp>

#include "spk_csr.hpp"

namespace spk {
namespace {

__global__ void asm_keys_kernel(int64_t m, int64_t nrows, int64_t ncols,
                                const int64_t* __restrict__ rows,
                                const int64_t* __restrict__ cols,
                                uint64_t* __restrict__ keys,
                                int32_t* __restrict__ idx, int* err) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int64_t r = rows[k], c = cols[k];
  if (r < 0 || c < 0) atomicOr(err, 1);               // Index: logic_error
  else if (r >= nrows || c >= ncols) atomicOr(err, 2);  // outside the shape
  const bool ok = r >= 0 && c >= 0 && r < nrows && c < ncols;
  keys[k] = ok ? (uint64_t)r * (uint64_t)ncols + (uint64_t)c : 0;
  idx[k] = (int32_t)k;
}

__global__ void asm_tail_kernel(int64_t m, const uint64_t* __restrict__ keys,
                                int32_t* __restrict__ tail) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m) return;
  tail[k] = (k == m - 1 || keys[k + 1] != keys[k]) ? 1 : 0;
}

__global__ void asm_scatter_kernel(int64_t m, int64_t ncols,
                                   const uint64_t* __restrict__ keys,
                                   const int32_t* __restrict__ idx,
                                   const int32_t* __restrict__ tail,
                                   const int32_t* __restrict__ pos,
                                   const double* __restrict__ vals,
                                   uint64_t* __restrict__ ukeys,
                                   int32_t* __restrict__ ci,
                                   double* __restrict__ v) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m || !tail[k]) return;
  const int32_t q = pos[k];  // exclusive rank among run ends
  ukeys[q] = keys[k];
  ci[q] = (int32_t)(keys[k] % (uint64_t)ncols);
  v[q] = vals ? vals[idx[k]] : 0.0;
}

__global__ void asm_rowptr_kernel(int64_t nrows, int64_t ncols, int64_t u,
                                  const uint64_t* __restrict__ ukeys,
                                  int32_t* __restrict__ rp) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > nrows) return;
  const uint64_t first = (uint64_t)r * (uint64_t)ncols;
  int64_t lo = 0, hi = u;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (ukeys[mid] < first) lo = mid + 1; else hi = mid;
  }
  rp[r] = (int32_t)(r == nrows ? u : lo);
}

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

spk_csr_s* assemble(int64_t nrows, int64_t ncols, int64_t m, const int64_t* rows,
                    const int64_t* cols, const double* vals) {
  SPK_REQUIRE(nrows > 1 && ncols > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
  SPK_REQUIRE(nrows < INT32_MAX && ncols < INT32_MAX && m >= 0 && m < INT32_MAX,
              SPK_INVALID_ARGUMENT, "exceeds int32 indexing");
  SPK_REQUIRE(m == 0 || (rows && cols), SPK_INVALID_ARGUMENT, "null index list");
  hipStream_t s = stream();
  const int64_t mm = m > 0 ? m : 1;
  DevBuf<uint64_t> k_in((size_t)mm), k_out((size_t)mm);
  DevBuf<int32_t> i_in((size_t)mm), i_out((size_t)mm), tail((size_t)mm),
      pos((size_t)mm);
  DevBuf<int> err(1);
  SPK_HIP_TRY(hipMemsetAsync(err.p, 0, 4, s));
  if (m > 0)
    hipLaunchKernelGGL(asm_keys_kernel, dim3(nb(m)), dim3(256), 0, s, m, nrows,
                       ncols, rows, cols, k_in.p, i_in.p, err.p);
  SPK_HIP_TRY(hipGetLastError());
  int herr = 0;
  SPK_HIP_TRY(hipMemcpyAsync(&herr, err.p, 4, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  SPK_REQUIRE(!(herr & 1), SPK_LOGIC_ERROR, "negative index");
  SPK_REQUIRE(!(herr & 2), SPK_INVALID_ARGUMENT, "index outside the matrix shape");
  int64_t U = 0;
  if (m > 0) {
    int bits = 1;
    while (bits < 63 && ((uint64_t)1 << bits) < (uint64_t)nrows * (uint64_t)ncols)
      ++bits;
    size_t tb = 0;
    SPK_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k_in.p, k_out.p,
                                                   i_in.p, i_out.p, (int)m, 0,
                                                   bits, s));
    DevBuf<unsigned char> tmp(tb);
    SPK_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k_in.p, k_out.p,
                                                   i_in.p, i_out.p, (int)m, 0,
                                                   bits, s));
    hipLaunchKernelGGL(asm_tail_kernel, dim3(nb(m)), dim3(256), 0, s, m, k_out.p,
                       tail.p);
    tb = 0;
    SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, tail.p, pos.p, (int)m, s));
    tmp.reserve(tb);
    SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, tail.p, pos.p, (int)m, s));
    int32_t lp = 0, lt = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&lp, pos.p + m - 1, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipMemcpyAsync(&lt, tail.p + m - 1, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    U = (int64_t)lp + lt;
  }
  auto* A = new spk_csr_s();
  try {
    A->allocate(nrows, ncols, U);
    DevBuf<uint64_t> ukeys((size_t)(U > 0 ? U : 1));
    if (m > 0)
      hipLaunchKernelGGL(asm_scatter_kernel, dim3(nb(m)), dim3(256), 0, s, m, ncols,
                         (const uint64_t*)k_out.p, (const int32_t*)i_out.p,
                         (const int32_t*)tail.p, (const int32_t*)pos.p, vals,
                         ukeys.p, A->ci.p, A->v64.p);
    hipLaunchKernelGGL(asm_rowptr_kernel, dim3(nb(nrows + 1)), dim3(256), 0, s,
                       nrows, ncols, U, (const uint64_t*)ukeys.p, A->rp.p);
    SPK_HIP_TRY(hipGetLastError());
    SPK_HIP_TRY(hipStreamSynchronize(s));
    finish_store(*A, false);
  } catch (...) {
    delete A;
    throw;
  }
  return A;
}

// (row, col) of every entry as int64 lists: the transpose is the assembly
// of the swapped pairs
__global__ void expand_kernel(int64_t n, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ ci,
                              int64_t* __restrict__ rows,
                              int64_t* __restrict__ cols) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  for (int k = rp[i] + lane; k < rp[i + 1]; k += 64) {
    rows[k] = i;
    cols[k] = ci[k];
  }
}

spk_csr_s* transpose(spk_csr_s& A) {
  hipStream_t s = stream();
  const int64_t m = A.nnz > 0 ? A.nnz : 1;
  DevBuf<int64_t> r((size_t)m), c((size_t)m);
  if (A.nnz > 0)
    hipLaunchKernelGGL(expand_kernel, dim3(nb(A.nrows * 64)), dim3(256), 0, s,
                       A.nrows, (const int32_t*)A.rp.p, (const int32_t*)A.ci.p,
                       r.p, c.p);
  SPK_HIP_TRY(hipGetLastError());
  return assemble(A.ncols, A.nrows, A.nnz, c.p, r.p, A.v64.p);
}

// (M + M^T) / 2 on M's own pattern: one thread per row finds the mirror
// entry of each of its entries by binary search in the mirror row
__global__ void sym_part_kernel(int64_t n, const int32_t* __restrict__ rp,
                                const int32_t* __restrict__ ci,
                                const double* __restrict__ v,
                                double* __restrict__ out, int* err) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int j = ci[k];
    int lo = rp[j], hi = rp[j + 1];
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (ci[mid] < (int)i) lo = mid + 1; else hi = mid;
    }
    if (lo < rp[j + 1] && ci[lo] == (int)i) {
      out[k] = 0.5 * (v[k] + v[lo]);
    } else {
      out[k] = 0.0;
      atomicOr(err, 1);
    }
  }
}

spk_csr_s* symmetric_part(spk_csr_s& M) {
  SPK_REQUIRE(M.nrows == M.ncols, SPK_INVALID_ARGUMENT,
              "the symmetric part needs a square matrix");
  hipStream_t s = stream();
  auto* S = new spk_csr_s();
  try {
    S->allocate(M.nrows, M.ncols, M.nnz);
    SPK_HIP_TRY(hipMemcpyAsync(S->rp.p, M.rp.p, (M.nrows + 1) * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, s));
    if (M.nnz > 0)
      SPK_HIP_TRY(hipMemcpyAsync(S->ci.p, M.ci.p, M.nnz * sizeof(int32_t),
                                 hipMemcpyDeviceToDevice, s));
    DevBuf<int> err(1);
    SPK_HIP_TRY(hipMemsetAsync(err.p, 0, 4, s));
    if (M.nrows > 0)
      hipLaunchKernelGGL(sym_part_kernel, dim3(nb(M.nrows)), dim3(256), 0, s,
                         M.nrows, (const int32_t*)M.rp.p, (const int32_t*)M.ci.p,
                         (const double*)M.v64.p, S->v64.p, err.p);
    SPK_HIP_TRY(hipGetLastError());
    int h = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&h, err.p, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    SPK_REQUIRE(!h, SPK_INVALID_ARGUMENT,
                "pattern is not structurally symmetric");
    finish_store(*S, false);
  } catch (...) {
    delete S;
    throw;
  }
  return S;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_csr_symmetric_part(spk_csr M, spk_csr* out) {
  return guarded([&] {
    TraceRange trace_("spk_csr_symmetric_part");
    SPK_REQUIRE(M && out, SPK_INVALID_ARGUMENT, "null argument");
    *out = symmetric_part(*M);
  });
}

extern "C" spk_status spk_csr_transpose(spk_csr A, spk_csr* out) {
  return guarded([&] {
    TraceRange trace_("spk_csr_transpose");
    SPK_REQUIRE(A && out, SPK_INVALID_ARGUMENT, "null argument");
    *out = transpose(*A);
  });
}

extern "C" spk_status spk_csr_assemble(int64_t nrows, int64_t ncols, int64_t count,
                                       const int64_t* d_rows, const int64_t* d_cols,
                                       const double* d_vals, spk_csr* out) {
  return guarded([&] {
    TraceRange trace_("spk_csr_assemble");
    SPK_REQUIRE(out, SPK_INVALID_ARGUMENT, "null output");
    *out = assemble(nrows, ncols, count, d_rows, d_cols, d_vals);
  });
}
