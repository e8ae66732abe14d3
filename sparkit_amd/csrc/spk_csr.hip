// spk_csr.hip -- compressed_row store: upload, plan, SpMV entry point,
// diagonal extraction and download.
#include "spk_csr.hpp"

#include <algorithm>
#include <climits>

namespace spk {

std::vector<int32_t> build_spmv_plan(const int32_t* rp, int64_t nrows,
                                     int* nlong) {
  std::vector<int32_t> blocks;
  blocks.reserve((size_t)(nrows / 64 + 2));
  blocks.push_back(0);
  int64_t r = 0;
  int longs = 0;
  while (r < nrows) {
    int64_t len = (int64_t)rp[r + 1] - rp[r];
    if (len > kNnzBlock) {
      ++r;
      ++longs;
      blocks.push_back((int32_t)r);
      continue;
    }
    int64_t start = r, nnz = 0;
    while (r < nrows && r - start < kSpmvBlock) {
      int64_t l = (int64_t)rp[r + 1] - rp[r];
      if (nnz + l > kNnzBlock) break;
      nnz += l;
      ++r;
    }
    blocks.push_back((int32_t)r);
  }
  *nlong = longs;
  return blocks;
}

namespace {

__global__ void check_cols_kernel(const int32_t* ci, int64_t nnz, int64_t ncols,
                                  int* bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    int32_t c = ci[k];
    if (c < 0 || c >= ncols) atomicOr(bad, 1);
  }
}

__global__ void f64_to_f32_kernel(const double* in, float* out, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    out[k] = (float)in[k];
}

// the FP32 value (as FP32 SpMVs read it) rounded to FP16, round to nearest
// even at each step
__global__ void f64_to_f16_kernel(const double* in, _Float16* out, int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    out[k] = (_Float16)(float)in[k];
}

// extract_diagonal: one thread per row, binary search (unary.hpp:29-37)
__global__ void diag_kernel(const int32_t* rp, const int32_t* ci,
                            const double* v, int64_t len, double* d) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  int lo = rp[i], hi = rp[i + 1], e = hi;
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    if (ci[mid] < i) lo = mid + 1; else hi = mid;
  }
  d[i] = (lo < e && ci[lo] == i) ? v[lo] : 0.0;
}

// flags a finite FP64 value whose FP16 rounding is not finite
__global__ void check_f16_kernel(const double* v, const _Float16* h, int64_t n,
                                 int* bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    if (isfinite(v[k]) && !isfinite((float)h[k])) atomicOr(bad, 1);
}

// short column indices: one workgroup per row block writes its entries'
// columns as offsets from the block's first row, flagging any beyond int16
__global__ void short_cols_kernel(const int2* __restrict__ plan,
                                  const int32_t* __restrict__ ci, int16_t* c16,
                                  int* bad) {
  const int2 p0 = plan[blockIdx.x], p1 = plan[blockIdx.x + 1];
  for (int k = p0.y + threadIdx.x; k < p1.y; k += blockDim.x) {
    const int d = ci[k] - p0.x;
    if (d < -32768 || d > 32767) atomicOr(bad, 1);
    else c16[k] = (int16_t)d;
  }
}

int grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 65536);
}

}  // namespace

void finish_store(spk_csr_s& A, bool validate_cols) {
  hipStream_t s = stream();
  A.h_rp.resize((size_t)A.nrows + 1);
  SPK_HIP_TRY(hipMemcpyAsync(A.h_rp.data(), A.rp.p,
                             (A.nrows + 1) * sizeof(int32_t),
                             hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  SPK_REQUIRE(A.h_rp[0] == 0 && A.h_rp[A.nrows] == A.nnz, SPK_INVALID_ARGUMENT,
              "row_ptr must start at 0 and end at nnz");
  for (int64_t r = 0; r < A.nrows; ++r)
    SPK_REQUIRE(A.h_rp[r] <= A.h_rp[r + 1], SPK_INVALID_ARGUMENT,
                "row_ptr must be non-decreasing");
  if (validate_cols && A.nnz > 0) {
    DevBuf<int> bad(1);
    SPK_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(check_cols_kernel, dim3(grid_for(A.nnz)), dim3(256), 0, s,
                       A.ci.p, A.nnz, A.ncols, bad.p);
    int hbad = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    SPK_REQUIRE(!hbad, SPK_INVALID_ARGUMENT, "column index out of range");
  }
  upload_plan(A, build_spmv_plan(A.h_rp.data(), A.nrows, &A.nlong));
}

void upload_plan(spk_csr_s& A, const std::vector<int32_t>& blocks) {
  std::vector<int2> plan(blocks.size());
  for (size_t b = 0; b < blocks.size(); ++b)
    plan[b] = make_int2(blocks[b], A.h_rp[blocks[b]]);
  A.nblocks = (int)blocks.size() - 1;
  A.plan.alloc(plan.size());
  SPK_HIP_TRY(hipMemcpyAsync(A.plan.p, plan.data(), plan.size() * sizeof(int2),
                             hipMemcpyHostToDevice, stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
}

}  // namespace spk

namespace spk {
namespace {

// first position in each row whose column is >= p * width, p = 0..P
__global__ void panel_bounds_kernel(int64_t n, const int32_t* __restrict__ rp,
                                    const int32_t* __restrict__ ci, int P,
                                    int64_t width, int32_t* __restrict__ start) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r0 = rp[i], r1 = rp[i + 1];
  for (int p = 0; p <= P; ++p) {
    const int64_t c0 = (int64_t)p * width;
    int lo = r0, hi = r1;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if ((int64_t)ci[mid] < c0) lo = mid + 1; else hi = mid;
    }
    start[(size_t)p * n + i] = p == P ? r1 : lo;
  }
}

__global__ void panel_len_kernel(int64_t n, const int32_t* __restrict__ start,
                                 int p, int32_t* __restrict__ len) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  len[i] = start[(size_t)(p + 1) * n + i] - start[(size_t)p * n + i];
}

__global__ void panel_copy_kernel(int64_t n, const int32_t* __restrict__ start,
                                  int p, const int32_t* __restrict__ ci,
                                  const double* __restrict__ v,
                                  const int32_t* __restrict__ prp,
                                  int32_t* __restrict__ pci,
                                  double* __restrict__ pv) {
  // one wavefront per row
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int a = start[(size_t)p * n + i], b = start[(size_t)(p + 1) * n + i];
  const int o = prp[i] - a;
  for (int k = a + lane; k < b; k += 64) {
    pci[k + o] = ci[k];
    pv[k + o] = v[k];
  }
}

// Window tiles: one thread per (group, owner thread) walks its S rows'
// cursors window by window; rows longer than lmax are skipped.  COUNT:
// writes each (window, thread) segment's length into off, the owner's
// long-row slots into skip, and flags rows whose columns do not ascend;
// FILL: writes the entries at base[g] + off[window * T + t] (off scanned by
// then).
template <int S, bool FILL>
__global__ void xt_walk_kernel(int64_t nrows, const int32_t* __restrict__ rp,
                               const int32_t* __restrict__ ci,
                               const double* __restrict__ v, int64_t groups,
                               int64_t windows, int window, int lmax,
                               int32_t* __restrict__ off,
                               const int64_t* __restrict__ base,
                               uint32_t* __restrict__ idx, double* __restrict__ xv,
                               uint32_t* __restrict__ skip, int* __restrict__ bad) {
  const int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gt >= groups * kXtThreads) return;
  const int64_t g = gt / kXtThreads;
  const int t = (int)(gt % kXtThreads);
  int cur[S], end[S];
  uint32_t mask = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int64_t r = g * (int64_t)(kXtThreads * S) + t + (int64_t)s * kXtThreads;
    cur[s] = r < nrows ? rp[r] : 0;
    end[s] = r < nrows ? rp[r + 1] : 0;
    if (end[s] - cur[s] > lmax) {
      mask |= 1u << s;
      cur[s] = end[s];
    }
  }
  if (!FILL) skip[gt] = mask;
  int32_t* o = off + g * (windows * kXtThreads + 1);
  for (int64_t w = 0; w < windows; ++w) {
    const int64_t c0 = w * (int64_t)window, limit = c0 + window;
    int64_t p = FILL ? base[g] + o[w * kXtThreads + t] : 0;
    int count = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      while (cur[s] < end[s] && (int64_t)ci[cur[s]] < limit) {
        const int k = cur[s];
        if (FILL) {
          idx[p] = ((uint32_t)s << 16) | (uint32_t)(ci[k] - c0);
          xv[p] = v[k];
          ++p;
        } else {
          ++count;
        }
        ++cur[s];
      }
    }
    if (!FILL) o[w * kXtThreads + t] = count;
  }
  if (!FILL) {
    // ascending columns within each owned row (the walk above relies on it)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int64_t r = g * (int64_t)(kXtThreads * S) + t + (int64_t)s * kXtThreads;
      if (r >= nrows || ((mask >> s) & 1u)) continue;
      for (int k = rp[r] + 1; k < rp[r + 1]; ++k)
        if (ci[k] <= ci[k - 1]) atomicOr(bad, 1);
      if (cur[s] != end[s]) atomicOr(bad, 2);  // a column beyond the last window
    }
  }
}

// In place: each group's segment lengths -> exclusive offsets, the group's
// total in the last slot.  One workgroup per group.
__global__ __launch_bounds__(1024) void xt_scan_kernel(int32_t* __restrict__ off,
                                                       int64_t N) {
  __shared__ int64_t sums[1024];
  int32_t* a = off + blockIdx.x * (N + 1);
  const int64_t chunk = (N + 1023) / 1024;
  const int64_t lo = threadIdx.x * chunk;
  const int64_t hi = lo + chunk < N ? lo + chunk : N;
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += a[i];
  sums[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t t = sums[i];
      sums[i] = run;
      run += t;
    }
    a[N] = (int32_t)run;
  }
  __syncthreads();
  int64_t run = sums[threadIdx.x];
  for (int64_t i = lo; i < hi; ++i) {
    const int32_t t = a[i];
    a[i] = (int32_t)run;
    run += t;
  }
}

__global__ void xt_totals_kernel(const int32_t* __restrict__ off, int64_t N,
                                 int64_t groups, int64_t* __restrict__ tot) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < groups) tot[g] = off[g * (N + 1) + N];
}

// rows[i]'s entries of (rp, ci, v) to position qrp[i] of (qci, qv): one
// wavefront per row
__global__ void rows_copy_kernel(int64_t m, const int32_t* __restrict__ rows,
                                 const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 const double* __restrict__ v,
                                 const int32_t* __restrict__ qrp,
                                 int32_t* __restrict__ qci, double* __restrict__ qv) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= m) return;
  const int a = rp[rows[i]], b = rp[rows[i] + 1];
  const int o = qrp[i] - a;
  for (int k = a + lane; k < b; k += 64) {
    qci[k + o] = ci[k];
    qv[k + o] = v[k];
  }
}

// The long rows' compressed store (rows with more than lmax entries; none:
// null), with its row map and block plan.
spk_csr_s* build_long_rows(spk_csr_s& A, int lmax) {
  std::vector<int32_t> lrow, lrp{0};
  for (int64_t i = 0; i < A.nrows; ++i) {
    const int32_t len = A.h_rp[(size_t)i + 1] - A.h_rp[(size_t)i];
    if (len > lmax) {
      lrow.push_back((int32_t)i);
      lrp.push_back(lrp.back() + len);
    }
  }
  if (lrow.empty()) return nullptr;
  hipStream_t s = stream();
  const int64_t m = (int64_t)lrow.size();
  auto* Q = new spk_csr_s();
  try {
    Q->allocate(m, A.ncols, lrp.back());
    Q->rows.alloc((size_t)m);
    SPK_HIP_TRY(hipMemcpyAsync(Q->rp.p, lrp.data(), (m + 1) * 4,
                               hipMemcpyHostToDevice, s));
    SPK_HIP_TRY(hipMemcpyAsync(Q->rows.p, lrow.data(), m * 4,
                               hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rows_copy_kernel, dim3((unsigned)((m * 64 + 255) / 256)),
                       dim3(256), 0, s, m, (const int32_t*)Q->rows.p,
                       (const int32_t*)A.rp.p, (const int32_t*)A.ci.p,
                       (const double*)A.v64.p, (const int32_t*)Q->rp.p, Q->ci.p,
                       Q->v64.p);
    SPK_HIP_TRY(hipGetLastError());
    SPK_HIP_TRY(hipStreamSynchronize(s));
    Q->h_rp = std::move(lrp);
    upload_plan(*Q, build_spmv_plan(Q->h_rp.data(), m, &Q->nlong));
  } catch (...) {
    delete Q;
    throw;
  }
  return Q;
}

}  // namespace

void build_tiles(spk_csr_s& A, int S, int window) {
  SPK_REQUIRE(S == 8 || S == 16, SPK_INVALID_ARGUMENT,
              "window tiles: 8 or 16 rows per thread");
  SPK_REQUIRE(window >= 256 && window <= kXtMaxWindow && window % 64 == 0,
              SPK_INVALID_ARGUMENT, "window tiles: window in [256, 8192], a multiple of 64");
  A.drop_panels();
  if (A.nrows == 0 || A.nnz == 0) return;
  hipStream_t s = stream();
  const int64_t rows_per_group = (int64_t)kXtThreads * S;
  const int64_t groups = (A.nrows + rows_per_group - 1) / rows_per_group;
  const int64_t windows = (A.ncols + window - 1) / window;
  const int64_t N = windows * kXtThreads;
  SPK_REQUIRE(N < ((int64_t)1 << 31) / 2, SPK_INVALID_ARGUMENT,
              "window tiles: too many windows");
  // long rows would hold up their owner's wavefront in every window
  const int lmax = (int)std::max<int64_t>(256, 8 * windows);
  int64_t long_nnz = 0;
  for (int64_t i = 0; i < A.nrows; ++i) {
    const int32_t len = A.h_rp[(size_t)i + 1] - A.h_rp[(size_t)i];
    if (len > lmax) long_nnz += len;
  }
  DevBuf<int32_t> off((size_t)(groups * (N + 1)));
  DevBuf<uint32_t> skip((size_t)(groups * kXtThreads));
  DevBuf<int> bad(1);
  DevBuf<int64_t> tot((size_t)groups);
  SPK_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  const int64_t threads = groups * kXtThreads;
  const unsigned grid = (unsigned)((threads + 255) / 256);
  auto walk = [&](bool fill, const int64_t* base, uint32_t* idx, double* xv) {
    const int32_t* rp = A.rp.p;
    const int32_t* ci = A.ci.p;
    const double* v = A.v64.p;
    if (S == 16) {
      if (fill)
        hipLaunchKernelGGL((xt_walk_kernel<16, true>), dim3(grid), dim3(256), 0, s,
                           A.nrows, rp, ci, v, groups, windows, window, lmax, off.p,
                           base, idx, xv, skip.p, bad.p);
      else
        hipLaunchKernelGGL((xt_walk_kernel<16, false>), dim3(grid), dim3(256), 0, s,
                           A.nrows, rp, ci, v, groups, windows, window, lmax, off.p,
                           base, idx, xv, skip.p, bad.p);
    } else {
      if (fill)
        hipLaunchKernelGGL((xt_walk_kernel<8, true>), dim3(grid), dim3(256), 0, s,
                           A.nrows, rp, ci, v, groups, windows, window, lmax, off.p,
                           base, idx, xv, skip.p, bad.p);
      else
        hipLaunchKernelGGL((xt_walk_kernel<8, false>), dim3(grid), dim3(256), 0, s,
                           A.nrows, rp, ci, v, groups, windows, window, lmax, off.p,
                           base, idx, xv, skip.p, bad.p);
    }
    SPK_HIP_TRY(hipGetLastError());
  };
  walk(false, nullptr, nullptr, nullptr);
  int hbad = 0;
  SPK_HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  SPK_REQUIRE(hbad == 0, SPK_INVALID_ARGUMENT,
              "window tiles need ascending columns within every row");
  hipLaunchKernelGGL(xt_scan_kernel, dim3((unsigned)groups), dim3(1024), 0, s, off.p, N);
  hipLaunchKernelGGL(xt_totals_kernel, dim3((unsigned)((groups + 255) / 256)),
                     dim3(256), 0, s, (const int32_t*)off.p, N, groups, tot.p);
  SPK_HIP_TRY(hipGetLastError());
  std::vector<int64_t> htot((size_t)groups), hbase((size_t)groups + 1);
  SPK_HIP_TRY(hipMemcpyAsync(htot.data(), tot.p, groups * 8, hipMemcpyDeviceToHost, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  hbase[0] = 0;
  for (int64_t g = 0; g < groups; ++g) hbase[(size_t)g + 1] = hbase[(size_t)g] + htot[(size_t)g];
  SPK_REQUIRE(hbase[(size_t)groups] == A.nnz - long_nnz, SPK_HIP_ERROR,
              "window tiles: entry count mismatch");
  A.xt_long = build_long_rows(A, lmax);
  A.xt_base.alloc((size_t)groups + 1);
  SPK_HIP_TRY(hipMemcpyAsync(A.xt_base.p, hbase.data(), (groups + 1) * 8,
                             hipMemcpyHostToDevice, s));
  const size_t ntile = (size_t)std::max<int64_t>(hbase[(size_t)groups], 1);
  A.xt_idx.alloc(ntile);
  A.xt_v64.alloc(ntile);
  walk(true, A.xt_base.p, A.xt_idx.p, A.xt_v64.p);
  SPK_HIP_TRY(hipStreamSynchronize(s));
  A.xt_off = std::move(off);
  A.xt_skip = std::move(skip);
  A.xt_lmax = lmax;
  A.xt_slots = S;
  A.xt_window = window;
  A.xt_groups = groups;
  A.xt_windows = windows;
}

void build_panels(spk_csr_s& A, int P) {
  A.drop_panels();
  if (P <= 1 || A.nrows == 0) return;
  hipStream_t s = stream();
  const int64_t n = A.nrows;
  const int64_t width = (A.ncols + P - 1) / P;
  DevBuf<int32_t> start((size_t)(P + 1) * (size_t)n);
  DevBuf<int32_t> len((size_t)n);
  hipLaunchKernelGGL(panel_bounds_kernel, dim3(grid_for(n)), dim3(256), 0, s, n,
                     (const int32_t*)A.rp.p, (const int32_t*)A.ci.p, P, width,
                     start.p);
  SPK_HIP_TRY(hipGetLastError());
  DevBuf<int32_t> frp((size_t)n + 1);  // an inner panel's full-row offsets
  std::vector<int32_t> hl((size_t)n), hrp((size_t)n + 1), crp, crow;
  try {
    for (int p = 0; p < P; ++p) {
      hipLaunchKernelGGL(panel_len_kernel, dim3(grid_for(n)), dim3(256), 0, s, n,
                         (const int32_t*)start.p, p, len.p);
      SPK_HIP_TRY(hipMemcpyAsync(hl.data(), len.p, n * 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      int64_t acc = 0;
      hrp[0] = 0;
      for (int64_t i = 0; i < n; ++i) {
        acc += hl[(size_t)i];
        hrp[(size_t)i + 1] = (int32_t)acc;
      }
      // the first and last panels keep every row; the inner ones only their
      // non-empty rows (same ci / v, offsets of the kept rows, a row map)
      const bool inner = p > 0 && p + 1 < P;
      crp.clear();
      crow.clear();
      if (inner) {
        crp.push_back(0);
        for (int64_t i = 0; i < n; ++i)
          if (hl[(size_t)i] > 0) {
            crow.push_back((int32_t)i);
            crp.push_back(hrp[(size_t)i + 1]);
          }
      }
      const int64_t m = inner ? (int64_t)crow.size() : n;
      auto* Q = new spk_csr_s();
      A.panels.push_back(Q);
      Q->allocate(m, A.ncols, acc);
      int32_t* dst_rp = inner ? frp.p : Q->rp.p;
      SPK_HIP_TRY(hipMemcpyAsync(dst_rp, hrp.data(), (n + 1) * 4,
                                 hipMemcpyHostToDevice, s));
      if (acc > 0)
        hipLaunchKernelGGL(panel_copy_kernel, dim3((unsigned)((n * 64 + 255) / 256)),
                           dim3(256), 0, s, n, (const int32_t*)start.p, p,
                           (const int32_t*)A.ci.p, (const double*)A.v64.p,
                           (const int32_t*)dst_rp, Q->ci.p, Q->v64.p);
      SPK_HIP_TRY(hipGetLastError());
      if (inner) {
        SPK_HIP_TRY(hipMemcpyAsync(Q->rp.p, crp.data(), (m + 1) * 4,
                                   hipMemcpyHostToDevice, s));
        if (m > 0) {
          Q->rows.alloc((size_t)m);
          SPK_HIP_TRY(hipMemcpyAsync(Q->rows.p, crow.data(), m * 4,
                                     hipMemcpyHostToDevice, s));
        }
        Q->h_rp = crp;
      } else {
        Q->h_rp = hrp;
      }
      // the host vectors are reused next pass: the copies must be done
      SPK_HIP_TRY(hipStreamSynchronize(s));
      upload_plan(*Q, build_spmv_plan(Q->h_rp.data(), m, &Q->nlong));
    }
  } catch (...) {
    A.drop_panels();
    throw;
  }
}

}  // namespace spk

void spk_csr_s::allocate(int64_t nrows_, int64_t ncols_, int64_t nnz_) {
  nrows = nrows_;
  ncols = ncols_;
  nnz = nnz_;
  rp.alloc((size_t)nrows + 1);
  ci.alloc((size_t)nnz + spk::kPad);
  v64.alloc((size_t)nnz + spk::kPad);
  hipStream_t s = spk::stream();
  SPK_HIP_TRY(hipMemsetAsync(ci.p + nnz, 0, spk::kPad * sizeof(int32_t), s));
  SPK_HIP_TRY(hipMemsetAsync(v64.p + nnz, 0, spk::kPad * sizeof(double), s));
}

const float* spk_csr_s::tiles32() {
  if (!xt_has32) {
    // the tiles hold the entries of the tiled rows only (nnz minus the long
    // rows' entries): convert what xt_v64 holds
    const int64_t m = (int64_t)xt_v64.n;
    xt_v32.alloc((size_t)m);
    if (m > 0)
      hipLaunchKernelGGL(spk::f64_to_f32_kernel, dim3(spk::grid_for(m)),
                         dim3(256), 0, spk::stream(), xt_v64.p, xt_v32.p, m);
    SPK_HIP_TRY(hipGetLastError());
    xt_has32 = true;
  }
  return xt_v32.p;
}

const float* spk_csr_s::vals32() {
  if (!has32) {
    v32.alloc((size_t)nnz + spk::kPad);
    hipStream_t s = spk::stream();
    SPK_HIP_TRY(hipMemsetAsync(v32.p + nnz, 0, spk::kPad * sizeof(float), s));
    if (nnz > 0)
      hipLaunchKernelGGL(spk::f64_to_f32_kernel, dim3(spk::grid_for(nnz)),
                         dim3(256), 0, s, v64.p, v32.p, nnz);
    SPK_HIP_TRY(hipGetLastError());
    has32 = true;
  }
  return v32.p;
}

const _Float16* spk_csr_s::vals16() {
  if (!has16) {
    v16.alloc((size_t)nnz + spk::kPad);
    hipStream_t s = spk::stream();
    SPK_HIP_TRY(hipMemsetAsync(v16.p + nnz, 0, spk::kPad * sizeof(_Float16), s));
    if (nnz > 0)
      hipLaunchKernelGGL(spk::f64_to_f16_kernel, dim3(spk::grid_for(nnz)),
                         dim3(256), 0, s, v64.p, v16.p, nnz);
    SPK_HIP_TRY(hipGetLastError());
    has16 = true;
  }
  return v16.p;
}

using namespace spk;

extern "C" {

spk_status spk_csr_create(int64_t nrows, int64_t ncols, int64_t nnz,
                          const int64_t* row_ptr, const int64_t* col_ind,
                          const double* vals, spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && row_ptr, SPK_INVALID_ARGUMENT, "null argument");
    // Shape(r, c) throws std::logic_error for r <= 1 or c <= 1 (Shape.cpp:8-10)
    SPK_REQUIRE(nrows > 1 && ncols > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
    SPK_REQUIRE(nnz >= 0, SPK_INVALID_ARGUMENT, "negative size");
    SPK_REQUIRE(nrows < INT_MAX && nnz < INT_MAX, SPK_INVALID_ARGUMENT,
                "nrows and nnz must fit in int32 for the device store");
    SPK_REQUIRE(nnz == 0 || (col_ind && vals), SPK_INVALID_ARGUMENT,
                "null col_ind/vals");
    std::vector<int32_t> rp32((size_t)nrows + 1), ci32((size_t)nnz);
    for (int64_t r = 0; r <= nrows; ++r) rp32[r] = (int32_t)row_ptr[r];
    SPK_REQUIRE(row_ptr[nrows] == nnz, SPK_INVALID_ARGUMENT,
                "row_ptr[nrows] != nnz");
    for (int64_t k = 0; k < nnz; ++k) {
      SPK_REQUIRE(col_ind[k] >= 0 && col_ind[k] < ncols, SPK_INVALID_ARGUMENT,
                  "column index out of range");
      ci32[k] = (int32_t)col_ind[k];
    }
    auto* A = new spk_csr_s();
    try {
      A->allocate(nrows, ncols, nnz);
      hipStream_t s = stream();
      SPK_HIP_TRY(hipMemcpyAsync(A->rp.p, rp32.data(), rp32.size() * 4,
                                 hipMemcpyHostToDevice, s));
      if (nnz) {
        SPK_HIP_TRY(hipMemcpyAsync(A->ci.p, ci32.data(), ci32.size() * 4,
                                   hipMemcpyHostToDevice, s));
        SPK_HIP_TRY(hipMemcpyAsync(A->v64.p, vals, (size_t)nnz * 8,
                                   hipMemcpyHostToDevice, s));
      }
      finish_store(*A, false);
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}

spk_status spk_csr_create_device(int64_t nrows, int64_t ncols, int64_t nnz,
                                 const int32_t* d_row_ptr,
                                 const int32_t* d_col_ind,
                                 const double* d_vals, spk_csr* out) {
  return guarded([&] {
    SPK_REQUIRE(out && d_row_ptr, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(nrows > 1 && ncols > 1, SPK_LOGIC_ERROR, "invalid matrix shape");
    SPK_REQUIRE(nnz >= 0 && nrows < INT_MAX && nnz < INT_MAX,
                SPK_INVALID_ARGUMENT, "bad size");
    auto* A = new spk_csr_s();
    try {
      A->allocate(nrows, ncols, nnz);
      hipStream_t s = stream();
      SPK_HIP_TRY(hipMemcpyAsync(A->rp.p, d_row_ptr, (nrows + 1) * 4,
                                 hipMemcpyDeviceToDevice, s));
      if (nnz) {
        SPK_HIP_TRY(hipMemcpyAsync(A->ci.p, d_col_ind, nnz * 4,
                                   hipMemcpyDeviceToDevice, s));
        SPK_HIP_TRY(hipMemcpyAsync(A->v64.p, d_vals, nnz * 8,
                                   hipMemcpyDeviceToDevice, s));
      }
      finish_store(*A, true);
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}

spk_status spk_csr_destroy(spk_csr A) {
  return guarded([&] {
    if (!A) return;
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    delete A;
  });
}

spk_status spk_csr_shape(spk_csr A, int64_t* nrows, int64_t* ncols,
                         int64_t* nnz) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
  });
}

spk_status spk_csr_download(spk_csr A, int64_t* row_ptr, int64_t* col_ind,
                            double* vals) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    hipStream_t s = stream();
    if (row_ptr)
      for (int64_t r = 0; r <= A->nrows; ++r) row_ptr[r] = A->h_rp[r];
    if (col_ind && A->nnz) {
      std::vector<int32_t> ci((size_t)A->nnz);
      SPK_HIP_TRY(hipMemcpyAsync(ci.data(), A->ci.p, A->nnz * 4,
                                 hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      for (int64_t k = 0; k < A->nnz; ++k) col_ind[k] = ci[k];
    }
    if (vals && A->nnz) {
      SPK_HIP_TRY(hipMemcpyAsync(vals, A->v64.p, A->nnz * 8,
                                 hipMemcpyDeviceToHost, s));
    }
    SPK_HIP_TRY(hipStreamSynchronize(s));
  });
}

spk_status spk_csr_device_arrays(spk_csr A, int dtype, const int32_t** rp,
                                 const int32_t** ci, const void** vals) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    if (rp) *rp = A->rp.p;
    if (ci) *ci = A->ci.p;
    if (vals) *vals = dtype == SPK_F32 ? (const void*)A->vals32()
                                       : (const void*)A->v64.p;
  });
}

spk_status spk_csr_plan_info(spk_csr A, int64_t* nblocks,
                             int64_t* nlong_blocks) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    if (nblocks) *nblocks = A->nblocks;
    if (nlong_blocks) *nlong_blocks = A->nlong;
  });
}

spk_status spk_csr_spmv(spk_csr A, int dtype, const void* d_x, void* d_y) {
  return guarded([&] {
    SPK_REQUIRE(A && d_x && d_y, SPK_INVALID_ARGUMENT, "null argument");
    hipStream_t s = stream();
    if (dtype == SPK_F32)
      launch_spmv(*A, (const float*)d_x, EpiStore<float>{(float*)d_y}, s);
    else
      launch_spmv(*A, (const double*)d_x, EpiStore<double>{(double*)d_y}, s);
  });
}

spk_status spk_csr_spmv_dot(spk_csr A, int dtype, const void* d_x, void* d_y,
                            double* d_result) {
  return guarded([&] {
    SPK_REQUIRE(A && d_x && d_y && d_result, SPK_INVALID_ARGUMENT,
                "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "<x, A x> needs a square matrix");
    SPK_REQUIRE(dtype == SPK_F32 || dtype == SPK_F64, SPK_INVALID_ARGUMENT,
                "dtype must be SPK_F32 or SPK_F64");
    hipStream_t s = stream();
    if (dtype == SPK_F32)
      launch_spmv(*A, (const float*)d_x,
                  EpiDotX<float>{(const float*)d_x, (float*)d_y, d_result}, s);
    else
      launch_spmv(*A, (const double*)d_x,
                  EpiDotX<double>{(const double*)d_x, (double*)d_y, d_result}, s);
  });
}

spk_status spk_csr_set_panels(spk_csr A, int64_t panels) {
  return guarded([&] {
    SPK_REQUIRE(A && panels >= 1 && panels <= 4096, SPK_INVALID_ARGUMENT,
                "bad argument");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    build_panels(*A, (int)panels);
  });
}

spk_status spk_csr_get_panels(spk_csr A, int64_t* panels) {
  return guarded([&] {
    SPK_REQUIRE(A && panels, SPK_INVALID_ARGUMENT, "null argument");
    if (A->xt_groups)
      *panels = -((int64_t)A->xt_slots * 65536 + A->xt_window);
    else
      *panels = A->panels.empty() ? 1 : (int64_t)A->panels.size();
  });
}

spk_status spk_csr_set_half_values(spk_csr A, int enable) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    if (enable) {
      // the FP16 copy is made here; a value beyond FP16's range becomes
      // infinite there, which would make every FP32 product with it inf
      (void)A->vals16();
      DevBuf<int> bad(1);
      hipStream_t s = stream();
      SPK_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
      if (A->nnz > 0)
        hipLaunchKernelGGL(check_f16_kernel, dim3(grid_for(A->nnz)), dim3(256), 0, s,
                           (const double*)A->v64.p, (const _Float16*)A->v16.p,
                           A->nnz, bad.p);
      SPK_HIP_TRY(hipGetLastError());
      int hbad = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      SPK_REQUIRE(!hbad, SPK_INVALID_ARGUMENT,
                  "a finite value is out of FP16's range");
    }
    A->half32 = enable != 0;
  });
}

spk_status spk_csr_set_short_indices(spk_csr A, int enable) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    if (!enable) {
      A->short_idx = false;
      return;
    }
    if (A->short_idx) return;
    hipStream_t s = stream();
    A->c16.alloc((size_t)A->nnz + kPad);
    SPK_HIP_TRY(hipMemsetAsync(A->c16.p + A->nnz, 0, kPad * sizeof(int16_t), s));
    DevBuf<int> bad(1);
    SPK_HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    if (A->nblocks > 0)
      hipLaunchKernelGGL(short_cols_kernel, dim3(A->nblocks), dim3(256), 0, s,
                         (const int2*)A->plan.p, (const int32_t*)A->ci.p, A->c16.p,
                         bad.p);
    SPK_HIP_TRY(hipGetLastError());
    int hbad = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    if (hbad) A->c16.release();
    SPK_REQUIRE(!hbad, SPK_INVALID_ARGUMENT,
                "a column lies more than 32767 from its row block's first row");
    A->short_idx = true;
  });
}

spk_status spk_csr_set_window_tiles(spk_csr A, int64_t slots, int64_t window) {
  return guarded([&] {
    SPK_REQUIRE(A, SPK_INVALID_ARGUMENT, "null store");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    build_tiles(*A, (int)slots, (int)window);
  });
}

// layout codes of spk_csr_tune_layout / spk_csr_set_layout: P > 0 column
// panels, -(slots * 65536 + window) window tiles
static void apply_layout(spk_csr_s& A, int64_t code) {
  if (code > 0) build_panels(A, (int)code);
  else build_tiles(A, (int)(-code / 65536), (int)(-code % 65536));
}

spk_status spk_csr_set_layout(spk_csr A, int64_t code) {
  return guarded([&] {
    SPK_REQUIRE(A && code != 0, SPK_INVALID_ARGUMENT, "bad argument");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    apply_layout(*A, code);
  });
}

spk_status spk_csr_tune_layout(spk_csr A, int64_t* chosen, int64_t* tried,
                               float* tried_us, int64_t* ntried) {
  return guarded([&] {
    TraceRange trace_("spk_csr_tune_layout");
    SPK_REQUIRE(A && chosen, SPK_INVALID_ARGUMENT, "null argument");
    hipStream_t s = stream();
    DevBuf<double> x((size_t)std::max<int64_t>(A->ncols, 1));
    DevBuf<double> y((size_t)std::max<int64_t>(A->nrows, 1));
    SPK_HIP_TRY(hipMemsetAsync(x.p, 0, x.n * 8, s));
    std::vector<int64_t> cands{1};
    for (int P : {4, 8, 12, 16, 24, 32, 64})
      if (A->ncols >= (int64_t)P * 32768) cands.push_back(P);
    // window tiles stage every x window of a row group in LDS, so one call
    // reads groups * ncols x values where a gather reads about nnz: they are
    // offered only where the windows are at least that dense (a power-law
    // matrix of 4.2M rows and 38 entries a row reads 6x more x through them
    // and runs 3x slower than 16 column panels)
    const int64_t groups16 = (A->nrows + 16 * kXtThreads - 1) / (16 * kXtThreads);
    if (A->ncols >= 16 * 8192 && A->nrows >= 2 * kXtThreads * 16 &&
        groups16 * A->ncols <= A->nnz)
      for (int64_t code : {16 * 65536 + 8192, 8 * 65536 + 8192, 16 * 65536 + 4096})
        cands.push_back(-code);
    hipEvent_t e0, e1;
    SPK_HIP_TRY(hipEventCreate(&e0));
    SPK_HIP_TRY(hipEventCreate(&e1));
    int64_t best = 1;
    float best_us = 1e30f;
    int64_t k = 0;
    try {
      for (int64_t P : cands) {
        apply_layout(*A, P);
        for (int w = 0; w < 2; ++w) launch_spmv(*A, (const double*)x.p, EpiStore<double>{y.p}, s);
        SPK_HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < 5; ++r)
          launch_spmv(*A, (const double*)x.p, EpiStore<double>{y.p}, s);
        SPK_HIP_TRY(hipEventRecord(e1, s));
        SPK_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        SPK_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const float us = ms * 1000.f / 5.f;
        if (tried) tried[k] = P;
        if (tried_us) tried_us[k] = us;
        ++k;
        if (us < best_us) {
          best_us = us;
          best = P;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    apply_layout(*A, best);
    if (ntried) *ntried = k;
    *chosen = best;
  });
}

spk_status spk_csr_extract_diagonal(spk_csr A, double* d_diag) {
  return guarded([&] {
    SPK_REQUIRE(A && d_diag, SPK_INVALID_ARGUMENT, "null argument");
    int64_t len = std::min(A->nrows, A->ncols);
    if (len == 0) return;
    hipLaunchKernelGGL(diag_kernel, dim3((unsigned)((len + 255) / 256)),
                       dim3(256), 0, stream(), A->rp.p, A->ci.p, A->v64.p, len,
                       d_diag);
    SPK_HIP_TRY(hipGetLastError());
  });
}

}  // extern "C"
