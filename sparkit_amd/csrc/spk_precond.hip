// spk_precond.hip -- Jacobi, ILU(0)/MILU(0) and SpMV preconditioners.
#include "spk_precond.hpp"

#include <algorithm>

#include "spk_vec.hpp"

namespace spk {
namespace {

// jacobi(A): d = extract_diagonal(A); zero -> error; d = 1/d (jacobi.hpp:27-35)
__global__ void invert_diag_kernel(const int32_t* rp, const int32_t* ci,
                                   const double* v, int64_t len, double* d,
                                   int* zero) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  int lo = rp[i], hi = rp[i + 1], e = hi;
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    if (ci[mid] < i) lo = mid + 1; else hi = mid;
  }
  double di = (lo < e && ci[lo] == i) ? v[lo] : 0.0;
  if (di == 0.0) {
    atomicOr(zero, 1);
    d[i] = 0.0;
  } else {
    d[i] = 1.0 / di;
  }
}

__global__ void to_f32_kernel(const double* in, float* out, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// One IKJ row of ilu0_impl (ilu.hpp:72-115) per thread, for the rows of one
// dependency level.  Multiply-adds are not contracted so that every update
// rounds exactly as the reference's w[pos] -= l_ik * w[q].
__global__ void ilu_factor_level_kernel(const int32_t* __restrict__ rows,
                                        int count,
                                        const int32_t* __restrict__ rp,
                                        const int32_t* __restrict__ ci,
                                        const int32_t* __restrict__ diag_pos,
                                        double* __restrict__ w, int modified) {
#pragma clang fp contract(off)
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int pe = rp[i + 1];
  double dropped = 0.0;
  for (int p = rp[i]; p < pe; ++p) {
    const int k = ci[p];
    if (k >= i) break;
    w[p] = w[p] / w[diag_pos[k]];
    const double l_ik = w[p];
    int pi = p + 1;  // row i positions with column > k
    const int qe = rp[k + 1];
    for (int q = diag_pos[k] + 1; q < qe; ++q) {
      const int j = ci[q];
      while (pi < pe && ci[pi] < j) ++pi;
      if (pi < pe && ci[pi] == j) {
        w[pi] = w[pi] - l_ik * w[q];
      } else if (modified) {
        dropped = dropped + l_ik * w[q];
      }
    }
  }
  if (modified) w[diag_pos[i]] = w[diag_pos[i]] - dropped;
}

// Split the factored values into L (unit diagonal last) and U (diagonal
// first): ilu.hpp:123-141.
__global__ void ilu_split_kernel(int64_t n, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 const double* __restrict__ w,
                                 const int32_t* __restrict__ lrp,
                                 double* __restrict__ lv,
                                 const int32_t* __restrict__ urp,
                                 double* __restrict__ uv) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lk = lrp[i], uk = urp[i];
  for (int p = rp[i]; p < rp[i + 1]; ++p) {
    if (ci[p] < i) lv[lk++] = w[p];
    else uv[uk++] = w[p];
  }
  lv[lk] = 1.0;
}

// forward_solve for the rows of one level (triangular_solve.hpp:42-53)
template <typename T>
__global__ void lower_level_kernel(const int32_t* __restrict__ rows, int count,
                                   const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci,
                                   const T* __restrict__ v, const T* b, T* x) {
#pragma clang fp contract(off)
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int pe = rp[i + 1] - 1;
  T s = b[i];
  for (int p = rp[i]; p < pe; ++p) s = s - v[p] * x[ci[p]];
  x[i] = s / v[pe];
}

// backward_solve for the rows of one level (triangular_solve.hpp:82-96)
template <typename T>
__global__ void upper_level_kernel(const int32_t* __restrict__ rows, int count,
                                   const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci,
                                   const T* __restrict__ v, const T* b, T* x) {
#pragma clang fp contract(off)
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int p0 = rp[i];
  T s = b[i];
  for (int p = p0 + 1; p < rp[i + 1]; ++p) s = s - v[p] * x[ci[p]];
  x[i] = s / v[p0];
}

// group rows by level (counting sort); returns level pointer array
std::vector<int32_t> group_levels(const std::vector<int32_t>& lvl,
                                  std::vector<int32_t>& rows_out) {
  int32_t nl = 0;
  for (int32_t l : lvl) nl = std::max(nl, l + 1);
  std::vector<int32_t> ptr((size_t)nl + 1, 0);
  for (int32_t l : lvl) ptr[l + 1]++;
  for (int32_t l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  rows_out.resize(lvl.size());
  for (size_t i = 0; i < lvl.size(); ++i) rows_out[fill[lvl[i]]++] = (int32_t)i;
  return ptr;
}

spk_csr_s* make_store(int64_t n, std::vector<int32_t>&& rp,
                      const std::vector<int32_t>& ci) {
  auto* S = new spk_csr_s();
  S->allocate(n, n, rp[n]);
  hipStream_t s = stream();
  SPK_HIP_TRY(hipMemcpyAsync(S->rp.p, rp.data(), rp.size() * 4,
                             hipMemcpyHostToDevice, s));
  SPK_HIP_TRY(hipMemcpyAsync(S->ci.p, ci.data(), ci.size() * 4,
                             hipMemcpyHostToDevice, s));
  SPK_HIP_TRY(hipStreamSynchronize(s));
  S->h_rp = std::move(rp);
  upload_plan(*S, build_spmv_plan(S->h_rp.data(), n, &S->nlong));
  return S;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace spk

using namespace spk;

spk_precond_s::~spk_precond_s() {
  delete L;
  delete U;
}

const float* spk_precond_s::inv_diag32() {
  if (!has32) {
    inv_d32.alloc((size_t)n);
    if (n > 0)
      hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(n)), dim3(256), 0,
                         stream(), inv_d64.p, inv_d32.p, n);
    SPK_HIP_TRY(hipGetLastError());
    has32 = true;
  }
  return inv_d32.p;
}

template <>
double* spk_precond_s::tmp<double>() {
  tmp64.reserve((size_t)n);
  return tmp64.p;
}
template <>
float* spk_precond_s::tmp<float>() {
  tmp32.reserve((size_t)n);
  return tmp32.p;
}

template <typename T>
void spk_precond_s::forward(const T* b, T* x) {
  hipStream_t s = stream();
  CsrView<T> v = L->view<T>();
  for (size_t l = 0; l + 1 < lvl_ptr_l.size(); ++l) {
    int cnt = lvl_ptr_l[l + 1] - lvl_ptr_l[l];
    hipLaunchKernelGGL((lower_level_kernel<T>), dim3(blocks_for(cnt)),
                       dim3(256), 0, s, lvl_rows_l.p + lvl_ptr_l[l], cnt, v.rp,
                       v.ci, v.v, b, x);
  }
  SPK_HIP_TRY(hipGetLastError());
}

template <typename T>
void spk_precond_s::backward(const T* b, T* x) {
  hipStream_t s = stream();
  CsrView<T> v = U->view<T>();
  for (size_t l = 0; l + 1 < lvl_ptr_u.size(); ++l) {
    int cnt = lvl_ptr_u[l + 1] - lvl_ptr_u[l];
    hipLaunchKernelGGL((upper_level_kernel<T>), dim3(blocks_for(cnt)),
                       dim3(256), 0, s, lvl_rows_u.p + lvl_ptr_u[l], cnt, v.rp,
                       v.ci, v.v, b, x);
  }
  SPK_HIP_TRY(hipGetLastError());
}

template <typename T>
void spk_precond_s::apply(const T* in, T* out) {
  switch (kind) {
    case kIdentity:
      if (in != out)
        SPK_HIP_TRY(hipMemcpyAsync(out, in, (size_t)n * sizeof(T),
                                   hipMemcpyDeviceToDevice, stream()));
      break;
    case kJacobi: {
      const T* d = inv_diag<T>();
      launch_map(n, [=] __device__(int64_t i) { out[i] = d[i] * in[i]; });
      break;
    }
    case kIlu: {
      // ilu_apply: y = forward_solve(L, r); z = backward_solve(U, y)
      T* y = tmp<T>();
      forward(in, y);
      backward((const T*)y, out);
      break;
    }
    case kSpmv:
      launch_spmv(*M, in, EpiStore<T>{out}, stream());
      break;
  }
}

template void spk_precond_s::apply<double>(const double*, double*);
template void spk_precond_s::apply<float>(const float*, float*);
template void spk_precond_s::forward<double>(const double*, double*);
template void spk_precond_s::forward<float>(const float*, float*);
template void spk_precond_s::backward<double>(const double*, double*);
template void spk_precond_s::backward<float>(const float*, float*);

extern "C" {

spk_status spk_precond_create_jacobi(spk_csr A, spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(A && out, SPK_INVALID_ARGUMENT, "null argument");
    auto* M = new spk_precond_s();
    try {
      M->kind = spk_precond_s::kJacobi;
      M->n = std::min(A->nrows, A->ncols);
      M->inv_d64.alloc((size_t)M->n);
      DevBuf<int> zero(1);
      hipStream_t s = stream();
      SPK_HIP_TRY(hipMemsetAsync(zero.p, 0, sizeof(int), s));
      if (M->n > 0)
        hipLaunchKernelGGL(invert_diag_kernel, dim3(blocks_for(M->n)),
                           dim3(256), 0, s, A->rp.p, A->ci.p, A->v64.p, M->n,
                           M->inv_d64.p, zero.p);
      int hz = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&hz, zero.p, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      SPK_REQUIRE(!hz, SPK_INVALID_ARGUMENT, "jacobi: zero diagonal entry");
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
  });
}

spk_status spk_precond_create_identity(int64_t n, spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(out && n >= 0, SPK_INVALID_ARGUMENT, "bad argument");
    auto* M = new spk_precond_s();
    M->kind = spk_precond_s::kIdentity;
    M->n = n;
    *out = M;
  });
}

spk_status spk_precond_create_spmv(spk_csr Mst, spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(Mst && out, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(Mst->nrows == Mst->ncols, SPK_INVALID_ARGUMENT,
                "approximate inverse must be square");
    auto* M = new spk_precond_s();
    M->kind = spk_precond_s::kSpmv;
    M->n = Mst->nrows;
    M->M = Mst;
    *out = M;
  });
}

spk_status spk_precond_create_ilu0(spk_csr A, int modified, spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(A && out, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "ilu0 requires a square matrix");
    const int64_t n = A->nrows;
    hipStream_t s = stream();
    std::vector<int32_t> ci((size_t)A->nnz);
    if (A->nnz)
      SPK_HIP_TRY(hipMemcpyAsync(ci.data(), A->ci.p, A->nnz * 4,
                                 hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    const std::vector<int32_t>& rp = A->h_rp;
    // diagonal positions (ilu.hpp:54-66) and factor patterns
    std::vector<int32_t> diag_pos((size_t)n);
    std::vector<int32_t> lrp((size_t)n + 1), urp((size_t)n + 1);
    lrp[0] = urp[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
      int32_t dp = -1, nl = 0;
      for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
        if (ci[p] == i && dp < 0) dp = p;
        if (ci[p] < i) ++nl;
      }
      SPK_REQUIRE(dp >= 0, SPK_INVALID_ARGUMENT, "ilu0: missing diagonal entry");
      diag_pos[i] = dp;
      lrp[i + 1] = lrp[i] + nl + 1;
      urp[i + 1] = urp[i] + (rp[i + 1] - rp[i] - nl);
    }
    std::vector<int32_t> lci((size_t)lrp[n]), uci((size_t)urp[n]);
    std::vector<int32_t> lvl_l((size_t)n), lvl_u((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      int32_t lk = lrp[i], uk = urp[i], lv = 0;
      for (int32_t p = rp[i]; p < rp[i + 1]; ++p) {
        if (ci[p] < i) {
          lci[lk++] = ci[p];
          lv = std::max(lv, lvl_l[ci[p]] + 1);
        } else {
          uci[uk++] = ci[p];
        }
      }
      lci[lk] = (int32_t)i;
      lvl_l[i] = lv;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
      int32_t lv = 0;
      for (int32_t p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] > i) lv = std::max(lv, lvl_u[ci[p]] + 1);
      lvl_u[i] = lv;
    }
    auto* M = new spk_precond_s();
    try {
      M->kind = spk_precond_s::kIlu;
      M->n = n;
      std::vector<int32_t> rows_l, rows_u;
      M->lvl_ptr_l = group_levels(lvl_l, rows_l);
      M->lvl_ptr_u = group_levels(lvl_u, rows_u);
      M->lvl_rows_l.alloc((size_t)n);
      M->lvl_rows_u.alloc((size_t)n);
      SPK_HIP_TRY(hipMemcpyAsync(M->lvl_rows_l.p, rows_l.data(), n * 4,
                                 hipMemcpyHostToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(M->lvl_rows_u.p, rows_u.data(), n * 4,
                                 hipMemcpyHostToDevice, s));
      M->L = make_store(n, std::move(lrp), lci);
      M->U = make_store(n, std::move(urp), uci);
      // numeric factorization on a copy of A's values
      DevBuf<double> w((size_t)A->nnz);
      DevBuf<int32_t> dpos((size_t)n);
      SPK_HIP_TRY(hipMemcpyAsync(w.p, A->v64.p, A->nnz * 8,
                                 hipMemcpyDeviceToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(dpos.p, diag_pos.data(), n * 4,
                                 hipMemcpyHostToDevice, s));
      for (size_t l = 0; l + 1 < M->lvl_ptr_l.size(); ++l) {
        int cnt = M->lvl_ptr_l[l + 1] - M->lvl_ptr_l[l];
        hipLaunchKernelGGL(ilu_factor_level_kernel, dim3(blocks_for(cnt)),
                           dim3(256), 0, s, M->lvl_rows_l.p + M->lvl_ptr_l[l],
                           cnt, A->rp.p, A->ci.p, dpos.p, w.p, modified);
      }
      if (n > 0)
        hipLaunchKernelGGL(ilu_split_kernel, dim3(blocks_for(n)), dim3(256), 0,
                           s, n, A->rp.p, A->ci.p, w.p, M->L->rp.p,
                           M->L->v64.p, M->U->rp.p, M->U->v64.p);
      SPK_HIP_TRY(hipGetLastError());
      SPK_HIP_TRY(hipStreamSynchronize(s));
    } catch (...) {
      delete M;
      throw;
    }
    *out = M;
  });
}

spk_status spk_precond_destroy(spk_precond M) {
  return guarded([&] {
    if (!M) return;
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    delete M;
  });
}

spk_status spk_precond_apply(spk_precond M, int dtype, const void* d_in,
                             void* d_out) {
  return guarded([&] {
    SPK_REQUIRE(M && d_in && d_out, SPK_INVALID_ARGUMENT, "null argument");
    if (dtype == SPK_F32) M->apply((const float*)d_in, (float*)d_out);
    else M->apply((const double*)d_in, (double*)d_out);
  });
}

spk_status spk_ilu0_factor_nnz(spk_precond M, int64_t* nnz_l, int64_t* nnz_u) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu, SPK_INVALID_ARGUMENT,
                "not an ILU preconditioner");
    if (nnz_l) *nnz_l = M->L->nnz;
    if (nnz_u) *nnz_u = M->U->nnz;
  });
}

spk_status spk_ilu0_download(spk_precond M, int64_t* l_row_ptr,
                             int64_t* l_col_ind, double* l_vals,
                             int64_t* u_row_ptr, int64_t* u_col_ind,
                             double* u_vals) {
  spk_status st = SPK_OK;
  if (!M || M->kind != spk_precond_s::kIlu) {
    set_error("not an ILU preconditioner");
    return SPK_INVALID_ARGUMENT;
  }
  st = spk_csr_download(M->L, l_row_ptr, l_col_ind, l_vals);
  if (st) return st;
  return spk_csr_download(M->U, u_row_ptr, u_col_ind, u_vals);
}

spk_status spk_ilu0_levels(spk_precond M, int64_t* levels_l,
                           int64_t* levels_u) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu, SPK_INVALID_ARGUMENT,
                "not an ILU preconditioner");
    if (levels_l) *levels_l = (int64_t)M->lvl_ptr_l.size() - 1;
    if (levels_u) *levels_u = (int64_t)M->lvl_ptr_u.size() - 1;
  });
}

spk_status spk_ilu0_forward(spk_precond M, int dtype, const void* d_b,
                            void* d_x) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu && d_b && d_x,
                SPK_INVALID_ARGUMENT, "bad argument");
    if (dtype == SPK_F32) M->forward((const float*)d_b, (float*)d_x);
    else M->forward((const double*)d_b, (double*)d_x);
  });
}

spk_status spk_ilu0_backward(spk_precond M, int dtype, const void* d_b,
                             void* d_x) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu && d_b && d_x,
                SPK_INVALID_ARGUMENT, "bad argument");
    if (dtype == SPK_F32) M->backward((const float*)d_b, (float*)d_x);
    else M->backward((const double*)d_b, (double*)d_x);
  });
}

}  // extern "C"
