// spk_precond.hip -- Jacobi, ILU(0)/MILU(0) and SpMV preconditioners.
#include "spk_precond.hpp"

#include <algorithm>
#include <type_traits>

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

// Triangular sweeps in ONE launch ("sync-free").  Rows are laid out in
// dependency-level order, one thread per row, and every level is padded to a
// multiple of the wavefront size, so the rows of one wavefront never depend
// on each other: a lane only ever waits for rows of EARLIER wavefronts.
// (Without the padding a wavefront straddling two levels deadlocks: the
// compiler may sink the final store below the wait loop, where it runs only
// after every lane of the wavefront has left the loop.)  Every value of the
// output doubles as its own ready flag: the output starts filled with a NaN
// sentinel that arithmetic never produces, a finished row stores its value
// with one agent-scope atomic store, and a row waiting on a dependency spins
// on an agent-scope atomic load of it.  Workgroups are dispatched in order,
// so the earliest unfinished wavefront always has its dependencies resident
// or finished.  A bounded spin sets an error flag and ends the row, so every
// wavefront exits whatever happens.
template <typename T>
struct Sentinel;
template <>
struct Sentinel<double> {
  typedef unsigned long long U;
  static constexpr U bits = 0xFFF8DEADBEEFCAFEull;
};
template <>
struct Sentinel<float> {
  typedef unsigned int U;
  static constexpr U bits = 0xFFC0DEADu;
};

template <typename T>
__device__ __forceinline__ bool load_ready(const T* p, T* out) {
  typedef typename Sentinel<T>::U U;
  U u = __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  T v;
  __builtin_memcpy(&v, &u, sizeof(T));
  *out = v;
  return u != Sentinel<T>::bits;
}

template <typename T>
__device__ __forceinline__ void store_ready(T* p, T v) {
  typedef typename Sentinel<T>::U U;
  U u;
  __builtin_memcpy(&u, &v, sizeof(T));
  // a NaN result carrying the sentinel's payload (propagated from a NaN
  // input) must not read as "not ready": publish it with another payload
  if (u == Sentinel<T>::bits) u ^= (U)1;
  __hip_atomic_store(reinterpret_cast<U*>(p), u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ T sentinel_value() {
  typedef typename Sentinel<T>::U U;
  U u = Sentinel<T>::bits;
  T v;
  __builtin_memcpy(&v, &u, sizeof(T));
  return v;
}

constexpr unsigned kSpinLimit = 1u << 22;
constexpr int kSweepBatch = 4;  // dependencies a sweep row loads together
constexpr int kErrorWords = 64;  // sweep status word + guard words
// sweep schedule thresholds (see spk_precond_s::level_launch)
constexpr int64_t kLevelLaunchRows = 65536;
constexpr int64_t kLevelLaunchMaxLevels = 4096;

// LOWER: forward_solve (triangular_solve.hpp:42-53), the diagonal is the
// row's last entry; otherwise backward_solve (:82-96), the diagonal is the
// first.  x[c] -= v * x[c'] in storage order, multiply-adds not contracted:
// bit-identical to the sequential reference.  Optionally resets b[i] to the
// sentinel after reading it (b is the previous sweep's output) and marks
// mark[i] with the sentinel (the next sweep's output).
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void sweep_kernel(
    const int32_t* __restrict__ rows, int n, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const T* __restrict__ v, T* b, T* x,
    bool reset_b, T* mark, int* error) {
#pragma clang fp contract(off)
  // A small persistent grid walks the sweep order: each thread takes the
  // rows t, t + W, t + 2W, ... (W = grid threads), so only a window of W
  // rows -- a few levels -- waits at any time.  A wavefront always holds 64
  // consecutive positions of one (padded) level, i.e. independent rows, and
  // every row waits only for earlier positions, so the earliest unfinished
  // row can always proceed.
  const int W = gridDim.x * blockDim.x;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += W) {
    const int i = rows[t];
    if (i < 0) continue;  // level padding
    const int p0 = rp[i], p1 = rp[i + 1];
    const int q1 = LOWER ? p1 - 1 : p1;
    const int pd = LOWER ? p1 - 1 : p0;
    T s = b[i];
    if (reset_b) b[i] = sentinel_value<T>();
    if (mark) mark[i] = sentinel_value<T>();
    // The dependencies are polled kSweepBatch at a time: all of a batch's
    // loads are issued together (one round trip instead of one per
    // dependency), and the batch is consumed in storage order once every
    // value in it is ready -- the reference's arithmetic, unchanged.
    int p = LOWER ? p0 : p0 + 1;
    unsigned spins = 0;
    while (p < q1) {
      const int cnt = q1 - p < kSweepBatch ? q1 - p : kSweepBatch;
      int c[kSweepBatch];
      T xv[kSweepBatch];
      bool ok[kSweepBatch];
#pragma unroll
      for (int k = 0; k < kSweepBatch; ++k) {
        ok[k] = k >= cnt;
        c[k] = k < cnt ? ci[p + k] : 0;
      }
      bool all = false;
      while (!all) {
        all = true;
#pragma unroll
        for (int k = 0; k < kSweepBatch; ++k) {
          if (!ok[k]) ok[k] = load_ready(x + c[k], &xv[k]);
          all = all && ok[k];
        }
        if (!all) {
          if (++spins > kSpinLimit) {
            atomicOr(error, LOWER ? 1 : 2);
            break;
          }
          __builtin_amdgcn_s_sleep(2);  // back off: fewer polls in flight
        }
      }
      if (!all) {  // expired wait: publish what there is (reported)
        p = q1;
        break;
      }
#pragma unroll
      for (int k = 0; k < kSweepBatch; ++k)
        if (k < cnt) s = s - v[p + k] * xv[k];
      p += cnt;
    }
    store_ready(x + i, s / v[pd]);
    // keep this row's publication ahead of the next row's waits
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
}

// One dependency level of a triangular sweep: every row of the level reads
// only rows of earlier levels (final, previous launches), so plain loads
// suffice.  Same operation order as sweep_kernel.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void level_sweep_kernel(
    const int32_t* __restrict__ rows, int count, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const T* __restrict__ v, const T* b, T* x) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int p0 = rp[i], p1 = rp[i + 1];
  const int q1 = LOWER ? p1 - 1 : p1;
  const int pd = LOWER ? p1 - 1 : p0;
  T s = b[i];
  for (int p = LOWER ? p0 : p0 + 1; p < q1; ++p) s = s - v[p] * x[ci[p]];
  x[i] = s / v[pd];
}

template <typename T>
__global__ void fill_sentinel_kernel(T* x, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = sentinel_value<T>();
}

// rows grouped by level, each level padded with -1 to a multiple of 64
std::vector<int32_t> pad_levels(const std::vector<int32_t>& rows,
                                const std::vector<int32_t>& ptr) {
  std::vector<int32_t> out;
  out.reserve(rows.size() + 64 * (ptr.size() - 1));
  for (size_t l = 0; l + 1 < ptr.size(); ++l) {
    for (int32_t k = ptr[l]; k < ptr[l + 1]; ++k) out.push_back(rows[k]);
    while (out.size() % 64) out.push_back(-1);
  }
  return out;
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
  delete Ls;
  delete Us;
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

template <typename T>
static T* sentinel_buffer(DevBuf<T>& buf, int64_t n) {
  if (buf.p == nullptr || buf.n < (size_t)n) {
    buf.alloc((size_t)n);
    hipLaunchKernelGGL((fill_sentinel_kernel<T>), dim3(blocks_for(n)),
                       dim3(256), 0, stream(), buf.p, n);
    SPK_HIP_TRY(hipGetLastError());
  }
  return buf.p;
}

template <>
double* spk_precond_s::tmp<double>() {
  return sentinel_buffer(tmp64, n);
}
template <>
float* spk_precond_s::tmp<float>() {
  return sentinel_buffer(tmp32, n);
}

void spk_precond_s::reset_sentinels() {
  if (tmp64.p)
    hipLaunchKernelGGL((fill_sentinel_kernel<double>), dim3(blocks_for(n)),
                       dim3(256), 0, stream(), tmp64.p, n);
  if (tmp32.p)
    hipLaunchKernelGGL((fill_sentinel_kernel<float>), dim3(blocks_for(n)),
                       dim3(256), 0, stream(), tmp32.p, n);
  SPK_HIP_TRY(hipGetLastError());
}

void spk_precond_s::check_sweeps() {
  if (kind != kIlu) return;
  // word 0: 1 = a forward, 2 = a backward sweep's bounded wait expired;
  // words 1..kErrorWords-1 must stay zero (a guard against stray writes)
  int h[kErrorWords];
  SPK_HIP_TRY(hipMemcpyAsync(h, sweep_error.p, sizeof(h), hipMemcpyDeviceToHost,
                             stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
  bool guard_ok = true;
  for (int k = 1; k < kErrorWords; ++k) guard_ok = guard_ok && h[k] == 0;
  if (h[0] == 0 && guard_ok) return;
  SPK_HIP_TRY(hipMemsetAsync(sweep_error.p, 0, sizeof(h), stream()));
  SPK_HIP_TRY(hipStreamSynchronize(stream()));
  SPK_REQUIRE(guard_ok, SPK_HIP_ERROR,
              "ILU sweep status words overwritten (a stray device write)");
  SPK_REQUIRE(false, SPK_HIP_ERROR,
              std::string("ILU triangular sweep: bounded wait expired in the ") +
                  (h[0] & 1 ? "forward" : "") + (h[0] == 3 ? " and " : "") +
                  (h[0] & 2 ? "backward" : "") + " sweep (results invalid)");
}

template <typename T, bool LOWER>
void spk_precond_s::level_sweep(const T* b, T* x) {
  hipStream_t s = stream();
  CsrView<T> v = (LOWER ? L : U)->view<T>();
  const std::vector<int32_t>& ptr = LOWER ? lvl_ptr_l : lvl_ptr_u;
  const int32_t* rows = LOWER ? lvl_rows_l.p : lvl_rows_u.p;
  for (size_t l = 0; l + 1 < ptr.size(); ++l) {
    const int cnt = ptr[l + 1] - ptr[l];
    hipLaunchKernelGGL((level_sweep_kernel<T, LOWER>), dim3(blocks_for(cnt)),
                       dim3(256), 0, s, rows + ptr[l], cnt, v.rp, v.ci, v.v, b, x);
  }
  SPK_HIP_TRY(hipGetLastError());
}

template <typename T>
void spk_precond_s::forward(const T* b, T* x) {
  if (level_launch) {
    level_sweep<T, true>(b, x);
    return;
  }
  hipStream_t s = stream();
  CsrView<T> v = L->view<T>();
  hipLaunchKernelGGL((fill_sentinel_kernel<T>), dim3(blocks_for(n)),
                     dim3(256), 0, s, x, n);
  hipLaunchKernelGGL((sweep_kernel<T, true>), dim3(sweep_grid_l),
                     dim3(256), 0, s, sweep_rows_l.p, (int)sweep_len_l, v.rp,
                     v.ci, v.v, (T*)b, x, false, (T*)nullptr, sweep_error.p);
  SPK_HIP_TRY(hipGetLastError());
}

template <typename T>
void spk_precond_s::backward(const T* b, T* x) {
  if (level_launch) {
    level_sweep<T, false>(b, x);
    return;
  }
  hipStream_t s = stream();
  CsrView<T> v = U->view<T>();
  hipLaunchKernelGGL((fill_sentinel_kernel<T>), dim3(blocks_for(n)),
                     dim3(256), 0, s, x, n);
  hipLaunchKernelGGL((sweep_kernel<T, false>), dim3(sweep_grid_u),
                     dim3(256), 0, s, sweep_rows_u.p, (int)sweep_len_u, v.rp,
                     v.ci, v.v, (T*)b, x, false, (T*)nullptr, sweep_error.p);
  SPK_HIP_TRY(hipGetLastError());
}

// strict part of one factor row: L drops its last entry (the unit
// diagonal), U its first (the diagonal, kept in d)
__global__ void strict_copy_kernel(int64_t n, const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci,
                                   const double* __restrict__ v, bool lower,
                                   int32_t* __restrict__ sci,
                                   double* __restrict__ sv,
                                   double* __restrict__ d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = rp[i], b = rp[i + 1];
  if (lower) {
    for (int k = a; k < b - 1; ++k) { sci[k - i] = ci[k]; sv[k - i] = v[k]; }
  } else {
    d[i] = v[a];
    for (int k = a + 1; k < b; ++k) { sci[k - i - 1] = ci[k]; sv[k - i - 1] = v[k]; }
  }
}

void spk_precond_s::build_strict() {
  if (Ls) return;
  hipStream_t s = stream();
  du64.alloc((size_t)(n > 0 ? n : 1));
  for (int f = 0; f < 2; ++f) {
    spk_csr_s* F = f == 0 ? L : U;
    std::vector<int32_t> hrp((size_t)n + 1);
    for (int64_t i = 0; i <= n; ++i) hrp[(size_t)i] = F->h_rp[(size_t)i] - (int32_t)i;
    auto* S = new spk_csr_s();
    try {
      S->allocate(n, n, F->nnz - n);
      SPK_HIP_TRY(hipMemcpyAsync(S->rp.p, hrp.data(), (n + 1) * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
      if (n > 0)
        hipLaunchKernelGGL(strict_copy_kernel, dim3(blocks_for(n)), dim3(256), 0,
                           s, n, (const int32_t*)F->rp.p, (const int32_t*)F->ci.p,
                           (const double*)F->v64.p, f == 0, S->ci.p, S->v64.p,
                           du64.p);
      SPK_HIP_TRY(hipGetLastError());
      SPK_HIP_TRY(hipStreamSynchronize(s));
      finish_store(*S, false);
    } catch (...) {
      delete S;
      throw;
    }
    (f == 0 ? Ls : Us) = S;
  }
}

template <typename T>
struct EpiJacLastForward {  // y = r - L_s y_old; z_0 = y / d (z may alias r)
  static constexpr bool kReduce = false;
  const T* r;
  T* __restrict__ y;
  const T* __restrict__ d;
  T* z;
  const int* done = nullptr;
  __device__ double row(int i, T s) const {
    const T yi = r[i] - s;
    y[i] = yi;
    z[i] = yi / d[i];
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
struct EpiJacBackward {  // z = (y - U_s z_old) / d
  static constexpr bool kReduce = false;
  const T* __restrict__ y;
  const T* __restrict__ d;
  T* __restrict__ z;
  const int* done = nullptr;
  __device__ double row(int i, T s) const {
    z[i] = (y[i] - s) / d[i];
    return 0.0;
  }
  __device__ void finalize(double) const {}
};

template <typename T>
static T* scratch_vec(DevBuf<T>& b, int64_t n) {
  if (b.n < (size_t)n) b.alloc((size_t)(n > 0 ? n : 1));
  return b.p;
}

template <typename T>
static void jac_buffers(spk_precond_s& P, T** a, T** b);
template <>
void jac_buffers<double>(spk_precond_s& P, double** a, double** b) {
  *a = scratch_vec(P.jb64a, P.n);
  *b = scratch_vec(P.jb64b, P.n);
}
template <>
void jac_buffers<float>(spk_precond_s& P, float** a, float** b) {
  *a = scratch_vec(P.jb32a, P.n);
  *b = scratch_vec(P.jb32b, P.n);
}
// a third buffer: z_0 for an odd sweep count >= 3, whose last forward pass
// gathers from b
template <typename T>
static T* jac_buffer_c(spk_precond_s& P);
template <>
double* jac_buffer_c<double>(spk_precond_s& P) { return scratch_vec(P.jb64c, P.n); }
template <>
float* jac_buffer_c<float>(spk_precond_s& P) { return scratch_vec(P.jb32c, P.n); }

template <typename T>
static const T* diag_u(spk_precond_s& P);
template <>
const double* diag_u<double>(spk_precond_s& P) { return P.du64.p; }
template <>
const float* diag_u<float>(spk_precond_s& P) {
  if (!P.du_has32) {
    P.du32.alloc((size_t)(P.n > 0 ? P.n : 1));
    if (P.n > 0)
      hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(P.n)), dim3(256), 0,
                         stream(), (const double*)P.du64.p, P.du32.p, P.n);
    SPK_HIP_TRY(hipGetLastError());
    P.du_has32 = true;
  }
  return P.du32.p;
}

template <typename T>
void spk_precond_s::jacobi_apply(const T* in, T* out) {
  hipStream_t s = stream();
  const int k = jac_sweeps;
  T *a, *b;
  jac_buffers<T>(*this, &a, &b);
  const T* d = diag_u<T>(*this);
  // forward: y_0 = r, y_j = r - L_s y_{j-1}; y_k ends in a
  // backward: z_0 = y / d, z_j = (y - U_s z_{j-1}) / d; z_k ends in out
  // z_0 is written by the last forward pass: into out for an even k (in
  // may alias out: that pass reads in[i] only for its own row, before
  // writing z_0[i], and gathers from b), into b for k = 1 (its one pass
  // gathers in), into a third buffer for an odd k >= 3 (its last pass
  // gathers from b)
  T* z0 = k % 2 == 0 ? out : (k == 1 ? b : jac_buffer_c<T>(*this));
  const T* y = in;
  for (int j = 0; j < k; ++j) {
    T* dst = (k - j) % 2 == 1 ? a : b;
    if (j == k - 1)
      launch_spmv(*Ls, y, EpiJacLastForward<T>{in, dst, d, z0}, s);
    else
      launch_spmv(*Ls, y, EpiResidual<T>{in, dst}, s);
    y = dst;
  }
  if (k == 0) launch_map(n, [=] __device__(int64_t i) { z0[i] = y[i] / d[i]; });
  const T* z = z0;
  for (int j = 0; j < k; ++j) {
    T* dst = (k - j) % 2 == 1 ? out : b;
    launch_spmv(*Us, z, EpiJacBackward<T>{y, d, dst}, s);
    z = dst;
  }
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
      // ilu_apply: y = forward_solve(L, r); z = backward_solve(U, y).
      // Two launches: the forward sweep marks `out` with the sentinel after
      // reading in[i] (in may alias out), the backward sweep returns y to
      // the sentinel as it consumes it.
      if (jac_sweeps > 0) {
        jacobi_apply(in, out);
        break;
      }
      T* y = tmp<T>();
      if (level_launch) {  // wide levels: one launch per level, no waiting
        level_sweep<T, true>(in, y);
        level_sweep<T, false>((const T*)y, out);
        break;
      }
      hipStream_t s = stream();
      CsrView<T> lv = L->view<T>(), uv = U->view<T>();
      hipLaunchKernelGGL((sweep_kernel<T, true>), dim3(sweep_grid_l),
                         dim3(256), 0, s, sweep_rows_l.p, (int)sweep_len_l,
                         lv.rp, lv.ci, lv.v, (T*)in, y, false, out,
                         sweep_error.p);
      hipLaunchKernelGGL((sweep_kernel<T, false>), dim3(sweep_grid_u),
                         dim3(256), 0, s, sweep_rows_u.p, (int)sweep_len_u,
                         uv.rp, uv.ci, uv.v, y, out, true, (T*)nullptr,
                         sweep_error.p);
      SPK_HIP_TRY(hipGetLastError());
      break;
    }
    case kSpmv:
      launch_spmv(*M, in, EpiStore<T>{out}, stream());
      break;
    case kInnerCg:
      inner_cg_apply(in, out);
      break;
    case kInnerIr:
      inner_ir_apply(in, out);
      break;
  }
}

// z = up to inner_iters PCG steps on A z = r from z = 0 (the inner solve's
// vectors in inner_precision; r and z converted at the boundary)
template <typename U, typename T>
static void inner_cg_run(spk_precond_s& P, const T* in, T* out,
                         spk::DevBuf<U>& bb, spk::DevBuf<U>& zb) {
  const int64_t n = P.n;
  const U* b;
  U* z;
  if constexpr (std::is_same<U, T>::value) {
    b = in;
    z = out;
    if (in == out) {  // the solve reads b while it writes z
      bb.reserve((size_t)(n > 0 ? n : 1));
      SPK_HIP_TRY(hipMemcpyAsync(bb.p, in, (size_t)n * sizeof(U),
                                 hipMemcpyDeviceToDevice, stream()));
      b = bb.p;
    }
  } else {
    bb.reserve((size_t)(n > 0 ? n : 1));
    zb.reserve((size_t)(n > 0 ? n : 1));
    U* bp = bb.p;
    launch_map(n, [=] __device__(int64_t i) { bp[i] = (U)in[i]; });
    b = bp;
    z = zb.p;
  }
  launch_map(n, [=] __device__(int64_t i) { z[i] = U(0); });
  spk_solver_config c{P.inner_tol, P.inner_iters, P.inner_iters, 0,
                      P.inner_precision, 0, 0};
  spk_summary sum{};
  const spk_status st = spk_cg_solve(P.inner_A, P.inner_M, nullptr, &c, b, z,
                                     &sum, nullptr);
  if (st != SPK_OK && st != SPK_BREAKDOWN)
    throw Error{st, std::string("inner PCG: ") + spk_last_error()};
  P.inner_total += sum.computed_iterations;
  if constexpr (!std::is_same<U, T>::value) {
    const U* zp = z;
    launch_map(n, [=] __device__(int64_t i) { out[i] = (T)zp[i]; });
  }
}

template <typename T>
void spk_precond_s::inner_cg_apply(const T* in, T* out) {
  if (inner_precision == SPK_F32)
    inner_cg_run<float, T>(*this, in, out, ib32, iz32);
  else
    inner_cg_run<double, T>(*this, in, out, ib64, iz64);
}

// z = spk_ir_solve(inner_A z = r) from z = 0 with inner_M and inner_ir (the
// solve's right-hand side and result are FP64: converted at the boundary
// when the caller's vectors are FP32)
template <typename T>
void spk_precond_s::inner_ir_apply(const T* in, T* out) {
  const double* b;
  double* z;
  const size_t m = (size_t)(n > 0 ? n : 1);
  if constexpr (std::is_same<T, double>::value) {
    b = in;
    z = out;
    if (in == out) {  // the solve reads b while it writes z
      ib64.reserve(m);
      SPK_HIP_TRY(hipMemcpyAsync(ib64.p, in, (size_t)n * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream()));
      b = ib64.p;
    }
  } else {
    ib64.reserve(m);
    iz64.reserve(m);
    double* bp = ib64.p;
    launch_map(n, [=] __device__(int64_t i) { bp[i] = (double)in[i]; });
    b = bp;
    z = iz64.p;
  }
  launch_map(n, [=] __device__(int64_t i) { z[i] = 0.0; });
  spk_ir_summary sum{};
  const spk_status st = spk_ir_solve(inner_A, inner_M, &inner_ir, b, z, &sum,
                                     nullptr);
  if (st != SPK_OK && st != SPK_BREAKDOWN)
    throw Error{st, std::string("inner mixed-precision solve: ") + spk_last_error()};
  inner_total += sum.inner_iterations;
  if constexpr (!std::is_same<T, double>::value) {
    const double* zp = z;
    launch_map(n, [=] __device__(int64_t i) { out[i] = (T)zp[i]; });
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

spk_status spk_ilu0_set_schedule(spk_precond M, int schedule) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu, SPK_INVALID_ARGUMENT,
                "not an ILU(0) preconditioner");
    SPK_REQUIRE(schedule >= 0 && schedule <= 2, SPK_INVALID_ARGUMENT,
                "bad schedule");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    if (schedule == 0) {
      const int64_t nl = (int64_t)M->lvl_ptr_l.size() - 1;
      const int64_t nu = (int64_t)M->lvl_ptr_u.size() - 1;
      M->level_launch = M->n >= kLevelLaunchRows &&
                        std::max(nl, nu) <= kLevelLaunchMaxLevels;
    } else {
      M->level_launch = schedule == 1;
    }
    // the per-level schedule leaves ordinary values in the scratch vector;
    // the sync-free one needs it filled with the ready sentinel
    if (!M->level_launch) M->reset_sentinels();
  });
}

spk_status spk_ilu0_set_jacobi_sweeps(spk_precond M, int sweeps) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu, SPK_INVALID_ARGUMENT,
                "not an ILU(0) preconditioner");
    SPK_REQUIRE(sweeps >= 0 && sweeps <= 1024, SPK_INVALID_ARGUMENT,
                "sweeps must be in [0, 1024]");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    if (sweeps > 0) M->build_strict();
    M->jac_sweeps = sweeps;
  });
}

spk_status spk_ilu0_set_half_factors(spk_precond M, int enable) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu, SPK_INVALID_ARGUMENT,
                "not an ILU(0) preconditioner");
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    M->build_strict();
    for (spk_csr_s* S : {M->Ls, M->Us}) {
      const spk_status st = spk_csr_set_half_values(S, enable);
      if (st != SPK_OK)
        throw Error{st, std::string("ILU(0) strict factors: ") + spk_last_error()};
    }
  });
}

spk_status spk_ilu0_get_half_factors(spk_precond M, int* enabled) {
  return guarded([&] {
    SPK_REQUIRE(M && enabled && M->kind == spk_precond_s::kIlu,
                SPK_INVALID_ARGUMENT, "not an ILU(0) preconditioner");
    *enabled = M->Ls && M->Ls->half32 ? 1 : 0;
  });
}

spk_status spk_ilu0_get_jacobi_sweeps(spk_precond M, int* sweeps) {
  return guarded([&] {
    SPK_REQUIRE(M && sweeps && M->kind == spk_precond_s::kIlu,
                SPK_INVALID_ARGUMENT, "not an ILU(0) preconditioner");
    *sweeps = M->jac_sweeps;
  });
}

spk_status spk_ilu0_get_schedule(spk_precond M, int* schedule) {
  return guarded([&] {
    SPK_REQUIRE(M && schedule && M->kind == spk_precond_s::kIlu,
                SPK_INVALID_ARGUMENT, "not an ILU(0) preconditioner");
    *schedule = M->level_launch ? 1 : 2;
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

spk_status spk_precond_create_inner_cg(spk_csr A, spk_precond M_inner,
                                       int64_t iterations, double tolerance,
                                       int precision, spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(A && out, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "inner PCG requires a square matrix");
    SPK_REQUIRE(!M_inner || M_inner->n == A->nrows, SPK_INVALID_ARGUMENT,
                "inner preconditioner size mismatch");
    SPK_REQUIRE(!M_inner || !M_inner->runs_solve(), SPK_INVALID_ARGUMENT,
                "inner PCG cannot nest another inner solve");
    SPK_REQUIRE(iterations >= 1 && tolerance >= 0.0, SPK_INVALID_ARGUMENT,
                "iterations >= 1, tolerance >= 0");
    SPK_REQUIRE(precision == SPK_F32 || precision == SPK_F64,
                SPK_INVALID_ARGUMENT, "precision must be SPK_F32 or SPK_F64");
    auto* P = new spk_precond_s();
    P->kind = spk_precond_s::kInnerCg;
    P->n = A->nrows;
    P->inner_A = A;
    P->inner_M = M_inner;
    P->inner_iters = iterations;
    P->inner_tol = tolerance;
    P->inner_precision = precision;
    *out = P;
  });
}

spk_status spk_precond_create_inner_ir(spk_csr A, spk_precond M_inner,
                                       const spk_ir_config* cfg,
                                       spk_precond* out) {
  return guarded([&] {
    SPK_REQUIRE(A && cfg && out, SPK_INVALID_ARGUMENT, "null argument");
    SPK_REQUIRE(A->nrows == A->ncols, SPK_INVALID_ARGUMENT,
                "inner solve requires a square matrix");
    SPK_REQUIRE(!M_inner || M_inner->n == A->nrows, SPK_INVALID_ARGUMENT,
                "inner preconditioner size mismatch");
    SPK_REQUIRE(!M_inner || !M_inner->runs_solve(), SPK_INVALID_ARGUMENT,
                "an inner solve cannot nest another inner solve");
    // the (F)GMRES this preconditioner serves keeps its basis in the GMRES
    // work vectors, which an inner (F)GMRES would reuse
    SPK_REQUIRE(cfg->inner_method == SPK_IR_CG ||
                    cfg->inner_method == SPK_IR_CG_RELIABLE ||
                    cfg->inner_method == SPK_IR_BICGSTAB,
                SPK_INVALID_ARGUMENT,
                "inner solve: inner_method must be SPK_IR_CG, "
                "SPK_IR_CG_RELIABLE or SPK_IR_BICGSTAB");
    SPK_REQUIRE(cfg->tolerance > 0.0 && cfg->max_outer >= 1 && cfg->max_inner >= 1,
                SPK_INVALID_ARGUMENT,
                "inner solve: tolerance > 0, max_outer >= 1, max_inner >= 1");
    auto* P = new spk_precond_s();
    P->kind = spk_precond_s::kInnerIr;
    P->n = A->nrows;
    P->inner_A = A;
    P->inner_M = M_inner;
    P->inner_ir = *cfg;
    *out = P;
  });
}

spk_status spk_precond_inner_cg_iterations(spk_precond M, int64_t* total) {
  return guarded([&] {
    SPK_REQUIRE(M && total && M->runs_solve(), SPK_INVALID_ARGUMENT,
                "not an inner-solve preconditioner");
    *total = M->inner_total;
  });
}

spk_status spk_precond_create_ilu0(spk_csr A, int modified, spk_precond* out) {
  return guarded([&] {
    TraceRange trace_("spk_precond_create_ilu0");
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
      M->sweep_error.alloc(kErrorWords);
      SPK_HIP_TRY(hipMemsetAsync(M->sweep_error.p, 0, sizeof(int) * kErrorWords, s));
      std::vector<int32_t> rows_l, rows_u;
      M->lvl_ptr_l = group_levels(lvl_l, rows_l);
      M->lvl_ptr_u = group_levels(lvl_u, rows_u);
      M->lvl_rows_l.alloc((size_t)n);
      M->lvl_rows_u.alloc((size_t)n);
      std::vector<int32_t> pad_l = pad_levels(rows_l, M->lvl_ptr_l);
      std::vector<int32_t> pad_u = pad_levels(rows_u, M->lvl_ptr_u);
      SPK_REQUIRE(pad_l.size() < (size_t)INT32_MAX && pad_u.size() < (size_t)INT32_MAX,
                  SPK_INVALID_ARGUMENT, "too many rows for the sweep order");
      M->sweep_len_l = (int64_t)pad_l.size();
      M->sweep_len_u = (int64_t)pad_u.size();
      {
        const int64_t nl = (int64_t)M->lvl_ptr_l.size() - 1;
        const int64_t nu = (int64_t)M->lvl_ptr_u.size() - 1;
        M->level_launch = n >= kLevelLaunchRows &&
                          std::max(nl, nu) <= kLevelLaunchMaxLevels;
        // sync-free window: about four padded levels of threads
        auto grid = [&](int64_t len, int64_t levels) {
          const int64_t w = levels > 0 ? len / levels : 64;
          int64_t g = (4 * w + 255) / 256;
          const int64_t cap = std::max<int64_t>(4, 4 * ctx().num_cus);
          g = std::max<int64_t>(4, std::min(g, cap));
          return (unsigned)std::min<int64_t>(g, blocks_for(std::max<int64_t>(len, 1)));
        };
        M->sweep_grid_l = grid(M->sweep_len_l, nl);
        M->sweep_grid_u = grid(M->sweep_len_u, nu);
      }
      M->sweep_rows_l.alloc(pad_l.size());
      M->sweep_rows_u.alloc(pad_u.size());
      SPK_HIP_TRY(hipMemcpyAsync(M->sweep_rows_l.p, pad_l.data(), pad_l.size() * 4,
                                 hipMemcpyHostToDevice, s));
      SPK_HIP_TRY(hipMemcpyAsync(M->sweep_rows_u.p, pad_u.data(), pad_u.size() * 4,
                                 hipMemcpyHostToDevice, s));
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
    TraceRange trace_("spk_precond_apply");
    SPK_REQUIRE(M && d_in && d_out, SPK_INVALID_ARGUMENT, "null argument");
    if (dtype == SPK_F32) M->apply((const float*)d_in, (float*)d_out);
    else M->apply((const double*)d_in, (double*)d_out);
    M->check_sweeps();
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

// The sync-free sweeps mark x with a sentinel before they read b, so an
// input that overlaps the output would be overwritten first.
static bool vectors_overlap(const void* b, const void* x, int64_t n, int dtype) {
  const size_t bytes = (size_t)(n > 0 ? n : 0) * (dtype == SPK_F32 ? 4 : 8);
  const uintptr_t pb = (uintptr_t)b, px = (uintptr_t)x;
  return pb == px || (pb < px + bytes && px < pb + bytes);
}

spk_status spk_ilu0_forward(spk_precond M, int dtype, const void* d_b,
                            void* d_x) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu && d_b && d_x,
                SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(dtype == SPK_F32 || dtype == SPK_F64, SPK_INVALID_ARGUMENT,
                "dtype must be SPK_F32 or SPK_F64");
    SPK_REQUIRE(!vectors_overlap(d_b, d_x, M->n, dtype), SPK_INVALID_ARGUMENT,
                "ilu0 triangular sweep: b and x must not overlap");
    if (dtype == SPK_F32) M->forward((const float*)d_b, (float*)d_x);
    else M->forward((const double*)d_b, (double*)d_x);
    M->check_sweeps();
  });
}

spk_status spk_ilu0_backward(spk_precond M, int dtype, const void* d_b,
                             void* d_x) {
  return guarded([&] {
    SPK_REQUIRE(M && M->kind == spk_precond_s::kIlu && d_b && d_x,
                SPK_INVALID_ARGUMENT, "bad argument");
    SPK_REQUIRE(dtype == SPK_F32 || dtype == SPK_F64, SPK_INVALID_ARGUMENT,
                "dtype must be SPK_F32 or SPK_F64");
    SPK_REQUIRE(!vectors_overlap(d_b, d_x, M->n, dtype), SPK_INVALID_ARGUMENT,
                "ilu0 triangular sweep: b and x must not overlap");
    if (dtype == SPK_F32) M->backward((const float*)d_b, (float*)d_x);
    else M->backward((const double*)d_b, (double*)d_x);
    M->check_sweeps();
  });
}

}  // extern "C"
