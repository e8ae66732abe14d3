// spk_csr.hpp -- the device-resident compressed_row store, its row-binned
// SpMV plan and the SpMV kernel template shared by the solvers.
//
// Layout in HBM (one store):
//   rp[nrows+1]      int32   row offsets      (Compressed_row_sparsity::row_ptr)
//   ci[nnz+kPad]     int32   column indices   (Compressed_row_sparsity::col_ind)
//   v64[nnz+kPad]    double  values           (Compressed_row_matrix::values)
//   v32[nnz+kPad]    float   FP32 copy, made on first FP32 use
//   plan[nb+1]       int2    {first row, first nonzero} of each SpMV row block
// The kPad zeroed tail lets the kernel read index/value PAIRS from an even
// start without leaving the allocation.
//
// Plan: rows are cut greedily into blocks of at most kSpmvBlock rows and at
// most kNnzBlock nonzeros; a row longer than kNnzBlock gets a block of its
// own (a "long" block).  One workgroup of kSpmvBlock threads runs per block:
//  - short/medium rows: the block's nonzeros are streamed with 16-byte
//    (pairwise) nontemporal loads of ci and v -- the matrix is read once per
//    SpMV, so it bypasses the caches that keep x --, multiplied by the
//    gathered x and staged in LDS; then each row is summed from LDS by L
//    lanes (L: power of two, at most 64, L * rows <= kSpmvBlock, about eight
//    products per lane).  With L = 1 a lane adds its row's products in
//    storage order, which is sparse_blas.hpp:26-31's order, so FP64 results
//    are bit-identical to the reference for such rows (all stencil rows).
//  - long rows: all threads stride over the row, then a block reduction.
// Blocks are dispatched in plan order: measured on MI355X, keeping
// consecutive row blocks on different XCDs (the hardware's round-robin) is
// faster than remapping contiguous ranges onto one XCD (the x working set
// of a 7-point stencil is then shared through the Infinity Cache).
#pragma once

#include "spk_internal.hpp"
#include "spk_vec.hpp"

namespace spk {

constexpr int kSpmvBlock = 512;
constexpr int kNnzBlock = 4096;
constexpr int kPad = 4;  // zeroed tail elements of ci / v

template <typename T>
struct CsrView {
  const int32_t* __restrict__ rp;
  const int32_t* __restrict__ ci;
  const T* __restrict__ v;
  const int2* __restrict__ plan;  // {first row, first nonzero} per block
  int nblocks;
  const int32_t* __restrict__ rows;  // matrix row of each stored row (null: same)
};

}  // namespace spk

struct spk_csr_s {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  spk::DevBuf<int32_t> rp, ci;
  spk::DevBuf<int2> plan;
  spk::DevBuf<double> v64;
  spk::DevBuf<float> v32;
  bool has32 = false;
  int nblocks = 0, nlong = 0;
  std::vector<int32_t> h_rp;  // host copy of the row offsets
  // Column panels (empty: none).  Panel p holds the entries with columns in
  // [p W, (p+1) W) as a store of its own; an SpMV runs the panels in column
  // order, each continuing every row's running sum from the previous panel
  // (pacc), so the gathers of one launch stay inside an L2-sized window of
  // x.  Values are copied when the panels are made.  The inner panels (all
  // but the first and the last) keep only their non-empty rows: `rows` maps
  // a stored row to its matrix row, and such a panel adds its sums to the
  // running sums in place, so a row costs nothing in a panel it has no
  // entries in.
  std::vector<spk_csr_s*> panels;
  spk::DevBuf<int32_t> rows;  // compressed store: matrix row of each row
  spk::DevBuf<double> pacc64;
  spk::DevBuf<float> pacc32;

  spk_csr_s() = default;
  spk_csr_s(const spk_csr_s&) = delete;
  spk_csr_s& operator=(const spk_csr_s&) = delete;
  ~spk_csr_s() { drop_panels(); }
  void drop_panels() {
    for (auto* q : panels) delete q;
    panels.clear();
  }
  // allocates rp, ci and v64 (with the zeroed kPad tail)
  void allocate(int64_t nrows_, int64_t ncols_, int64_t nnz_);
  const float* vals32();
  template <typename T>
  spk::CsrView<T> view();
  template <typename T>
  T* panel_acc();
};

template <>
inline double* spk_csr_s::panel_acc<double>() {
  if (pacc64.n < (size_t)nrows) pacc64.alloc((size_t)nrows);
  return pacc64.p;
}
template <>
inline float* spk_csr_s::panel_acc<float>() {
  if (pacc32.n < (size_t)nrows) pacc32.alloc((size_t)nrows);
  return pacc32.p;
}

template <>
inline spk::CsrView<double> spk_csr_s::view<double>() {
  return {rp.p, ci.p, v64.p, plan.p, nblocks, rows.p};
}
template <>
inline spk::CsrView<float> spk_csr_s::view<float>() {
  return {rp.p, ci.p, vals32(), plan.p, nblocks, rows.p};
}

namespace spk {

// Build the row-block plan from host row offsets: row-block boundaries.
std::vector<int32_t> build_spmv_plan(const int32_t* rp, int64_t nrows,
                                     int* nlong);
// Upload a plan ({first row, first nonzero} per boundary) into A.plan.
void upload_plan(spk_csr_s& A, const std::vector<int32_t>& blocks);
// Finish a store whose rp/ci/v64 are filled on the device: copies rp back,
// validates it and builds the plan.
void finish_store(spk_csr_s& A, bool validate_cols);
// Split A into `panels` column panels (1: none); values are copied.
void build_panels(spk_csr_s& A, int panels);

// ------------------------------------------------------------- epilogues
//
// An epilogue receives each row's sum: row(i, s) stores what it needs and
// returns the row's contribution to an optional grid-wide sum (kReduce);
// finalize(total) runs once, in the finisher kernel.  `done` (may be null)
// is a device flag that turns the launch into a no-op.

// y = A x
template <typename T>
struct EpiStore {
  static constexpr bool kReduce = false;
  T* __restrict__ y;
  const int* done = nullptr;
  __device__ double row(int i, T s) const { y[i] = s; return 0.0; }
  __device__ void finalize(double) const {}
};

// y = A x and <x, y> (square A): CG's q = A p with <p, q>
template <typename T>
struct EpiDotX {
  static constexpr bool kReduce = true;
  const T* __restrict__ x;
  T* __restrict__ y;
  double* out;
  const int* done = nullptr;
  __device__ double row(int i, T s) const {
    y[i] = s;
    return (double)x[i] * (double)s;
  }
  __device__ void finalize(double t) const { *out = t; }
};

// y = b - A x
template <typename T>
struct EpiResidual {
  static constexpr bool kReduce = false;
  const T* __restrict__ b;
  T* __restrict__ y;
  const int* done = nullptr;
  __device__ double row(int i, T s) const { y[i] = b[i] - s; return 0.0; }
  __device__ void finalize(double) const {}
};

// y += A x, in place: the running sum arrives as init == y, so y is not
// restrict-qualified
template <typename T>
struct EpiAccum {
  static constexpr bool kReduce = false;
  T* y;
  const int* done = nullptr;
  __device__ double row(int i, T s) const { y[i] = s; return 0.0; }
  __device__ void finalize(double) const {}
};

template <typename Epi>
struct EpiFinalize {  // adapts Epi::finalize to the finisher's fin(totals)
  Epi epi;
  __device__ void operator()(const double* t) const { epi.finalize(t[0]); }
};

// ------------------------------------------------------------- kernel

template <typename T>
struct Pair;
template <>
struct Pair<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
typedef int IdxPair __attribute__((ext_vector_type(2)));

// init (may be null): each row's sum starts from init[row] (column panels;
// init may be the epilogue's own output).  Stored row r is matrix row
// A.rows[r] when the store is compressed.
template <typename T, typename Epi>
__global__ __launch_bounds__(kSpmvBlock) void spmv_kernel(
    CsrView<T> A, const T* __restrict__ x, Epi epi, double* partials,
    const T* init) {
  __shared__ T prod[kNnzBlock + 2];
  __shared__ double red[kSpmvBlock / 64];
  if (epi.done && *epi.done) return;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int2 p0 = A.plan[b], p1 = A.plan[b + 1];
  const int r0 = p0.x, base = p0.y, nrows = p1.x - p0.x, cnt = p1.y - p0.y;
  double acc = 0.0;
  if (cnt > kNnzBlock) {
    // long row: the block holds exactly one row
    T s = T(0);
    for (int k = base + tid; k < base + cnt; k += kSpmvBlock)
      s += __builtin_nontemporal_load(A.v + k) * x[__builtin_nontemporal_load(A.ci + k)];
    s = block_reduce_sum<T, kSpmvBlock>(s, reinterpret_cast<T*>(red));
    if (tid == 0) {
      const int ri = A.rows ? A.rows[r0] : r0;
      acc = epi.row(ri, init ? init[ri] + s : s);
    }
  } else {
    // lanes per row: enough rows to occupy the block, about eight products
    // per lane; short rows (stencils) get one lane each
    int L = 1;
    while (L < 64 && 2 * L * nrows <= kSpmvBlock && 16 * L * nrows <= cnt) L *= 2;
    // pairs of (index, value) from the even-aligned start k0 <= base; every
    // load of the block and the row bounds are issued before the gathers
    const int k0 = base & ~1, off = base - k0, span = cnt + off;
    constexpr int U = (kNnzBlock + 2) / (2 * kSpmvBlock) + 1;
    typedef typename Pair<T>::type TP;
    IdxPair c[U];
    TP a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 2 * tid + 2 * u * kSpmvBlock;
      if (k < span) {
        c[u] = __builtin_nontemporal_load((const IdxPair*)(A.ci + k0 + k));
        a[u] = __builtin_nontemporal_load((const TP*)(A.v + k0 + k));
      } else {
        c[u] = IdxPair{0, 0};
        a[u] = TP{T(0), T(0)};
      }
    }
    const int g = tid / L, lane = tid % L;
    int s0 = 0, s1 = 0, ri = r0 + g;
    if (g < nrows) {
      s0 = A.rp[r0 + g] - base + off;
      s1 = A.rp[r0 + g + 1] - base + off;
      if (A.rows) ri = A.rows[r0 + g];
    }
    T xv0[U], xv1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 2 * tid + 2 * u * kSpmvBlock;
      xv0[u] = (k < span) ? x[c[u].x] : T(0);
      xv1[u] = (k + 1 < span) ? x[c[u].y] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 2 * tid + 2 * u * kSpmvBlock;
      if (k < span) prod[k] = a[u].x * xv0[u];
      if (k + 1 < span) prod[k + 1] = a[u].y * xv1[u];
    }
    __syncthreads();
    if (L == 1) {
      // storage-order sum (sparse_blas.hpp:26-31)
      if (g < nrows) {
        T s = init ? init[ri] : T(0);
        for (int j = s0; j < s1; ++j) s += prod[j];
        acc = epi.row(ri, s);
      }
    } else {
      T s = T(0);
      if (g < nrows)
        for (int j = s0 + lane; j < s1; j += L) s += prod[j];
      for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (g < nrows && lane == 0)
        acc = epi.row(ri, init ? init[ri] + s : s);
    }
  }
  if constexpr (Epi::kReduce) {
    const double t = block_sum_to_thread0<kSpmvBlock>(acc, red);
    if (tid == 0) partials[b] = t;
  }
}

template <typename T, typename Epi>
void launch_spmv_one(spk_csr_s& A, const T* x, const Epi& epi, const T* init,
                     hipStream_t s) {
  double* parts = Epi::kReduce ? reduce_partials(A.nblocks) : nullptr;
  hipLaunchKernelGGL((spmv_kernel<T, Epi>), dim3(A.nblocks), dim3(kSpmvBlock),
                     0, s, A.view<T>(), x, epi, parts, init);
  SPK_HIP_TRY(hipGetLastError());
  if constexpr (Epi::kReduce) {
    hipLaunchKernelGGL((finish_kernel<1, EpiFinalize<Epi>>), dim3(1),
                       dim3(kFinishBlock), 0, s, (const double*)parts,
                       A.nblocks, EpiFinalize<Epi>{epi}, epi.done);
    SPK_HIP_TRY(hipGetLastError());
  }
}

// All column panels but the last: the first stores every row's sum, the
// compressed inner ones add theirs in place.  Returns the running sums.
template <typename T>
const T* launch_panels(spk_csr_s& A, const T* x, const int* done, hipStream_t s) {
  T* acc = A.panel_acc<T>();
  for (size_t p = 0; p + 1 < A.panels.size(); ++p) {
    spk_csr_s& Q = *A.panels[p];
    if (p == 0) {
      EpiStore<T> part;
      part.y = acc;
      part.done = done;
      launch_spmv_one(Q, x, part, (const T*)nullptr, s);
    } else if (Q.nblocks > 0) {
      EpiAccum<T> part;
      part.y = acc;
      part.done = done;
      launch_spmv_one(Q, x, part, (const T*)acc, s);
    }
  }
  return acc;
}

// The reduction's block partials (A.nblocks of them, or the last panel's)
// go to `parts` and no finisher runs: the next kernel sums them itself.
template <typename T, typename Epi>
int launch_spmv_partials(spk_csr_s& A, const T* x, const Epi& epi, double* parts,
                         hipStream_t s) {
  static_assert(Epi::kReduce, "an epilogue with a sum");
  spk_csr_s& last = A.panels.empty() ? A : *A.panels.back();
  const T* init = nullptr;
  if (!A.panels.empty()) init = launch_panels(A, x, epi.done, s);
  hipLaunchKernelGGL((spmv_kernel<T, Epi>), dim3(last.nblocks), dim3(kSpmvBlock),
                     0, s, last.view<T>(), x, epi, parts, init);
  SPK_HIP_TRY(hipGetLastError());
  return last.nblocks;
}

// the number of partials launch_spmv_partials writes
inline int spmv_partials_count(const spk_csr_s& A) {
  return A.panels.empty() ? A.nblocks : A.panels.back()->nblocks;
}

template <typename T, typename Epi>
void launch_spmv(spk_csr_s& A, const T* x, const Epi& epi, hipStream_t s) {
  if (A.nrows == 0) return;
  if (A.panels.empty()) {
    launch_spmv_one(A, x, epi, (const T*)nullptr, s);
    return;
  }
  const T* init = launch_panels(A, x, epi.done, s);
  launch_spmv_one(*A.panels.back(), x, epi, init, s);
}

}  // namespace spk
