// spk_csr.hpp -- the device-resident compressed_row store, its row-binned
// SpMV plan and the SpMV kernel template shared by the solvers.
//
// Layout in HBM (one store):
//   rp[nrows+1]      int32   row offsets      (Compressed_row_sparsity::row_ptr)
//   ci[nnz+kPad]     int32   column indices   (Compressed_row_sparsity::col_ind)
//   v64[nnz+kPad]    double  values           (Compressed_row_matrix::values)
//   v32[nnz+kPad]    float   FP32 copy, made on first FP32 use
//   plan[nb+1]       int2    {first row, first nonzero} of each SpMV row block
//   c16[nnz+kPad]    int16   optional: column - first row of the entry's block
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

#include <type_traits>

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
  // null, or short column indices: column = the block's first row + c16[k]
  const int16_t* __restrict__ c16;
};

}  // namespace spk

struct spk_csr_s {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  spk::DevBuf<int32_t> rp, ci;
  spk::DevBuf<int2> plan;
  spk::DevBuf<double> v64;
  spk::DevBuf<float> v32;
  bool has32 = false;
  // FP16 values (half32): FP32 SpMVs of the plain layout read the values
  // rounded to FP16 -- half the value bytes, for a preconditioner such as
  // an approximate inverse, whose rounding the FP64 true residual of the
  // solve it serves absorbs.  FP64 SpMVs keep v64.
  spk::DevBuf<_Float16> v16;
  bool has16 = false;
  bool half32 = false;
  // Short column indices (short_idx): c16[k] = column - the first row of
  // entry k's row block, for stores whose every entry lies within 32767
  // columns of its block's first row (banded matrices in natural order,
  // such as stencils): the row kernel then streams 2 index bytes per entry
  // instead of 4, in every precision.  Column panels and window tiles keep
  // their own indices.
  spk::DevBuf<int16_t> c16;
  bool short_idx = false;
  const int16_t* short_cols() const { return short_idx ? c16.p : nullptr; }
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
  // Window tiles (xt_groups == 0: none), for matrices whose rows scatter
  // over all columns.  Rows are cut into groups of kXtThreads * xt_slots;
  // thread t of a group owns its rows t + s * kXtThreads (slot s).  For
  // every window of xt_window columns each thread's entries of its rows in
  // that window are stored together, in (slot, column) order, as
  // (slot << 16 | column - window start, value), and a window's segments
  // follow each other in thread order; the SpMV stages the x window and the
  // window's entries in LDS with coalesced loads, so every gather is an LDS
  // read, and a row's sum runs in column order (the CSR storage order when
  // rows are sorted).  Rows longer than xt_lmax stay out of the tiles (their
  // owner's bit in xt_skip is set): they form the compressed store xt_long,
  // run by the row kernel after the tiles.
  int xt_slots = 0, xt_window = 0, xt_lmax = 0;
  int64_t xt_groups = 0, xt_windows = 0;
  spk::DevBuf<int64_t> xt_base;  // first entry of each group (groups + 1)
  spk::DevBuf<int32_t> xt_off;   // per group, (windows * threads + 1) offsets
  spk::DevBuf<uint32_t> xt_idx;
  spk::DevBuf<double> xt_v64;
  spk::DevBuf<float> xt_v32;
  spk::DevBuf<uint32_t> xt_skip;  // per (group, thread): slots of long rows
  spk_csr_s* xt_long = nullptr;
  bool xt_has32 = false;

  spk_csr_s() = default;
  spk_csr_s(const spk_csr_s&) = delete;
  spk_csr_s& operator=(const spk_csr_s&) = delete;
  ~spk_csr_s() { drop_panels(); }
  void drop_panels() {
    for (auto* q : panels) delete q;
    panels.clear();
    drop_tiles();
  }
  void drop_tiles() {
    xt_slots = xt_window = xt_lmax = 0;
    xt_groups = xt_windows = 0;
    xt_base.release();
    xt_off.release();
    xt_idx.release();
    xt_v64.release();
    xt_v32.release();
    xt_skip.release();
    delete xt_long;
    xt_long = nullptr;
    xt_has32 = false;
  }
  const float* tiles32();
  // allocates rp, ci and v64 (with the zeroed kPad tail)
  void allocate(int64_t nrows_, int64_t ncols_, int64_t nnz_);
  const float* vals32();
  const _Float16* vals16();
  template <typename T>
  spk::CsrView<T> view();
  spk::CsrView<_Float16> view16() {
    return {rp.p, ci.p, vals16(), plan.p, nblocks, rows.p, short_cols()};
  }
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
  return {rp.p, ci.p, v64.p, plan.p, nblocks, rows.p, short_cols()};
}
template <>
inline spk::CsrView<float> spk_csr_s::view<float>() {
  return {rp.p, ci.p, vals32(), plan.p, nblocks, rows.p, short_cols()};
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
// Window tiles with `slots` rows per thread (8 or 16) and `window` columns
// per x window (<= kXtMaxWindow); drops panels.  Rows longer than
// max(256, 8 * windows) go to the long-row store.  SPK_INVALID_ARGUMENT if a
// tiled row's columns are not ascending.
void build_tiles(spk_csr_s& A, int slots, int window);

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
template <>
struct Pair<_Float16> {
  typedef _Float16 type __attribute__((ext_vector_type(2)));
};
typedef int IdxPair __attribute__((ext_vector_type(2)));

// The row kernel's work for one block: the products of its entries in
// storage order and one sum per row.  x[j] is read as gx(j).  pre() runs
// after the block's index and value loads and its row bounds are issued and
// before the first gather (a hand-off kernel sums the previous kernel's
// partials there, so their latency overlaps the loads); every thread of the
// block calls it, and false ends the block.  row(ri, s, acc) gets each
// row's sum; W: the number of block sums in acc.  init (may be null): each
// row's sum starts from init[row] (column panels; init may be the
// epilogue's own output).  Stored row r is matrix row A.rows[r] when the
// store is compressed.  VT: the stored value type (T, or FP16 values
// widened to T on load).
template <int W, typename T, typename VT, typename Pre, typename Gx, typename Row>
__device__ __forceinline__ void spmv_rows(const CsrView<VT>& A, const T* init,
                                          Pre&& pre, const Gx& gx, const Row& row,
                                          T* prod, double* red, double* acc) {
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int2 p0 = A.plan[b], p1 = A.plan[b + 1];
  const int r0 = p0.x, base = p0.y, nrows = p1.x - p0.x, cnt = p1.y - p0.y;
  if (cnt > kNnzBlock) {
    // long row: the block holds exactly one row
    if (!pre()) return;
    T s = T(0);
    for (int k = base + tid; k < base + cnt; k += kSpmvBlock) {
      const int j = A.c16 ? r0 + (int)__builtin_nontemporal_load(A.c16 + k)
                          : __builtin_nontemporal_load(A.ci + k);
      s += (T)__builtin_nontemporal_load(A.v + k) * gx(j);
    }
    s = block_reduce_sum<T, kSpmvBlock>(s, reinterpret_cast<T*>(red));
    if (tid == 0) {
      const int ri = A.rows ? A.rows[r0] : r0;
      row(ri, init ? init[ri] + s : s, acc);
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
    typedef typename Pair<VT>::type TP;
    IdxPair c[U];
    TP a[U];
    if (A.c16) {
      // short indices, two per 32-bit load; the entry before base that an
      // odd base brings in belongs to the previous block (its offset is
      // from that block's first row): its column is replaced by 0, and its
      // product is never summed
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = 2 * tid + 2 * u * kSpmvBlock;
        if (k < span) {
          const int w = __builtin_nontemporal_load((const int*)(A.c16 + k0 + k));
          c[u] = IdxPair{r0 + (int)(int16_t)(w & 0xffff), r0 + (w >> 16)};
          a[u] = __builtin_nontemporal_load((const TP*)(A.v + k0 + k));
        } else {
          c[u] = IdxPair{0, 0};
          a[u] = TP{VT(0), VT(0)};
        }
      }
      if (2 * tid < off) c[0].x = 0;
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = 2 * tid + 2 * u * kSpmvBlock;
        if (k < span) {
          c[u] = __builtin_nontemporal_load((const IdxPair*)(A.ci + k0 + k));
          a[u] = __builtin_nontemporal_load((const TP*)(A.v + k0 + k));
        } else {
          c[u] = IdxPair{0, 0};
          a[u] = TP{VT(0), VT(0)};
        }
      }
    }
    const int g = tid / L, lane = tid % L;
    int s0 = 0, s1 = 0, ri = r0 + g;
    if (g < nrows) {
      s0 = A.rp[r0 + g] - base + off;
      s1 = A.rp[r0 + g + 1] - base + off;
      if (A.rows) ri = A.rows[r0 + g];
    }
    if (!pre()) return;
    T xv0[U], xv1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 2 * tid + 2 * u * kSpmvBlock;
      xv0[u] = (k < span) ? gx(c[u].x) : T(0);
      xv1[u] = (k + 1 < span) ? gx(c[u].y) : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = 2 * tid + 2 * u * kSpmvBlock;
      if (k < span) prod[k] = (T)a[u].x * xv0[u];
      if (k + 1 < span) prod[k + 1] = (T)a[u].y * xv1[u];
    }
    __syncthreads();
    if (L == 1) {
      // storage-order sum (sparse_blas.hpp:26-31)
      if (g < nrows) {
        T s = init ? init[ri] : T(0);
        for (int j = s0; j < s1; ++j) s += prod[j];
        row(ri, s, acc);
      }
    } else {
      T s = T(0);
      if (g < nrows)
        for (int j = s0 + lane; j < s1; j += L) s += prod[j];
      for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (g < nrows && lane == 0) row(ri, init ? init[ri] + s : s, acc);
    }
  }
}

template <typename T, typename Epi, typename VT = T>
__global__ __launch_bounds__(kSpmvBlock) void spmv_kernel(
    CsrView<VT> A, const T* __restrict__ x, Epi epi, double* partials,
    const T* init) {
  __shared__ T prod[kNnzBlock + 2];
  __shared__ double red[kSpmvBlock / 64];
  if (epi.done && *epi.done) return;
  double acc = 0.0;
  spmv_rows<1>(
      A, init, [] { return true; }, [x](int j) { return x[j]; },
      [&epi](int ri, T s, double* a) { a[0] = epi.row(ri, s); }, prod, red, &acc);
  if constexpr (Epi::kReduce) {
    const double t = block_sum_to_thread0<kSpmvBlock>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

// ------------------------------------------------------------- window tiles

constexpr int kXtThreads = 1024;
constexpr int kXtMaxWindow = 8192;
constexpr int kXtChunk = 7168;  // entries staged in LDS per pass

template <typename T>
struct XtView {
  const int64_t* __restrict__ base;
  const int32_t* __restrict__ off;
  const uint32_t* __restrict__ idx;
  const T* __restrict__ v;
  const uint32_t* __restrict__ skip;
  int64_t nrows, ncols, windows;
  int window;
};

// One workgroup per row group.  Per window: x's window and the window's
// entries (up to kXtChunk per pass) go to LDS with coalesced loads, then
// every thread runs through its own entries of the window (its S rows, slot
// by slot, columns ascending) with both served by LDS.  The loads of the
// next window's x, first entry chunk and thread ranges are issued into
// registers before the current window's arithmetic, and that arithmetic
// reads only registers and LDS, so the loads are in flight while it runs;
// the window's entry range reaches every thread through LDS.  A row's sum
// starts at 0 and adds a * x in column order with the product and the sum
// rounded separately, the reference's sparse_blas.hpp:26-31 arithmetic.  A
// thread's entries of a window come slot by slot, so it walks the slots in
// order and adds each run to its slot's accumulator.  Rows marked in `skip`
// (long rows) are left to the row kernel.
template <typename T, int S, typename Epi>
__global__ __launch_bounds__(kXtThreads) void xtile_spmv_kernel(
    XtView<T> A, const T* __restrict__ x, Epi epi, double* partials) {
#pragma clang fp contract(off)
  constexpr int XJ = kXtMaxWindow / kXtThreads;
  constexpr int EJ = kXtChunk / kXtThreads;
  static_assert(XJ * kXtThreads == kXtMaxWindow && EJ * kXtThreads == kXtChunk,
                "staging divides evenly");
  __shared__ T win[kXtMaxWindow];
  __shared__ T eval[kXtChunk];
  __shared__ uint32_t eidx[kXtChunk];
  __shared__ double red[kXtThreads / 64];
  __shared__ int range[2];
  if (epi.done && *epi.done) return;
  const int tid = threadIdx.x;
  const int64_t g = blockIdx.x;
  const int64_t gbase = A.base[g];
  const int gtot = (int)(A.base[g + 1] - gbase);  // the group's entries
  const int32_t* __restrict__ off = A.off + g * (A.windows * kXtThreads + 1);
  const uint32_t* __restrict__ idx = A.idx + gbase;
  const T* __restrict__ val = A.v + gbase;
  T acc[S];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s] = T(0);
  // the next window's x, thread range and first entry chunk (the entries
  // from the window's start on, up to kXtChunk of them), in registers
  T xr[XJ], vr[EJ];
  uint32_t ir[EJ];
  int nk0 = 0, nk1 = 0;
  auto fetch = [&](int64_t w, int nb) {
    const int64_t c0 = w * A.window;
    const int64_t left = A.ncols - c0;
    const int wl = left < A.window ? (int)left : A.window;
    nk0 = off[w * kXtThreads + tid];
    nk1 = off[w * kXtThreads + tid + 1];
    const int cend = gtot - nb < kXtChunk ? gtot : nb + kXtChunk;
#pragma unroll
    for (int j = 0; j < XJ; ++j) {
      const int k = tid + j * kXtThreads;
      xr[j] = k < wl ? x[c0 + k] : T(0);
    }
#pragma unroll
    for (int j = 0; j < EJ; ++j) {
      const int k = nb + tid + j * kXtThreads;
      const bool in = k < cend;
      ir[j] = in ? __builtin_nontemporal_load(idx + k) : 0u;
      vr[j] = in ? __builtin_nontemporal_load(val + k) : T(0);
    }
  };
  fetch(0, 0);
  for (int64_t w = 0; w < A.windows; ++w) {
    const int k0 = nk0, k1 = nk1;
    __syncthreads();  // every thread is done with the previous window
#pragma unroll
    for (int j = 0; j < XJ; ++j) win[tid + j * kXtThreads] = xr[j];
#pragma unroll
    for (int j = 0; j < EJ; ++j) {
      eidx[tid + j * kXtThreads] = ir[j];
      eval[tid + j * kXtThreads] = vr[j];
    }
    if (tid == 0) range[0] = k0;
    if (tid == kXtThreads - 1) range[1] = k1;
    __syncthreads();
    const int wb = range[0], we = range[1];  // block-uniform
    if (w + 1 < A.windows) fetch(w + 1, we);
    if (we > wb) {
      int cb = wb;
      for (;;) {
        const int ce = we - cb < kXtChunk ? we : cb + kXtChunk;
        const int a1 = k1 < ce ? k1 : ce;
        int k = k0 > cb ? k0 : cb;
        uint32_t e = k < a1 ? eidx[k - cb] : 0xFFFFFFFFu;  // slot 0xFFFF: none
#pragma unroll
        for (int s = 0; s < S; ++s) {
          while ((e >> 16) == (uint32_t)s) {
            acc[s] += eval[k - cb] * win[e & 0xFFFFu];
            ++k;
            e = k < a1 ? eidx[k - cb] : 0xFFFFFFFFu;
          }
        }
        if (ce == we) break;
        // a further chunk of this window (rare: the window is unusually full)
        cb = ce;
        const int ce2 = we - cb < kXtChunk ? we : cb + kXtChunk;
        __syncthreads();
        for (int kk = cb + tid; kk < ce2; kk += kXtThreads) {
          eidx[kk - cb] = __builtin_nontemporal_load(idx + kk);
          eval[kk - cb] = __builtin_nontemporal_load(val + kk);
        }
        __syncthreads();
      }
    }
  }
  const uint32_t skip = A.skip[g * kXtThreads + tid];
  double part = 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int64_t r = g * (int64_t)(kXtThreads * S) + tid + (int64_t)s * kXtThreads;
    if (r < A.nrows && !((skip >> s) & 1u)) part += epi.row((int)r, acc[s]);
  }
  if constexpr (Epi::kReduce) {
    const double t = block_sum_to_thread0<kXtThreads>(part, red);
    if (tid == 0) partials[g] = t;
  }
}

// the number of partials a tiled SpMV writes: one per group, then the long
// rows' blocks
inline int xtile_partials_count(const spk_csr_s& A) {
  return (int)A.xt_groups + (A.xt_long ? A.xt_long->nblocks : 0);
}

template <typename T, typename Epi>
void launch_xtiles(spk_csr_s& A, const T* x, const Epi& epi, double* parts,
                   hipStream_t s) {
  XtView<T> v{A.xt_base.p, A.xt_off.p, A.xt_idx.p, nullptr, A.xt_skip.p,
              A.nrows, A.ncols, A.xt_windows, A.xt_window};
  if constexpr (std::is_same<T, float>::value) v.v = A.tiles32();
  else v.v = A.xt_v64.p;
  const dim3 grid((unsigned)A.xt_groups), block(kXtThreads);
  if (A.xt_slots == 16)
    hipLaunchKernelGGL((xtile_spmv_kernel<T, 16, Epi>), grid, block, 0, s, v, x,
                       epi, parts);
  else
    hipLaunchKernelGGL((xtile_spmv_kernel<T, 8, Epi>), grid, block, 0, s, v, x,
                       epi, parts);
  SPK_HIP_TRY(hipGetLastError());
  if (A.xt_long && A.xt_long->nblocks > 0) {
    // the long rows: their store maps each stored row to its matrix row
    spk_csr_s& L = *A.xt_long;
    hipLaunchKernelGGL((spmv_kernel<T, Epi>), dim3(L.nblocks), dim3(kSpmvBlock), 0,
                       s, L.view<T>(), x, epi, parts ? parts + A.xt_groups : nullptr,
                       (const T*)nullptr);
    SPK_HIP_TRY(hipGetLastError());
  }
}

// the row kernel over one store: FP32 SpMVs of a store with half32 set
// read its FP16 values
template <typename T, typename Epi>
void launch_row_kernel(spk_csr_s& A, const T* x, const Epi& epi, double* parts,
                       const T* init, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (A.half32) {
      hipLaunchKernelGGL((spmv_kernel<float, Epi, _Float16>), dim3(A.nblocks),
                         dim3(kSpmvBlock), 0, s, A.view16(), x, epi, parts, init);
      SPK_HIP_TRY(hipGetLastError());
      return;
    }
  }
  hipLaunchKernelGGL((spmv_kernel<T, Epi>), dim3(A.nblocks), dim3(kSpmvBlock), 0,
                     s, A.view<T>(), x, epi, parts, init);
  SPK_HIP_TRY(hipGetLastError());
}

template <typename T, typename Epi>
void launch_spmv_one(spk_csr_s& A, const T* x, const Epi& epi, const T* init,
                     hipStream_t s) {
  const int nparts = A.xt_groups ? xtile_partials_count(A) : A.nblocks;
  double* parts = Epi::kReduce ? reduce_partials(nparts) : nullptr;
  if (A.xt_groups) launch_xtiles(A, x, epi, parts, s);
  else launch_row_kernel(A, x, epi, parts, init, s);
  if constexpr (Epi::kReduce) {
    hipLaunchKernelGGL((finish_kernel<1, EpiFinalize<Epi>>), dim3(1),
                       dim3(kFinishBlock), 0, s, (const double*)parts,
                       nparts, EpiFinalize<Epi>{epi}, epi.done);
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
  if (A.xt_groups) {
    launch_xtiles(A, x, epi, parts, s);
    return xtile_partials_count(A);
  }
  spk_csr_s& last = A.panels.empty() ? A : *A.panels.back();
  const T* init = nullptr;
  if (!A.panels.empty()) init = launch_panels(A, x, epi.done, s);
  launch_row_kernel(last, x, epi, parts, init, s);
  return last.nblocks;
}

// the number of partials launch_spmv_partials writes
inline int spmv_partials_count(const spk_csr_s& A) {
  if (A.xt_groups) return xtile_partials_count(A);
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
