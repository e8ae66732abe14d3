// spk_rcm.hip -- reverse Cuthill-McKee on the device (reordering.cpp:19-203).
//
// The Cuthill-McKee order is a breadth-first order in which every node
// appends its unvisited neighbours sorted by degree (ties in column order,
// the order of a stable sort; the reference's std::sort keeps that order
// for up to 16 neighbours).  Level by level that is one sort: a node of
// the next level goes to the earliest node of the current level that has
// it as a neighbour, so the next level is the newly reached nodes sorted
// by (position of that earliest parent, degree, index).  Per level:
//   discover  one wavefront per current-level node; atomicMin of the
//             parent's position into each unvisited neighbour's claim; the
//             first claimant appends the neighbour to a candidate list
//   sort      64-bit keys (claim, degree, index), device radix sort
//   place     the sorted nodes extend the order; they become visited and
//             their claims return to the unclaimed value
// The start nodes follow the reference: a pseudo-peripheral node (repeated
// breadth-first sweeps from the first minimum-degree node) for the first
// component, the first unvisited node of minimum degree for the others
// (isolated nodes, taken in index order, are appended at once).  The
// result is reversed: perm[old] = new.
#include <hipcub/hipcub.hpp>

#include "spk_csr.hpp"

namespace spk {
namespace {

constexpr int32_t kUnclaimed = 0x7F7F7F7F;  // a memset of 0x7F bytes

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

// (i, j) and (j, i) for every entry: symmetrize_pattern's index list; a
// row's entries before position k take two slots each, its diagonal one
__global__ void sym_expand_kernel(int64_t n, const int32_t* __restrict__ rp,
                                  const int32_t* __restrict__ ci,
                                  const int32_t* __restrict__ off,
                                  int64_t* __restrict__ r, int64_t* __restrict__ c) {
  const int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const int a = rp[i], b = rp[i + 1];
  int lo = a, hi = b;  // the diagonal's position (columns are sorted)
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (ci[mid] < (int)i) lo = mid + 1; else hi = mid;
  }
  const int dpos = (lo < b && ci[lo] == (int)i) ? lo : -1;
  const int64_t o = off[i];
  for (int k = a + lane; k < b; k += 64) {
    const int j = ci[k];
    const int64_t at = o + 2 * (int64_t)(k - a) - ((dpos >= 0 && dpos < k) ? 1 : 0);
    r[at] = i;
    c[at] = j;
    if (j != (int)i) {
      r[at + 1] = j;
      c[at + 1] = i;
    }
  }
}

__global__ void sym_count_kernel(int64_t n, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ ci,
                                 int32_t* __restrict__ cnt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) c += ci[k] == (int)i ? 1 : 2;
  cnt[i] = c;
}

// adjacency_degree (reordering.cpp:40-56)
__global__ void degree_kernel(int64_t n, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ ci, int32_t* __restrict__ deg,
                              unsigned* __restrict__ maxdeg) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int d = 0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) d += ci[k] != (int)i;
  deg[i] = d;
  atomicMax(maxdeg, (unsigned)d);
}

// first unvisited node of minimum degree: min of (degree, index)
__global__ void argmin_kernel(int64_t n, const int32_t* __restrict__ deg,
                              const int32_t* __restrict__ visited,
                              unsigned long long* __restrict__ best) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || visited[i]) return;
  atomicMin(best, ((unsigned long long)(uint32_t)deg[i] << 32) | (uint64_t)i);
}

__global__ void start_kernel(int32_t v, int64_t at, int32_t* __restrict__ order,
                             int32_t* __restrict__ visited) {
  order[at] = v;
  visited[v] = 1;
}

__global__ void discover_kernel(int64_t lo, int64_t cnt, const int32_t* __restrict__ rp,
                                const int32_t* __restrict__ ci,
                                const int32_t* __restrict__ order,
                                const int32_t* __restrict__ visited,
                                int32_t* __restrict__ claim, int32_t* __restrict__ cand,
                                unsigned* __restrict__ ncand) {
  const int64_t k = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (k >= cnt) return;
  const int32_t pos = (int32_t)(lo + k);
  const int u = order[pos];
  for (int t = rp[u] + lane; t < rp[u + 1]; t += 64) {
    const int v = ci[t];
    if (v == u || visited[v]) continue;
    const int32_t old = atomicMin(claim + v, pos);
    if (old == kUnclaimed) cand[atomicAdd(ncand, 1u)] = v;
  }
}

__global__ void key_kernel(int64_t cnt, const int32_t* __restrict__ cand,
                           const int32_t* __restrict__ claim,
                           const int32_t* __restrict__ deg, int bd, int bv,
                           unsigned long long* __restrict__ keys) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int v = cand[i];
  keys[i] = ((unsigned long long)(uint32_t)claim[v] << (bd + bv)) |
            ((unsigned long long)(uint32_t)deg[v] << bv) | (unsigned long long)v;
}

__global__ void place_kernel(int64_t cnt, int64_t at, int bv,
                             const unsigned long long* __restrict__ keys,
                             int32_t* __restrict__ order, int32_t* __restrict__ visited,
                             int32_t* __restrict__ claim) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int v = (int)(keys[i] & ((1ull << bv) - 1));
  order[at + i] = v;
  visited[v] = 1;
  claim[v] = kUnclaimed;
}

__global__ void isolated_flags_kernel(int64_t n, const int32_t* __restrict__ deg,
                                      const int32_t* __restrict__ visited,
                                      char* __restrict__ flag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flag[i] = !visited[i] && deg[i] == 0;
}

__global__ void mark_kernel(int64_t cnt, const int32_t* __restrict__ list,
                            int32_t* __restrict__ visited) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < cnt) visited[list[i]] = 1;
}

__global__ void reverse_kernel(int64_t n, const int32_t* __restrict__ order,
                               int64_t* __restrict__ perm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) perm[order[i]] = n - 1 - i;
}

int bits_for(uint64_t v) {  // bits to hold values 0..v
  int b = 1;
  while (b < 63 && (1ull << b) <= v) ++b;
  return b;
}

struct Rcm {
  spk_csr_s* S = nullptr;  // symmetrised pattern
  int64_t n = 0;
  int bd = 1, bv = 1;
  DevBuf<int32_t> deg, visited, claim, order, cand, ilist;
  DevBuf<unsigned> counter;
  DevBuf<unsigned long long> keys, keys2, best;
  DevBuf<unsigned char> tmp;
  DevBuf<char> flags;
  std::vector<int32_t> hdeg;
  ~Rcm() { delete S; }

  void setup(spk_csr_s& A) {
    SPK_REQUIRE(A.nrows == A.ncols, SPK_INVALID_ARGUMENT,
                "reordering needs a square matrix");
    SPK_REQUIRE(2 * A.nnz < INT32_MAX, SPK_INVALID_ARGUMENT,
                "the symmetrised pattern exceeds int32 indexing");
    n = A.nrows;
    hipStream_t s = stream();
    // symmetrize_pattern (reordering.cpp:19-38) by device assembly
    DevBuf<int32_t> cnt((size_t)n), off((size_t)n);
    hipLaunchKernelGGL(sym_count_kernel, dim3(nb(n)), dim3(256), 0, s, n,
                       (const int32_t*)A.rp.p, (const int32_t*)A.ci.p, cnt.p);
    size_t tb = 0;
    SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.p, off.p, (int)n, s));
    tmp.reserve(tb);
    SPK_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.p, off.p, (int)n, s));
    int32_t lo = 0, lc = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&lo, off.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipMemcpyAsync(&lc, cnt.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    const int64_t m = (int64_t)lo + lc;
    DevBuf<int64_t> r((size_t)(m > 0 ? m : 1)), c((size_t)(m > 0 ? m : 1));
    hipLaunchKernelGGL(sym_expand_kernel, dim3(nb(n * 64)), dim3(256), 0, s, n,
                       (const int32_t*)A.rp.p, (const int32_t*)A.ci.p,
                       (const int32_t*)off.p, r.p, c.p);
    SPK_HIP_TRY(hipGetLastError());
    spk_csr out = nullptr;
    const spk_status rc = spk_csr_assemble(n, n, m, r.p, c.p, nullptr, &out);
    if (rc != SPK_OK) throw Error{rc, spk_last_error()};
    S = out;
    deg.alloc((size_t)n);
    visited.alloc((size_t)n);
    claim.alloc((size_t)n);
    order.alloc((size_t)n);
    cand.alloc((size_t)n);
    keys.alloc((size_t)n);
    keys2.alloc((size_t)n);
    counter.alloc(2);
    best.alloc(1);
    SPK_HIP_TRY(hipMemsetAsync(counter.p, 0, 8, s));
    hipLaunchKernelGGL(degree_kernel, dim3(nb(n)), dim3(256), 0, s, n,
                       (const int32_t*)S->rp.p, (const int32_t*)S->ci.p, deg.p,
                       counter.p + 1);
    SPK_HIP_TRY(hipGetLastError());
    unsigned md = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&md, counter.p + 1, 4, hipMemcpyDeviceToHost, s));
    hdeg.resize((size_t)n);
    SPK_HIP_TRY(hipMemcpyAsync(hdeg.data(), deg.p, n * 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    bd = bits_for(md);
    bv = bits_for((uint64_t)n);
    SPK_REQUIRE(bd + 2 * bv <= 64, SPK_INVALID_ARGUMENT,
                "matrix too large for 64-bit ordering keys");
    SPK_HIP_TRY(hipMemsetAsync(claim.p, 0x7F, n * 4, s));
    size_t sb = 0;
    SPK_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, sb, keys.p, keys2.p,
                                                  (int)n, 0, 2 * bv + bd, s));
    tmp.reserve(sb);
  }

  void reset_visited() {
    SPK_HIP_TRY(hipMemsetAsync(visited.p, 0, n * 4, stream()));
  }

  // Cuthill-McKee sweep from v, appending at position at; returns the end
  // position and the start of the last level
  int64_t sweep(int32_t v, int64_t at, int64_t* last_lo) {
    hipStream_t s = stream();
    hipLaunchKernelGGL(start_kernel, dim3(1), dim3(1), 0, s, v, at, order.p, visited.p);
    int64_t lo = at, hi = at + 1;
    *last_lo = lo;
    level_count = 1;
    for (;;) {
      SPK_HIP_TRY(hipMemsetAsync(counter.p, 0, 4, s));
      hipLaunchKernelGGL(discover_kernel, dim3(nb((hi - lo) * 64)), dim3(256), 0, s,
                         lo, hi - lo, (const int32_t*)S->rp.p, (const int32_t*)S->ci.p,
                         (const int32_t*)order.p, (const int32_t*)visited.p,
                         claim.p, cand.p, counter.p);
      SPK_HIP_TRY(hipGetLastError());
      unsigned cnt = 0;
      SPK_HIP_TRY(hipMemcpyAsync(&cnt, counter.p, 4, hipMemcpyDeviceToHost, s));
      SPK_HIP_TRY(hipStreamSynchronize(s));
      if (cnt == 0) return hi;
      hipLaunchKernelGGL(key_kernel, dim3(nb(cnt)), dim3(256), 0, s, (int64_t)cnt,
                         (const int32_t*)cand.p, (const int32_t*)claim.p,
                         (const int32_t*)deg.p, bd, bv, keys.p);
      size_t sb = tmp.n;
      SPK_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, sb, keys.p, keys2.p,
                                                    (int)cnt, 0, 2 * bv + bd, s));
      hipLaunchKernelGGL(place_kernel, dim3(nb(cnt)), dim3(256), 0, s, (int64_t)cnt,
                         hi, bv, (const unsigned long long*)keys2.p, order.p,
                         visited.p, claim.p);
      SPK_HIP_TRY(hipGetLastError());
      lo = hi;
      hi += cnt;
      *last_lo = lo;
      ++level_count;
    }
  }

  // every unvisited node of degree zero, in index order, at position at
  int64_t append_isolated(int64_t at) {
    hipStream_t s = stream();
    if (flags.n < (size_t)n) flags.alloc((size_t)n);
    if (ilist.n < (size_t)n) ilist.alloc((size_t)n);
    hipLaunchKernelGGL(isolated_flags_kernel, dim3(nb(n)), dim3(256), 0, s, n,
                       (const int32_t*)deg.p, (const int32_t*)visited.p, flags.p);
    hipcub::CountingInputIterator<int32_t> idx(0);
    size_t tb = 0;
    SPK_HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, idx, flags.p, order.p + at,
                                              counter.p, (int)n, s));
    tmp.reserve(tb);
    SPK_HIP_TRY(hipcub::DeviceSelect::Flagged(tmp.p, tb, idx, flags.p, order.p + at,
                                              counter.p, (int)n, s));
    unsigned cnt = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&cnt, counter.p, 4, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    hipLaunchKernelGGL(mark_kernel, dim3(nb(cnt > 0 ? cnt : 1)), dim3(256), 0, s,
                       (int64_t)cnt, (const int32_t*)(order.p + at), visited.p);
    SPK_HIP_TRY(hipGetLastError());
    return at + cnt;
  }

  int32_t argmin_unvisited() {
    hipStream_t s = stream();
    SPK_HIP_TRY(hipMemsetAsync(best.p, 0xFF, 8, s));
    hipLaunchKernelGGL(argmin_kernel, dim3(nb(n)), dim3(256), 0, s, n,
                       (const int32_t*)deg.p, (const int32_t*)visited.p, best.p);
    unsigned long long b = 0;
    SPK_HIP_TRY(hipMemcpyAsync(&b, best.p, 8, hipMemcpyDeviceToHost, s));
    SPK_HIP_TRY(hipStreamSynchronize(s));
    return (int32_t)(b & 0xFFFFFFFFull);
  }

  // pseudo_peripheral_node (reordering.cpp:112-150)
  int32_t pseudo_peripheral() {
    reset_visited();
    int32_t start = argmin_unvisited();
    auto levels = [&](int32_t v, int64_t* ecc, std::vector<int32_t>* last) {
      reset_visited();
      int64_t last_lo = 0;
      const int64_t end = sweep(v, 0, &last_lo);
      // eccentricity: the number of levels after the first
      last->resize((size_t)(end - last_lo));
      SPK_HIP_TRY(hipMemcpyAsync(last->data(), order.p + last_lo,
                                 (end - last_lo) * 4, hipMemcpyDeviceToHost, stream()));
      SPK_HIP_TRY(hipStreamSynchronize(stream()));
      *ecc = level_count - 1;
    };
    int64_t ecc = 0;
    std::vector<int32_t> last;
    levels(start, &ecc, &last);
    for (;;) {
      int32_t cand = last[0];
      for (size_t k = 1; k < last.size(); ++k)
        if (hdeg[(size_t)last[k]] < hdeg[(size_t)cand]) cand = last[k];
      int64_t new_ecc = 0;
      std::vector<int32_t> new_last;
      levels(cand, &new_ecc, &new_last);
      if (new_ecc <= ecc) return start;
      start = cand;
      ecc = new_ecc;
      last.swap(new_last);
    }
  }
  int64_t level_count = 0;
};

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_csr_pseudo_peripheral_node(spk_csr A, int64_t* node) {
  return guarded([&] {
    SPK_REQUIRE(A && node, SPK_INVALID_ARGUMENT, "null argument");
    Rcm R;
    R.setup(*A);
    *node = R.pseudo_peripheral();
  });
}

// reverse_cuthill_mckee (reordering.cpp:152-203)
extern "C" spk_status spk_csr_rcm(spk_csr A, int64_t* d_perm) {
  return guarded([&] {
    TraceRange trace_("spk_csr_rcm");
    SPK_REQUIRE(A && d_perm, SPK_INVALID_ARGUMENT, "null argument");
    Rcm R;
    R.setup(*A);
    const int64_t n = R.n;
    const int32_t s0 = R.pseudo_peripheral();
    R.reset_visited();
    int64_t last = 0;
    int64_t total = R.sweep(s0, 0, &last);
    while (total < n) {
      const int32_t v = R.argmin_unvisited();
      if (R.hdeg[(size_t)v] == 0) total = R.append_isolated(total);
      else total = R.sweep(v, total, &last);
    }
    hipLaunchKernelGGL(reverse_kernel, dim3(nb(n)), dim3(256), 0, stream(), n,
                       (const int32_t*)R.order.p, d_perm);
    SPK_HIP_TRY(hipGetLastError());
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
  });
}
