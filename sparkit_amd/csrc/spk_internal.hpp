// spk_internal.hpp -- shared runtime pieces of the sparkit MI355X library:
// status/error plumbing, the device context (stream, reduction scratch),
// wavefront/block reductions and the XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "sparkit_hip.h"

namespace spk {

// ------------------------------------------------------------- tracing

// A named roctx range for the span of a scope: the solvers mark their
// entry points and phases (setup, restart cycles, refinement steps,
// reliable updates, the learned solve's check / solve / fallback) so that
// `rocprofv3 --marker-trace` places the kernels of a trace under them.
// Without a tool attached a range costs a call that returns at once.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

// ------------------------------------------------------------- errors

void set_error(const std::string& msg);

struct Error {
  spk_status status;
  std::string msg;
};

#define SPK_HIP_TRY(expr)                                                  \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      throw ::spk::Error{SPK_HIP_ERROR,                                    \
                         std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define SPK_REQUIRE(cond, status, msg)                     \
  do {                                                     \
    if (!(cond)) throw ::spk::Error{(status), (msg)};      \
  } while (0)

// Runs body, converting exceptions into a status + last-error string.
template <typename F>
spk_status guarded(F&& body) {
  try {
    body();
    return SPK_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return SPK_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SPK_INVALID_ARGUMENT;
  }
}

// ------------------------------------------------------------- device

// Device memory owned by the library (RAII, move-only).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{SPK_OUT_OF_MEMORY, "hipMalloc of " +
                  std::to_string(count * sizeof(T)) + " bytes failed"};
    }
    n = count;
  }
  // grow-only reallocation (contents are not preserved)
  void reserve(size_t count) { if (count > n || p == nullptr) alloc(count); }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  T* get() const { return p; }
};

// Per-process device context: the stream every launch goes on and the
// scratch used by grid-wide deterministic reductions.
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cus = 0;
};

Context& ctx();             // initialises device 0 on first use
hipStream_t stream();       // current library stream

// Grid-wide reductions write one partial per (logical) block into a shared,
// grow-only scratch array.  Kernels on the library stream run one after
// another, so a single array serves every reduction site.
double* reduce_partials(size_t count);

// ------------------------------------------------------------- launch

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------- reductions

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a 256-thread block; result valid in every thread.  Fixed
// reduction tree, so results are deterministic.
template <typename T, int BLOCK>
__device__ __forceinline__ T block_reduce_sum(T v, T* smem /* BLOCK/64 */) {
  constexpr int W = BLOCK / 64;
  v = wave_reduce_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T s = smem[0];
#pragma unroll
  for (int i = 1; i < W; ++i) s += smem[i];
  return s;
}

// Wavefront sum with DPP lane moves (quad swaps, half-row and row mirrors,
// row broadcasts) instead of LDS permutes: a fixed order, the total in lane
// 63 (other lanes hold partial sums).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xF>(v);  // row_mirror: every lane holds its row's sum
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Sum over a block of BLOCK threads, valid in thread 0 only (fixed order).
template <int BLOCK>
__device__ __forceinline__ double block_sum_to_thread0(double v, double* smem) {
  constexpr int W = BLOCK / 64;
  v = wave_sum_dpp(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    s = smem[0];
#pragma unroll
    for (int i = 1; i < W; ++i) s += smem[i];
  }
  return s;
}

// Coherent load of a value another block may have written during the same
// kernel (different XCDs do not share an L2).
__device__ __forceinline__ double load_partial(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide reductions are two-phase: every block of the producing kernel
// stores its partial sums with plain stores, and a one-block finisher kernel
// (launched right after, so the kernel boundary orders the stores) adds them
// in index order -- deterministic, and free of per-block release fences,
// which on this architecture write back the L2 and cost far more than the
// extra launch.
constexpr int kFinishBlock = 1024;

// ------------------------------------------------------------- workspace

// Grow-only pool of device vectors reused across solver calls so the timed
// region of a solve performs no hipMalloc.
struct Workspace {
  std::vector<DevBuf<unsigned char>> bufs;
  void* get(int slot, size_t bytes);
};
Workspace& workspace();

// ------------------------------------------------------------- scratch arena
//
// While an ArenaScope is alive, Workspace::get and reduce_partials carve
// their buffers out of one device region given to the scope instead of
// allocating: each workspace slot gets a 256-byte-aligned piece on first
// use, the reduction partials one piece sized up front.  A request that
// does not fit throws SPK_INVALID_ARGUMENT; nothing is ever freed (the
// region belongs to the caller).  Host-side bookkeeping is a fixed array.
constexpr int kArenaSlots = 160;
constexpr size_t kArenaAlign = 256;
inline size_t arena_round(size_t b) {
  return (b + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
}
class ArenaScope {
 public:
  // partials: room reserved for reduce_partials (doubles)
  ArenaScope(void* base, size_t bytes, size_t partials);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;
};
// The library's static scratch array on the current device: its address
// (looked up with hipGetSymbolAddress once per device, when the library
// initialises that device) and its size.
void* scratch_ptr();
size_t scratch_bytes();

}  // namespace spk
