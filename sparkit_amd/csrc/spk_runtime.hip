// spk_runtime.hip -- device selection, the library stream, device memory,
// timers, the workspace pool and the reduction scratch.
#include "spk_internal.hpp"

#include <cstring>
#include <memory>
#include <mutex>

namespace spk {

// Device scratch for the entry points that allocate nothing
// (spk_cg_solve_noalloc).  In host code `scratch` names a host-side stand-in,
// not device memory: every host use goes through the address
// hipGetSymbolAddress returned for the current device (g_scratch_dev).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};

struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* slot_ptr[kArenaSlots] = {};
  size_t slot_bytes[kArenaSlots] = {};
  double* partials = nullptr;
  size_t partials_count = 0;
};
Arena* g_arena = nullptr;

void* arena_carve(Arena& a, size_t bytes) {
  const size_t need = arena_round(bytes > 0 ? bytes : 1);
  if (need > a.cap - a.used)
    throw Error{SPK_INVALID_ARGUMENT,
                "no-allocation scratch exhausted: " + std::to_string(a.used) +
                    " + " + std::to_string(need) + " bytes > " +
                    std::to_string(a.cap)};
  void* p = a.base + a.used;
  a.used += need;
  return p;
}

thread_local std::string g_last_error;
Context g_ctx;
std::mutex g_init_mutex;
hipStream_t g_user_stream = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
std::unique_ptr<Workspace> g_ws;
DevBuf<double>* g_partials = nullptr;

void init_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw Error{SPK_NO_DEVICE, "no HIP device available"};
  SPK_REQUIRE(device >= 0 && device < count, SPK_INVALID_ARGUMENT,
              "device index out of range");
  SPK_HIP_TRY(hipSetDevice(device));
  SPK_REQUIRE(device < kMaxDevices, SPK_INVALID_ARGUMENT, "device index too large");
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    SPK_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = p;
  }
  hipDeviceProp_t prop;
  SPK_HIP_TRY(hipGetDeviceProperties(&prop, device));
  g_ctx.device = device;
  g_ctx.num_cus = prop.multiProcessorCount;
  SPK_HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  SPK_HIP_TRY(hipEventCreate(&g_ev0));
  SPK_HIP_TRY(hipEventCreate(&g_ev1));
  g_partials = new DevBuf<double>(1 << 16);
  g_ws.reset(new Workspace());
}
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

Context& ctx() {
  if (g_ctx.device < 0) {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if (g_ctx.device < 0) init_device(0);
  }
  return g_ctx;
}

hipStream_t stream() {
  Context& c = ctx();
  return g_user_stream ? g_user_stream : c.stream;
}

void* scratch_ptr() {
  Context& c = ctx();
  void* p = g_scratch_dev[c.device];
  SPK_REQUIRE(p, SPK_HIP_ERROR, "scratch address unknown for this device");
  return p;
}

size_t scratch_bytes() { return kScratchBytes; }

ArenaScope::ArenaScope(void* base, size_t bytes, size_t partials) {
  SPK_REQUIRE(!g_arena, SPK_INVALID_ARGUMENT, "nested no-allocation calls");
  static Arena arena;  // fixed host bookkeeping, reset per scope
  arena = Arena{};
  arena.base = (char*)base;
  arena.cap = bytes;
  arena.partials = (double*)arena_carve(arena, partials * sizeof(double));
  arena.partials_count = partials;
  g_arena = &arena;
}

ArenaScope::~ArenaScope() { g_arena = nullptr; }

void* Workspace::get(int slot, size_t bytes) {
  if (g_arena) {
    Arena& a = *g_arena;
    SPK_REQUIRE(slot >= 0 && slot < kArenaSlots, SPK_INVALID_ARGUMENT,
                "workspace slot out of range");
    if (!a.slot_ptr[slot] || a.slot_bytes[slot] < bytes) {
      a.slot_ptr[slot] = arena_carve(a, bytes);
      a.slot_bytes[slot] = bytes;
    }
    return a.slot_ptr[slot];
  }
  if ((size_t)slot >= bufs.size()) bufs.resize(slot + 1);
  bufs[slot].reserve(bytes);
  return bufs[slot].p;
}

Workspace& workspace() {
  ctx();
  return *g_ws;
}

double* reduce_partials(size_t count) {
  ctx();
  if (g_arena) {
    SPK_REQUIRE(count <= g_arena->partials_count, SPK_INVALID_ARGUMENT,
                "no-allocation scratch: reduction partials exceed the reserve");
    return g_arena->partials;
  }
  if (count > g_partials->n) {
    // the old array may still be read by queued kernels
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    g_partials->alloc(count + count / 4);
  }
  return g_partials->p;
}


}  // namespace spk

using namespace spk;

extern "C" {

const char* spk_last_error(void) { return g_last_error.c_str(); }

spk_status spk_init(int device) {
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if (g_ctx.device < 0) init_device(device);
    else SPK_REQUIRE(g_ctx.device == device, SPK_INVALID_ARGUMENT,
                     "library already initialised on another device");
  });
}

spk_status spk_finalize(void) {
  return guarded([&] {
    if (g_ctx.device < 0) return;
    SPK_HIP_TRY(hipDeviceSynchronize());
    g_ws.reset();
    delete g_partials;
    g_partials = nullptr;
    (void)hipEventDestroy(g_ev0);
    (void)hipEventDestroy(g_ev1);
    (void)hipStreamDestroy(g_ctx.stream);
    g_ctx = Context{};
  });
}

spk_status spk_set_stream(void* s) {
  return guarded([&] {
    ctx();
    g_user_stream = (hipStream_t)s;
  });
}

void* spk_get_stream(void) {
  void* s = nullptr;
  guarded([&] { s = (void*)stream(); });
  return s;
}

spk_status spk_synchronize(void) {
  return guarded([&] { SPK_HIP_TRY(hipStreamSynchronize(stream())); });
}

spk_status spk_device_info(int* num_cus, char* name, size_t name_len) {
  return guarded([&] {
    Context& c = ctx();
    if (num_cus) *num_cus = c.num_cus;
    if (name && name_len) {
      hipDeviceProp_t prop;
      SPK_HIP_TRY(hipGetDeviceProperties(&prop, c.device));
      std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
  });
}

spk_status spk_malloc(void** dptr, size_t bytes) {
  return guarded([&] {
    ctx();
    SPK_REQUIRE(dptr, SPK_INVALID_ARGUMENT, "null output pointer");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess)
      throw Error{SPK_OUT_OF_MEMORY, "hipMalloc failed: " +
                  std::string(hipGetErrorString(e))};
  });
}

spk_status spk_free(void* dptr) {
  return guarded([&] {
    ctx();
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    SPK_HIP_TRY(hipFree(dptr));
  });
}

spk_status spk_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    SPK_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
  });
}

spk_status spk_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    SPK_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
  });
}

spk_status spk_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    SPK_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  });
}

spk_status spk_memset(void* dst, int value, size_t bytes) {
  return guarded([&] { SPK_HIP_TRY(hipMemsetAsync(dst, value, bytes, stream())); });
}

spk_status spk_timer_start(void) {
  return guarded([&] {
    ctx();
    SPK_HIP_TRY(hipEventRecord(g_ev0, stream()));
  });
}

spk_status spk_timer_stop(float* ms) {
  return guarded([&] {
    ctx();
    SPK_HIP_TRY(hipEventRecord(g_ev1, stream()));
    SPK_HIP_TRY(hipEventSynchronize(g_ev1));
    SPK_HIP_TRY(hipEventElapsedTime(ms, g_ev0, g_ev1));
  });
}

}  // extern "C"
