// spk_runtime.hip -- device selection, the library stream, device memory,
// timers, the workspace pool and the reduction scratch.
#include "spk_internal.hpp"

#include <cstring>
#include <memory>
#include <mutex>

namespace spk {

namespace {
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

void* Workspace::get(int slot, size_t bytes) {
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
