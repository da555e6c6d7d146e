// Runtime of the library: device selection, the library stream, error reporting and the
// memory entry points of the C ABI.
#include <mutex>

#include "../../include/qed.h"
#include "qed_common.hpp"

namespace qed {

namespace {
thread_local std::string t_err;
std::once_flag g_once;
hipStream_t g_stream = nullptr;
int g_device = 0;
int g_cus = 256;
bool g_ready = false;
std::string g_init_err;

std::vector<uint32_t> make_binom() {
  std::vector<uint32_t> b(33 * 33, 0);
  for (int n = 0; n <= 32; ++n) {
    b[n * 33] = 1;
    for (int k = 1; k <= n; ++k) {
      uint64_t v = (uint64_t)b[(n - 1) * 33 + k - 1] + (k <= n - 1 ? b[(n - 1) * 33 + k] : 0);
      b[n * 33 + k] = (uint32_t)v;
    }
  }
  return b;
}

DevBuf* g_binom_dev = nullptr;
}  // namespace

void set_error(const std::string& msg) { t_err = msg; }

static void do_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    g_init_err = "no HIP device available (the MI355X library needs a GPU)";
    return;
  }
  if (device < 0 || device >= n) device = 0;
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_init_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return;
  }
  g_device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) g_cus = prop.multiProcessorCount;
  // default-priority stream (no priority changes)
  e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_init_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    return;
  }
  g_ready = true;
}

void ensure_init() {
  std::call_once(g_once, [] { do_init(0); });
  if (!g_ready) fail(g_init_err);
  // the calling thread may not have made the library's device current yet
  const hipError_t e = hipSetDevice(g_device);
  if (e != hipSuccess) fail(std::string("hipSetDevice: ") + hipGetErrorString(e));
}

hipStream_t stream() { return g_stream; }
int device_cu_count() { return g_cus; }

const std::vector<uint32_t>& binom_host() {
  static std::vector<uint32_t> b = make_binom();
  return b;
}

const uint32_t* binom_dev() {
  static std::mutex m;
  std::lock_guard<std::mutex> lk(m);
  if (!g_binom_dev) {
    g_binom_dev = new DevBuf(upload(binom_host()));
    QED_HIP(hipStreamSynchronize(stream()));
  }
  return g_binom_dev->as<uint32_t>();
}

}  // namespace qed

using namespace qed;

extern "C" {

int qed_init(int device) {
  try {
    std::call_once(g_once, [device] { do_init(device); });
    if (!g_ready) fail(g_init_err);
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 1;
  }
}

const char* qed_last_error(void) { return t_err.c_str(); }

int qed_sync(void) {
  return guard([] { QED_HIP(hipStreamSynchronize(stream())); });
}

int qed_alloc(void** ptr, uint64_t bytes) {
  return guard([&] {
    *ptr = nullptr;
    if (bytes) QED_HIP(hipMalloc(ptr, bytes));
  });
}

int qed_free(void* ptr) {
  return guard([&] {
    if (ptr) {
      QED_HIP(hipStreamSynchronize(stream()));
      QED_HIP(hipFree(ptr));
    }
  });
}

int qed_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  return guard([&] {
    QED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  return guard([&] {
    QED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_memcpy_d2d(void* dst, const void* src, uint64_t bytes) {
  return guard([&] { QED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream())); });
}

int qed_memset(void* ptr, int value, uint64_t bytes) {
  return guard([&] { QED_HIP(hipMemsetAsync(ptr, value, bytes, stream())); });
}

}  // extern "C"
