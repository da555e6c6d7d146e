// Shared runtime pieces of the MI355X ED library: error handling, the library stream,
// device buffers, wave64 reductions and the combinatorial-number-system tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace qed {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] inline void fail(const std::string& msg) { throw Error(msg); }

#define QED_HIP(call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      ::qed::fail(std::string(#call) + " failed: " + hipGetErrorString(e_));               \
  } while (0)

void ensure_init();
hipStream_t stream();
void set_error(const std::string& msg);
int device_cu_count();

// Runs f and converts exceptions into the C ABI's status code + thread-local message.
template <class F>
int guard(F&& f) {
  try {
    ensure_init();
    f();
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 1;
  }
}

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: a view into memory owned elsewhere
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), owned(o.owned) {
    o.p = nullptr;
    o.bytes = 0;
    o.owned = true;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      owned = o.owned;
      o.p = nullptr;
      o.bytes = 0;
      o.owned = true;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  static DevBuf view(void* ptr, size_t n) {
    DevBuf b;
    b.p = ptr;
    b.bytes = n;
    b.owned = false;
    return b;
  }
  void alloc(size_t n) {
    release();
    if (n) QED_HIP(hipMalloc(&p, n));
    bytes = n;
    owned = true;
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

template <class T>
DevBuf upload(const std::vector<T>& v) {
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty())
    QED_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream()));
  return b;
}

// Binomial coefficients C(n, k) for n, k <= 32 (C(32,16) < 2^32), host copy.
const std::vector<uint32_t>& binom_host();
// Device copy (33 x 33 row-major, C(n,k) at n*33+k).
const uint32_t* binom_dev();

// ---- device helpers ---------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum; every thread gets the result. scratch: >= blockDim/64 doubles in LDS.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += scratch[i];
  return t;
}

// Colex (= ascending-integer) rank of a half state by the combinatorial number system:
// rank = sum_k C(p_k, k+1) over the set bits p_0 < p_1 < ... (binom: 33 x 33 table).
__device__ __forceinline__ uint32_t cns_rank(uint32_t s, const uint32_t* binom) {
  uint32_t r = 0;
  int k = 1;
  while (s) {
    int p = __ffs(s) - 1;
    r += binom[p * 33 + k];
    ++k;
    s &= s - 1;
  }
  return r;
}

// Inverse of cns_rank for N particles on L orbitals.
__device__ __forceinline__ uint32_t cns_unrank(uint32_t r, int L, int N, const uint32_t* binom) {
  uint32_t s = 0;
  int p = L - 1;
  for (int k = N; k >= 1; --k) {
    while (binom[p * 33 + k] > r) --p;
    s |= 1u << p;
    r -= binom[p * 33 + k];
    --p;
  }
  return s;
}

}  // namespace qed
