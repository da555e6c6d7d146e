// Batched Krylov solvers: many sectors (cluster instances x Hilbert-space sectors) advance
// together, so that every step of the recurrence is one launch per operation for all of
// them instead of one per sector.
//
//  * Ground state: Lanczos for all sectors in lockstep. Each iteration is a fixed kernel
//    sequence (batched H.v, fused dot, fused axpy+norm) driven by a device-side
//    iteration counter; two iterations (both parities of the vector rotation) are
//    captured once in a hipGraph and replayed chunk by chunk. The recurrence scalars stay
//    on the device; after each chunk the host reads them, applies the stopping test of
//    the scalar Lanczos (|beta_j e_j| <= accur, src_ed/Hamiltonian/Lanczos.hpp) to every
//    new step, and retires the converged sectors. The ground-state vector is then
//    rebuilt for all sectors by a batched replay of the recurrence.
//  * Green function: Freund's band Lanczos (src_ed/Hamiltonian/Lanczos.hpp:214-395) for
//    all (instance, particle/hole) sectors. The host keeps each sector's band matrix and
//    deflation logic; one synchronization per step serves all sectors, and every vector
//    operation of the step is a single batched launch over the sectors.
//
// Descriptor-based kernels: each sector ("job") has a descriptor with its vectors and a
// block range; block -> job by binary search. Reductions are two-pass and deterministic.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/qed.h"
#include "qed_common.hpp"
#include "qed_linalg.hpp"
#include "qed_sector.hpp"

namespace qed {
namespace {

constexpr int kBT = 256;               // threads per block
constexpr uint64_t kChunk = kBT * 8;   // elements per block
constexpr int kBM = 20;                // vectors per multi-operation

__host__ __device__ inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kChunk - 1) / kChunk); }

template <class D>
__device__ __forceinline__ int find_job(const D* __restrict__ d, int nj, uint32_t b) {
  const uint32_t* b0 = reinterpret_cast<const uint32_t*>(d + nj);  // compact first blocks
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// ---- multi-dot: out[q] = <x, y[q]> -------------------------------------------------------
struct MDot {
  const double* x;
  const double* y[kBM];
  double* out;  // device, m values
  int m;
  uint64_t n;
  uint32_t blk0;
};

__global__ __launch_bounds__(kBT) void k_mdot_part(const MDot* __restrict__ d, int nj,
                                                   double* __restrict__ part) {
  __shared__ double red[kBT / 64];
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const MDot& D = d[j];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  const int m = D.m;
  double acc[kBM];
#pragma unroll
  for (int q = 0; q < kBM; ++q) acc[q] = 0.0;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) {  // x read once
    const double xv = D.x[i];
#pragma unroll
    for (int q = 0; q < kBM; ++q)
      if (q < m) acc[q] += xv * D.y[q][i];
  }
#pragma unroll
  for (int q = 0; q < kBM; ++q) {
    if (q < m) {  // uniform across the block
      const double s = block_sum(acc[q], red);
      if (threadIdx.x == 0) part[(size_t)b * kBM + q] = s;
    }
  }
}

// one wavefront per (job, q): lane-strided partial sums, then a fixed-order wave sum
__global__ void k_mdot_final(const MDot* __restrict__ d, int nj, const double* __restrict__ part) {
  const int id = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (id >= nj * kBM) return;
  const int j = id / kBM, q = id % kBM;
  const MDot& D = d[j];
  if (q >= D.m) return;
  const uint32_t nb = nblocks(D.n);
  double s = 0.0;
  for (uint32_t b = lane; b < nb; b += 64) s += part[(size_t)(D.blk0 + b) * kBM + q];
  s = wave_sum(s);
  if (lane == 0) D.out[q] = s;
}

// ---- multi-axpy: y += sgn * sum_q (*c[q]) x[q]; optionally y *= scale first -------------
struct MAxpy {
  double* y;
  const double* x[kBM];
  const double* c[kBM];  // device scalars
  double sgn;
  double scale;  // applied to y before the sum
  int m;
  uint64_t n;
  uint32_t blk0;
};

__global__ __launch_bounds__(kBT) void k_maxpy(const MAxpy* __restrict__ d, int nj) {
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const MAxpy& D = d[j];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  const int m = D.m;
  double c[kBM];
#pragma unroll
  for (int q = 0; q < kBM; ++q) c[q] = (q < m) ? D.sgn * *D.c[q] : 0.0;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) {
    double v = D.y[i] * D.scale;
#pragma unroll
    for (int q = 0; q < kBM; ++q)
      if (q < m) v += c[q] * D.x[q][i];
    D.y[i] = v;
  }
}

// ---- reverse multi-axpy: y[q] -= (*c[q]) x ----------------------------------------------
struct MRev {
  const double* x;
  double* y[kBM];
  const double* c[kBM];
  int m;
  uint64_t n;
  uint32_t blk0;
};

__global__ __launch_bounds__(kBT) void k_mrev(const MRev* __restrict__ d, int nj) {
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const MRev& D = d[j];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  const int m = D.m;
  double c[kBM];
#pragma unroll
  for (int q = 0; q < kBM; ++q) c[q] = (q < m) ? *D.c[q] : 0.0;
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) {  // x read once
    const double xv = D.x[i];
#pragma unroll
    for (int q = 0; q < kBM; ++q)
      if (q < m) D.y[q][i] -= c[q] * xv;
  }
}

// ---- fill: x = 0 ------------------------------------------------------------------------
struct VFill {
  double* x;
  uint64_t n;
  uint32_t blk0;
};

__global__ __launch_bounds__(kBT) void k_fill0(const VFill* __restrict__ d, int nj) {
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const VFill& D = d[j];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) D.x[i] = 0.0;
}

// ---- Lanczos (ground state) ------------------------------------------------------------
// vectors A, B per sector; at iteration k (parity p = k & 1) r = p ? B : A, q = p ? A : B.
struct LJob {
  double* v[2];
  double* alpha;  // per-iteration scalars (device)
  double* beta;
  uint64_t n;
  uint32_t blk0;
};

// k > 0: r /= beta[k-1]; q *= -beta[k-1]
__global__ __launch_bounds__(kBT) void k_lz_scal(const LJob* __restrict__ d, int nj, int parity,
                                                 const int* __restrict__ kdev,
                                                 const int* __restrict__ active) {
  const int k = *kdev;
  if (k == 0) return;
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  if (!active[j]) return;
  const LJob& D = d[j];
  double* r = D.v[parity];
  double* q = D.v[parity ^ 1];
  const double bt = D.beta[k - 1];
  // a breakdown (beta = 0: the Krylov space is exhausted) leaves zero vectors, not NaN
  const double inv = bt != 0.0 ? 1.0 / bt : 0.0;
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) {
    r[i] *= inv;
    q[i] *= -bt;
  }
}

// partials of <r, q> (r = v_k, q = H v_k - beta v_{k-1})
__global__ __launch_bounds__(kBT) void k_lz_dot(const LJob* __restrict__ d, int nj, int parity,
                                                const int* __restrict__ active, double* __restrict__ part) {
  __shared__ double red[kBT / 64];
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const LJob& D = d[j];
  double s = 0.0;
  if (active[j]) {
    const double* r = D.v[parity];
    const double* q = D.v[parity ^ 1];
    const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
    const uint64_t i1 = min(i0 + kChunk, D.n);
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) s += r[i] * q[i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[b] = s;
}

// sum of one job's block partials: one wavefront per job, lane-strided, fixed-order wave sum
__device__ __forceinline__ double job_partial_sum(const LJob& D, const double* __restrict__ part, int lane) {
  const uint32_t nb = nblocks(D.n);
  double s = 0.0;
  for (uint32_t b = lane; b < nb; b += 64) s += part[D.blk0 + b];
  return wave_sum(s);
}

// alpha[k] = sum of the partials
__global__ void k_lz_alpha(const LJob* __restrict__ d, int nj, const int* __restrict__ kdev,
                           const int* __restrict__ active, const double* __restrict__ part) {
  const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nj || !active[j]) return;
  const LJob& D = d[j];
  const double s = job_partial_sum(D, part, lane);
  if (lane == 0) D.alpha[*kdev] = s;
}

// q -= alpha[k] r, partials of |q|^2
__global__ __launch_bounds__(kBT) void k_lz_axpy(const LJob* __restrict__ d, int nj, int parity,
                                                 const int* __restrict__ kdev,
                                                 const int* __restrict__ active,
                                                 double* __restrict__ part) {
  __shared__ double red[kBT / 64];
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const LJob& D = d[j];
  double s = 0.0;
  if (active[j]) {
    const double a = D.alpha[*kdev];
    const double* r = D.v[parity];
    double* q = D.v[parity ^ 1];
    const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
    const uint64_t i1 = min(i0 + kChunk, D.n);
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) {
      const double v = q[i] - a * r[i];
      q[i] = v;
      s += v * v;
    }
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[b] = s;
}

// beta[k] = sqrt(sum of partials)
__global__ void k_lz_beta(const LJob* __restrict__ d, int nj, const int* __restrict__ kdev,
                          const int* __restrict__ active, const double* __restrict__ part) {
  const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nj || !active[j]) return;
  const LJob& D = d[j];
  const double s = job_partial_sum(D, part, lane);
  if (lane == 0) D.beta[*kdev] = std::sqrt(s);
}

__global__ void k_lz_next(int* __restrict__ kdev) {
  if (threadIdx.x == 0) *kdev = *kdev + 1;
}

// eigenvector replay: q -= a r; (r, q) <- (q / b, -b r); psi += c r_new, for step i < len
struct RJob {
  double* r;
  double* q;
  double* psi;
  const double* alpha;  // device, len
  const double* beta;
  const double* evec;
  int len;
  uint64_t n;
  uint32_t blk0;
};

__global__ __launch_bounds__(kBT) void k_replay_batch(const RJob* __restrict__ d, int nj, int i) {
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const RJob& D = d[j];
  if (i >= D.len) return;
  const double a = D.alpha[i - 1], bt = D.beta[i - 1], c = D.evec[i];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  for (uint64_t t = i0 + threadIdx.x; t < i1; t += kBT) {
    const double rv = D.r[t];
    const double qv = D.q[t] - a * rv;
    const double rn = qv / bt;
    D.r[t] = rn;
    D.q[t] = -bt * rv;
    D.psi[t] += c * rn;
  }
}

struct PsiScale {
  double* psi;
  double c;
  uint64_t n;
  uint32_t blk0;
};
__global__ __launch_bounds__(kBT) void k_psi_scale(const PsiScale* __restrict__ d, int nj) {
  const uint32_t b = blockIdx.x;
  const int j = find_job(d, nj, b);
  const PsiScale& D = d[j];
  const uint64_t i0 = (uint64_t)(b - D.blk0) * kChunk;
  const uint64_t i1 = min(i0 + kChunk, D.n);
  for (uint64_t i = i0 + threadIdx.x; i < i1; i += kBT) D.psi[i] *= D.c;
}

// ---- host helpers ------------------------------------------------------------------------
template <class D>
struct Table {
  std::vector<D> h;
  DevBuf d;
  uint32_t total = 0;
  void clear() {
    h.clear();
    total = 0;
  }
  void add(D v, uint64_t n) {
    v.blk0 = total;
    total += nblocks(n);
    h.push_back(v);
  }
  // uploads the descriptors followed by their first blocks as a compact uint32 array
  // (read by find_job: the search touches a few cache lines instead of the descriptors)
  D* upload() {
    const size_t n = h.size();
    const size_t bytes = std::max<size_t>(1, n) * (sizeof(D) + sizeof(uint32_t));
    if (d.bytes < bytes) d.alloc(bytes * 2);
    if (n) {
      stage.resize(n * sizeof(D) + n * sizeof(uint32_t));
      std::memcpy(stage.data(), h.data(), n * sizeof(D));
      uint32_t* b0 = reinterpret_cast<uint32_t*>(stage.data() + n * sizeof(D));
      for (size_t j = 0; j < n; ++j) b0[j] = h[j].blk0;
      QED_HIP(hipMemcpyAsync(d.p, stage.data(), stage.size(), hipMemcpyHostToDevice, stream()));
    }
    return d.as<D>();
  }
  std::vector<unsigned char> stage;
};

void ensure(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) b.alloc(bytes);
}

// One host-to-device copy for a group of descriptor tables: they are written into a pinned
// host buffer at 256-byte aligned offsets whose device addresses are known at once, and
// sent together by flush(). The host buffer is rewritten only after the stream has passed
// a synchronization that follows the previous flush (the callers' step structure).
struct Stager {
  unsigned char* host = nullptr;
  size_t cap = 0, off = 0;
  DevBuf dev;
  Stager() = default;
  Stager(const Stager&) = delete;
  Stager& operator=(const Stager&) = delete;
  ~Stager() {
    if (host) (void)hipHostFree(host);
  }
  void init(size_t bytes) {
    cap = bytes;
    QED_HIP(hipHostMalloc((void**)&host, cap, hipHostMallocDefault));
    dev.alloc(cap);
  }
  // space for n bytes: (host address, device address)
  std::pair<unsigned char*, unsigned char*> reserve(size_t n) {
    off = (off + 255) & ~(size_t)255;
    if (off + n > cap) fail("batched band Lanczos: descriptor staging overflow");
    std::pair<unsigned char*, unsigned char*> r{host + off, dev.as<unsigned char>() + off};
    off += n;
    return r;
  }
  template <class D>
  D* put(const Table<D>& t) {  // descriptors followed by their first blocks (find_job)
    const size_t n = t.h.size();
    auto r = reserve(std::max<size_t>(1, n) * (sizeof(D) + sizeof(uint32_t)));
    if (n) {
      std::memcpy(r.first, t.h.data(), n * sizeof(D));
      uint32_t* b0 = reinterpret_cast<uint32_t*>(r.first + n * sizeof(D));
      for (size_t j = 0; j < n; ++j) b0[j] = t.h[j].blk0;
    }
    return reinterpret_cast<D*>(r.second);
  }
  void flush() {
    if (off) QED_HIP(hipMemcpyAsync(dev.p, host, off, hipMemcpyHostToDevice, stream()));
    off = 0;
  }
};

// Persistent worker pool for the per-step host work of the batched passes (one call per
// Krylov step: spawning threads per call costs more than the work at small batches).
class HostPool {
 public:
  static HostPool& get() {
    static HostPool p;
    return p;
  }
  void run(int n, const std::function<void(int)>& f) {
    std::unique_lock<std::mutex> lk(mu_);
    idle_.wait(lk, [&] { return active_ == 0; });
    fn_ = &f;
    n_ = n;
    next_ = 0;
    err_ = nullptr;
    ++gen_;
    lk.unlock();
    wake_.notify_all();
    drain();
    lk.lock();
    idle_.wait(lk, [&] { return active_ == 0; });
    fn_ = nullptr;
    if (err_) std::rethrow_exception(err_);
  }
  int threads() const { return (int)th_.size() + 1; }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : th_) t.join();
  }

 private:
  HostPool() {
    const int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    for (int t = 1; t < nt; ++t) th_.emplace_back([this] { worker(); });
  }
  void worker() {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      wake_.wait(lk, [&] { return stop_ || (gen_ != seen && fn_ != nullptr); });
      if (stop_) return;
      seen = gen_;
      ++active_;
      lk.unlock();
      drain();
      lk.lock();
      if (--active_ == 0) idle_.notify_all();
    }
  }
  void drain() {
    try {
      for (int i = next_++; i < n_; i = next_++) (*fn_)(i);
    } catch (...) {
      std::lock_guard<std::mutex> lk(err_mu_);
      if (!err_) err_ = std::current_exception();
      next_ = n_;
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_, err_mu_;
  std::condition_variable wake_, idle_;
  const std::function<void(int)>* fn_ = nullptr;
  int n_ = 0;
  std::atomic<int> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr err_;
};

template <class F>
void parallel_for(int n, F f) {
  if (n < 4 || HostPool::get().threads() <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  const std::function<void(int)> g = f;
  HostPool::get().run(n, g);
}

}  // namespace

// ========================================================================================
// Batched Lanczos ground state. psi[j] holds the start vector (device) and receives the
// ground-state vector. E0[j], niter[j] out. Returns the H.v launch configuration used.
void lanczos_gs_batch(const std::vector<qed_sector*>& S, const std::vector<double*>& psi, double accur,
                      int max_iter, bool use_graph, int chunk, std::vector<double>& E0,
                      std::vector<int>& niter, BatchStats* stats) {
  const int nj = (int)S.size();
  E0.assign(nj, 0.0);
  niter.assign(nj, 0);
  if (nj == 0) return;
  if (chunk < 2) chunk = 2;
  chunk += chunk & 1;
  const int ld = max_iter + chunk + 2;
  // vectors: A = copy of the start vector, B = 0
  std::vector<std::unique_ptr<DevBuf>> A(nj), Bv(nj);
  DevBuf scal((size_t)nj * ld * 2 * sizeof(double));
  double* alpha = scal.as<double>();
  double* beta = alpha + (size_t)nj * ld;
  Table<LJob> T;
  Table<VFill> F;
  for (int j = 0; j < nj; ++j) {
    const uint64_t n = S[j]->D;
    A[j].reset(new DevBuf(n * 8));
    Bv[j].reset(new DevBuf(n * 8));
    QED_HIP(hipMemcpyAsync(A[j]->p, psi[j], n * 8, hipMemcpyDeviceToDevice, stream()));
    T.add(LJob{{A[j]->as<double>(), Bv[j]->as<double>()}, alpha + (size_t)j * ld, beta + (size_t)j * ld, n, 0}, n);
    F.add(VFill{Bv[j]->as<double>(), n, 0}, n);
  }
  LJob* dT = T.upload();
  VFill* dF = F.upload();
  k_fill0<<<F.total, kBT, 0, stream()>>>(dF, nj);
  // H.v batch tables for both parities: parity p: x = r, y = q
  std::vector<const qed_sector*> cs(S.begin(), S.end());
  HvBatch HB;
  hv_batch_prepare(HB, cs);
  for (int j = 0; j < nj; ++j) {
    HB.host[j].x[0] = A[j]->as<double>();
    HB.host[j].y[0] = Bv[j]->as<double>();
    HB.host[j].x[1] = Bv[j]->as<double>();
    HB.host[j].y[1] = A[j]->as<double>();
  }
  hv_batch_upload(HB);
  std::vector<int> active(nj, 1);
  DevBuf dact((size_t)nj * sizeof(int)), dk(sizeof(int)), part((size_t)T.total * sizeof(double));
  QED_HIP(hipMemcpyAsync(dact.p, active.data(), nj * sizeof(int), hipMemcpyHostToDevice, stream()));
  const int zero = 0;
  QED_HIP(hipMemcpyAsync(dk.p, &zero, sizeof(int), hipMemcpyHostToDevice, stream()));
  // autotune the batched H.v on scratch vectors (results discarded)
  HvBatchConfig cfg;
  {
    std::vector<std::unique_ptr<DevBuf>> sx(nj), sy(nj);
    HvBatch tb;
    hv_batch_prepare(tb, cs);
    for (int j = 0; j < nj; ++j) {
      sx[j].reset(new DevBuf(S[j]->D * 8));
      sy[j].reset(new DevBuf(S[j]->D * 8));
      tb.host[j].x[0] = tb.host[j].x[1] = A[j]->as<double>();
      tb.host[j].y[0] = tb.host[j].y[1] = sy[j]->as<double>();
    }
    hv_batch_upload(tb);
    cfg = hv_batch_tune(tb, dact.as<int>());
    QED_HIP(hipStreamSynchronize(stream()));
  }
  if (stats) stats->hv_config = cfg;
  const int* act = dact.as<int>();
  int* kd = dk.as<int>();
  double* pd = part.as<double>();
  const int nb1 = (nj + 3) / 4;  // one wavefront per job
  auto iteration = [&](int parity) {
    k_lz_scal<<<T.total, kBT, 0, stream()>>>(dT, nj, parity, kd, act);
    hv_batch(HB, cfg, parity, act);
    k_lz_dot<<<T.total, kBT, 0, stream()>>>(dT, nj, parity, act, pd);
    k_lz_alpha<<<nb1, 256, 0, stream()>>>(dT, nj, kd, act, pd);
    k_lz_axpy<<<T.total, kBT, 0, stream()>>>(dT, nj, parity, kd, act, pd);
    k_lz_beta<<<nb1, 256, 0, stream()>>>(dT, nj, kd, act, pd);
    k_lz_next<<<1, 64, 0, stream()>>>(kd);
  };
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    hipGraph_t g;
    QED_HIP(hipStreamSynchronize(stream()));
    QED_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    iteration(0);
    iteration(1);
    QED_HIP(hipStreamEndCapture(stream(), &g));
    QED_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    QED_HIP(hipGraphDestroy(g));
  }
  std::vector<double> ha((size_t)nj * ld), hb((size_t)nj * ld);
  std::vector<std::vector<double>> evec(nj);
  std::vector<int> done_at(nj, -1);
  int k = 0;
  int remaining = nj;
  while (remaining > 0) {
    for (int c = 0; c < chunk; c += 2) {
      if (exec) {
        QED_HIP(hipGraphLaunch(exec, stream()));
      } else {
        iteration(0);
        iteration(1);
      }
    }
    QED_HIP(hipGetLastError());
    const int k1 = k + chunk;
    QED_HIP(hipMemcpyAsync(ha.data(), alpha, ha.size() * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipMemcpyAsync(hb.data(), beta, hb.size() * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    // stopping test of the scalar Lanczos for the new steps m = k+1 .. k1 of every active
    // sector: ritz(m) = |beta_m e_m^T y_min(T_m)|; the sector retires at its first converged
    // step, as the sequential solver stops there. A sector whose Krylov space can be
    // exhausted within the chunk (D <= k1: the steps after it follow a zero vector) has
    // every step tested in order; a larger one has its last step tested first and the
    // earlier ones only when that has converged (one small eigenproblem per chunk instead
    // of one per step)
    std::vector<int> newly(nj, -1);
    parallel_for(nj, [&](int j) {
      if (!active[j]) return;
      std::vector<double> a, b, ev;
      double e0 = 0.0;
      auto ritz_at = [&](int m) {
        a.assign(ha.begin() + (size_t)j * ld, ha.begin() + (size_t)j * ld + m);
        b.assign(hb.begin() + (size_t)j * ld, hb.begin() + (size_t)j * ld + m);
        tridiag_lowest(a, b, &e0, ev);
        return std::fabs(ev.back() * b.back());
      };
      if (S[j]->D > (uint64_t)k1 && ritz_at(k1) > accur) return;
      for (int m = k + 1; m <= k1; ++m)
        if (ritz_at(m) <= accur) {
          newly[j] = m;
          evec[j] = ev;
          E0[j] = e0;
          return;
        }
    });
    for (int j = 0; j < nj; ++j)
      if (newly[j] > 0) {
        if (newly[j] > max_iter) {  // converged too late: a failure, as in the scalar solver
          evec[j].clear();
          E0[j] = std::nan("");
          newly[j] = -1;
        }
        done_at[j] = newly[j];
        active[j] = 0;
        --remaining;
      }
    k = k1;
    if (remaining > 0 && k >= max_iter) {
      // "Lanczos procedure exceeded max_iter iterations" for the sectors still running:
      // they are reported as failed (E0 = NaN, niter = -1); the others are unaffected
      for (int j = 0; j < nj; ++j)
        if (active[j]) {
          active[j] = 0;
          done_at[j] = -1;
          evec[j].clear();
          E0[j] = std::nan("");
          --remaining;
        }
    }
    QED_HIP(hipMemcpyAsync(dact.p, active.data(), nj * sizeof(int), hipMemcpyHostToDevice, stream()));
  }
  if (exec) QED_HIP(hipGraphExecDestroy(exec));
  if (stats) {
    stats->lanczos_iterations = k;
  }
  // ground-state vectors: replay the recurrence with the stored scalars
  int maxlen = 0;
  std::vector<DevBuf> ev_dev;
  Table<RJob> R;
  Table<PsiScale> PS;
  std::vector<int> ones(nj, 1);
  for (int j = 0; j < nj; ++j) {
    niter[j] = done_at[j];
    const uint64_t n = S[j]->D;
    ev_dev.emplace_back(std::max<size_t>(1, evec[j].size()) * 8);
    if (!evec[j].empty())
      QED_HIP(hipMemcpyAsync(ev_dev.back().p, evec[j].data(), evec[j].size() * 8, hipMemcpyHostToDevice, stream()));
    // r = start vector (copy kept in psi), q = 0, psi *= evec[0]
    QED_HIP(hipMemcpyAsync(A[j]->p, psi[j], n * 8, hipMemcpyDeviceToDevice, stream()));
    R.add(RJob{A[j]->as<double>(), Bv[j]->as<double>(), psi[j], alpha + (size_t)j * ld, beta + (size_t)j * ld,
               ev_dev.back().as<double>(), (int)evec[j].size(), n, 0},
          n);
    PS.add(PsiScale{psi[j], evec[j].empty() ? 0.0 : evec[j][0], n, 0}, n);
    maxlen = std::max(maxlen, (int)evec[j].size());
  }
  k_fill0<<<F.total, kBT, 0, stream()>>>(dF, nj);
  PsiScale* dPS = PS.upload();
  k_psi_scale<<<PS.total, kBT, 0, stream()>>>(dPS, nj);
  RJob* dR = R.upload();
  QED_HIP(hipMemcpyAsync(dact.p, ones.data(), nj * sizeof(int), hipMemcpyHostToDevice, stream()));
  // replay H.v: x = r (A), y = q (B): parity 0 of HB
  for (int i = 1; i < maxlen; ++i) {
    for (int j = 0; j < nj; ++j) active[j] = i < (int)evec[j].size() ? 1 : 0;
    QED_HIP(hipMemcpyAsync(dact.p, active.data(), nj * sizeof(int), hipMemcpyHostToDevice, stream()));
    hv_batch(HB, cfg, 0, act);
    k_replay_batch<<<R.total, kBT, 0, stream()>>>(dR, nj, i);
  }
  QED_HIP(hipGetLastError());
  QED_HIP(hipStreamSynchronize(stream()));
}

// ========================================================================================
// Batched band Lanczos (Freund), one host state per job, one synchronization per step.
namespace {

struct BandJob {
  const qed_sector* s = nullptr;
  uint64_t n = 0;
  int p0 = 0, pc = 0, M0 = 0, M0p = 0, j = 0;
  bool done = false, fresh = true;
  std::vector<double*> v;
  std::vector<const double*> ph;
  std::vector<double> t, P;
  std::vector<int> defl, I;
  double ev_old = 1e12, ev_old2 = 1e12;
  std::vector<double> evec;  // lowest eigenvector of the last test (warm start)
  std::vector<std::unique_ptr<DevBuf>> owned;
  std::vector<double*> free_list;
  // device scalars of the step: [0..p0] multi-dot, [p0+1 .. p0+pc] pending-block overlaps,
  // then the deflated-vector overlaps and T(j, j)
  double* scd = nullptr;  // this job's kSc scalars in the batch's device array
  double* sch = nullptr;  // and their host copy
  // pending host reads from the previous step
  int prev_j = -1, prev_nk = 0, prev_pc = 0, prev_nI = 0, nk = 0;
  bool check_pending = false;
  double z = 0.0;
  std::vector<double> hcoef;  // host coefficients uploaded for phase 4
  double& T(int a, int b) { return t[(size_t)b * M0 + a]; }
  double* get() {
    if (!free_list.empty()) {
      double* p = free_list.back();
      free_list.pop_back();
      return p;
    }
    owned.emplace_back(new DevBuf(n * 8));
    return owned.back()->as<double>();
  }
  void put(double* p) {
    if (p) free_list.push_back(p);
  }
};

constexpr int kSc = 4 * kBM + 8;  // device scalars per job and step

}  // namespace

void band_lanczos_batch(const std::vector<qed_sector*>& S, const std::vector<const double*>& phi,
                        const std::vector<int>& p, int max_iter_BL, double accur_deflation,
                        double accur_band_lanczos, std::vector<std::vector<double>>& evals,
                        std::vector<std::vector<double>>& Q, BatchStats* stats, double min_gap) {
  const int nj = (int)S.size();
  std::atomic<int> gap_fail{0};  // a degenerate pair found by a worker (no_degenerate_BL)
  evals.assign(nj, {});
  Q.assign(nj, {});
  std::vector<BandJob> J(nj);
  std::vector<const qed_sector*> cs;
  std::vector<int> hv_of(nj, -1);
  for (int i = 0; i < nj; ++i) {
    BandJob& b = J[i];
    b.s = S[i];
    b.n = S[i]->D;
    b.p0 = b.pc = p[i];
    if (b.n == 0 || b.p0 == 0) {
      b.done = true;
      continue;
    }
    if (b.p0 + 1 > kBM) fail("batched band Lanczos: too many start vectors");
    int M0 = (int)(14 * b.p0 * std::log(1.0 * b.n));
    if (M0 <= 1) fail("band Lanczos: too few iterations requested");
    if (M0 > max_iter_BL) M0 = max_iter_BL;
    b.M0p = (M0 / b.p0 + 1) * b.p0;
    b.M0 = b.M0p - b.pc;
    b.v.assign(b.M0p + 1, nullptr);
    for (int k = 0; k < b.p0; ++k) b.ph.push_back(phi[i] + (uint64_t)k * b.n);
    for (int k = 0; k < b.p0; ++k) {
      b.v[k] = b.get();
      QED_HIP(hipMemcpyAsync(b.v[k], b.ph[k], b.n * 8, hipMemcpyDeviceToDevice, stream()));
    }
    b.t.assign((size_t)b.M0 * b.M0, 0.0);
    b.P.assign((size_t)b.p0 * b.M0, 0.0);
    b.defl.assign(b.M0, 0);
    hv_of[i] = (int)cs.size();
    cs.push_back(S[i]);
  }
  // the step scalars of all jobs in one device array (one download per step, into pinned
  // host memory); the descriptor tables of a step go up in two staged copies: phase 1's,
  // and those of phases 2-6 together with the host-supplied phase-4 coefficients
  DevBuf sc_all((size_t)std::max(1, nj) * kSc * sizeof(double));
  double* sch_all = nullptr;
  QED_HIP(hipHostMalloc((void**)&sch_all, (size_t)std::max(1, nj) * kSc * sizeof(double), hipHostMallocDefault));
  std::unique_ptr<double, void (*)(double*)> sch_keep(sch_all, [](double* p) { (void)hipHostFree(p); });
  std::fill(sch_all, sch_all + (size_t)std::max(1, nj) * kSc, 0.0);
  for (int i = 0; i < nj; ++i) {
    J[i].scd = sc_all.as<double>() + (size_t)i * kSc;
    J[i].sch = sch_all + (size_t)i * kSc;
  }
  // staging bounds: per step at most 8 + 2 kBM tables of nj descriptors, the H.v job table,
  // the activity flags and the coefficients
  const size_t per_table = (size_t)std::max(1, nj) * (sizeof(MAxpy) + sizeof(uint32_t)) + 256;
  Stager stage1, stage2;
  stage1.init(per_table + 256);
  stage2.init((8 + 2 * (size_t)kBM) * per_table + (size_t)std::max(1, nj) * (sizeof(HvJob) + sizeof(int)) +
              (size_t)std::max(1, nj) * kBM * sizeof(double) + 4096);
  std::vector<std::function<void()>> launches;  // the staged group's kernels, in order
  HvBatch HB;
  if (!cs.empty()) hv_batch_prepare(HB, cs);
  DevBuf part;
  Table<MDot> TD;
  Table<MAxpy> TA;
  Table<MRev> TR;
  Table<VFill> TF;
  HvBatchConfig cfg;
  bool tuned = false;
  int steps = 0;
  std::vector<int> act(cs.size(), 0);
  // the run_* helpers stage their table into stage2 and queue the launch; flush_launches
  // sends the group in one copy and then launches the queued kernels in order
  auto run_mdot = [&]() {
    if (TD.h.empty()) return;
    MDot* d = stage2.put(TD);
    const uint32_t total = TD.total;
    const int n = (int)TD.h.size();
    ensure(part, (size_t)total * kBM * sizeof(double));
    launches.push_back([&part, d, total, n] {
      k_mdot_part<<<total, kBT, 0, stream()>>>(d, n, part.as<double>());
      k_mdot_final<<<(n * kBM + 3) / 4, 256, 0, stream()>>>(d, n, part.as<double>());
    });
  };
  auto run_maxpy = [&]() {
    if (TA.h.empty()) return;
    MAxpy* d = stage2.put(TA);
    const uint32_t total = TA.total;
    const int n = (int)TA.h.size();
    launches.push_back([d, total, n] { k_maxpy<<<total, kBT, 0, stream()>>>(d, n); });
  };
  auto run_mrev = [&]() {
    if (TR.h.empty()) return;
    MRev* d = stage2.put(TR);
    const uint32_t total = TR.total;
    const int n = (int)TR.h.size();
    launches.push_back([d, total, n] { k_mrev<<<total, kBT, 0, stream()>>>(d, n); });
  };
  auto flush_launches = [&]() {
    stage2.flush();
    for (auto& f : launches) f();
    launches.clear();
  };
  while (true) {
    // ---- phase 1: norm of v[j] and overlaps with the start vectors (all live jobs) ----
    TD.clear();
    bool live = false;
    for (auto& b : J) {
      if (b.done) continue;
      live = true;
      if (b.j >= b.M0) continue;
      if (!b.v[b.j]) fail("batched band Lanczos: missing Krylov vector");
      MDot d{};
      d.x = b.v[b.j];
      d.y[0] = b.v[b.j];
      for (int k = 0; k < b.p0; ++k) d.y[k + 1] = b.ph[k];
      d.m = b.p0 + 1;
      d.n = b.n;
      d.out = b.scd;
      TD.add(d, b.n);
    }
    if (!live) break;
    if (!TD.h.empty()) {  // (every live job may have reached M0: it retires below)
      MDot* d = stage1.put(TD);
      ensure(part, (size_t)TD.total * kBM * sizeof(double));
      stage1.flush();
      k_mdot_part<<<TD.total, kBT, 0, stream()>>>(d, (int)TD.h.size(), part.as<double>());
      k_mdot_final<<<((int)TD.h.size() * kBM + 3) / 4, 256, 0, stream()>>>(d, (int)TD.h.size(), part.as<double>());
    }
    QED_HIP(hipMemcpyAsync(sch_all, sc_all.p, (size_t)nj * kSc * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    ++steps;
    // ---- host: complete the previous step's band entries and its eigenvalue test, then
    //      take this step's deflation decision
    std::vector<int> proceed(nj, 0);
    std::atomic<int> n_checks{0}, n_dense{0};
    const auto tc0 = std::chrono::steady_clock::now();
    parallel_for(nj, [&](int i) {
      BandJob& b = J[i];
      if (b.done) return;
      const double* s = b.sch;
      if (b.prev_j >= 0) {
        // previous step: [kBM, kBM + nk) pending-block overlaps, [2 kBM, 2 kBM + nI)
        // deflated-vector overlaps, [3 kBM] T(j, j); phase 1 wrote only [0, p0]
        const int pj = b.prev_j;
        for (int q = 0; q < b.prev_nk; ++q) {
          const int kk = pj + 1 + q;
          if (kk >= b.prev_pc) b.T(pj, kk - b.prev_pc) = s[kBM + q];
        }
        for (int q = 0; q < b.prev_nI; ++q) b.T(b.I[q], pj) = s[2 * kBM + q];
        b.T(pj, pj) = s[3 * kBM];
        for (int kk : b.I) b.T(pj, kk) = b.T(kk, pj);
        b.prev_j = -1;
        if (b.check_pending) {
          b.check_pending = false;
          // lowest eigenvalue of the pj x pj projected matrix: shifted inverse iteration on
          // the band (the deflated vectors' rows and columns as a border) from the previous
          // eigenvector, the shift below the previous value by more than its last decrease;
          // dense eigensolver otherwise
          double e0 = 0.0;
          bool ok = false;
          n_checks++;
          if (min_gap <= 0.0 && !b.evec.empty() && b.ev_old2 < 1e11) {
            // a second try with a lower shift and more solves when the first shift was not
            // below the spectrum or the residual test failed
            const double drop = std::fabs(b.ev_old2 - b.ev_old);
            const double scale = std::max(1.0, std::fabs(b.ev_old));
            // the band part of T is stored in the lower triangle only
            auto Tb = [&](int r, int c) { return r >= c ? b.T(r, c) : b.T(c, r); };
            std::vector<int> bd;
            for (int k : b.I)
              if (k < pj) bd.push_back(k);
            std::vector<double> x0 = b.evec;
            ok = band_lowest(Tb, pj, b.p0, b.ev_old - std::max(4.0 * drop, 1e-9 * scale), 3, b.evec, e0, bd);
            if (!ok) {
              b.evec = x0;
              ok = band_lowest(Tb, pj, b.p0, b.ev_old - std::max(64.0 * drop, 1e-6 * scale), 10, b.evec, e0,
                               bd);
            }
          }
          if (!ok) {
            n_dense++;
            std::vector<double> H((size_t)pj * pj), e(pj);
            for (int c = 0; c < pj; ++c)
              for (int r = 0; r < pj; ++r) H[(size_t)c * pj + r] = b.T(r, c);
            sym_eigen(pj, H.data(), e.data(), true);
            e0 = e[0];
            b.evec.assign(H.begin(), H.begin() + pj);
            // no_degenerate_BL (src_ed/Hamiltonian/Lanczos.hpp:368-370): the guard needs the
            // second eigenvalue, so the dense eigensolver runs at every test when it is on
            if (min_gap > 0.0 && pj > 1 && e[1] - e[0] < min_gap) {
              gap_fail = 1;
              b.done = true;
              return;
            }
          }
          const double test = std::fabs(e0 - b.ev_old);
          b.ev_old2 = b.ev_old;
          b.ev_old = e0;
          if (test < accur_band_lanczos) {
            b.j = pj;  // stop with the pj x pj projected matrix
            b.done = true;
            return;
          }
        }
        if (b.j >= b.M0) {
          b.done = true;
          return;
        }
      }
      const double z = std::sqrt(std::max(s[0], 0.0));
      if (z < accur_deflation) {
        if (b.j >= b.pc) {
          b.I.push_back(b.j - b.pc);
          b.defl[b.j - b.pc] = 1;
        }
        b.pc--;
        if (b.pc == 0) {
          b.done = true;
          return;
        }
        b.put(b.v[b.j]);
        for (int k = b.j; k < b.j + b.pc; ++k) b.v[k] = b.v[k + 1];
        b.v[b.j + b.pc] = nullptr;
        return;  // phase 1 again at the same j
      }
      if (b.j >= b.pc) b.T(b.j, b.j - b.pc) = z;
      for (int k = 0; k < b.p0; ++k) b.P[(size_t)b.j * b.p0 + k] = s[k + 1] / z;
      b.z = z;
      proceed[i] = 1;
    });
    if (gap_fail)
      fail("band Lanczos: the gap between the first two eigenvalues is smaller than " + std::to_string(min_gap));
    if (stats) {
      stats->t_band_checks += std::chrono::duration<double>(std::chrono::steady_clock::now() - tc0).count();
      stats->band_checks += n_checks;
      stats->band_dense_checks += n_dense;
    }
    // ---- phase 2: v[j] /= z; pending block orthogonalized against v[j] ----
    // scale: y = y * (1/z) with no sum (MAxpy with m = 0)
    TA.clear();
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      BandJob& b = J[i];
      MAxpy a{};
      a.y = b.v[b.j];
      a.m = 0;
      a.sgn = 1.0;
      a.scale = 1.0 / b.z;
      a.n = b.n;
      TA.add(a, b.n);
    }
    run_maxpy();
    TD.clear();
    TR.clear();
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      BandJob& b = J[i];
      const int nk = b.pc - 1;
      b.nk = nk;
      if (nk <= 0) continue;
      MDot d{};
      d.x = b.v[b.j];
      for (int q = 0; q < nk; ++q) d.y[q] = b.v[b.j + 1 + q];
      d.m = nk;
      d.n = b.n;
      d.out = b.scd + kBM;
      TD.add(d, b.n);
      MRev r{};
      r.x = b.v[b.j];
      for (int q = 0; q < nk; ++q) {
        r.y[q] = b.v[b.j + 1 + q];
        r.c[q] = b.scd + kBM + q;
      }
      r.m = nk;
      r.n = b.n;
      TR.add(r, b.n);
    }
    run_mdot();
    run_mrev();
    // ---- phase 3: w = H v[j] ----
    TF.clear();
    std::vector<double*> w(nj, nullptr);
    for (int i = 0; i < nj; ++i) {
      BandJob& b = J[i];
      if (proceed[i]) {
        w[i] = b.get();
        TF.add(VFill{w[i], b.n, 0}, b.n);
      }
    }
    if (!TF.h.empty()) {
      VFill* d = stage2.put(TF);
      const uint32_t total = TF.total;
      const int n = (int)TF.h.size();
      launches.push_back([d, total, n] { k_fill0<<<total, kBT, 0, stream()>>>(d, n); });
    }
    for (size_t q = 0; q < cs.size(); ++q) act[q] = 0;
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      const int h = hv_of[i];
      act[h] = 1;
      HB.host[h].x[0] = J[i].v[J[i].j];
      HB.host[h].y[0] = w[i];
    }
    if (!cs.empty()) {
      auto jr = stage2.reserve(HB.host.size() * sizeof(HvJob));
      std::memcpy(jr.first, HB.host.data(), HB.host.size() * sizeof(HvJob));
      auto ar = stage2.reserve(act.size() * sizeof(int));
      std::memcpy(ar.first, act.data(), act.size() * sizeof(int));
      HvJob* jobs_dev = reinterpret_cast<HvJob*>(jr.second);
      const int* act_dev = reinterpret_cast<const int*>(ar.second);
      launches.push_back([&, jobs_dev, act_dev] {
        if (!tuned) {
          // tune on scratch outputs (the trial products must not land in w)
          std::vector<std::unique_ptr<DevBuf>> sy(cs.size());
          HvBatch tb;
          hv_batch_prepare(tb, cs);
          for (size_t q = 0; q < cs.size(); ++q) {
            sy[q].reset(new DevBuf(cs[q]->D * 8));
            tb.host[q].x[0] = tb.host[q].x[1] = HB.host[q].x[0] ? HB.host[q].x[0] : sy[q]->as<double>();
            tb.host[q].y[0] = tb.host[q].y[1] = sy[q]->as<double>();
          }
          hv_batch_upload(tb);
          cfg = hv_batch_tune(tb, act_dev);
          QED_HIP(hipStreamSynchronize(stream()));
          tuned = true;
          if (stats) stats->hv_config = cfg;
        }
        HB.jobs = jobs_dev;
        hv_batch(HB, cfg, 0, act_dev);
      });
    }
    // ---- phase 4: w -= sum_k T(k, j) v[k] over the band; release the oldest vector ----
    TA.clear();
    auto hcr = stage2.reserve((size_t)std::max(1, nj) * kBM * sizeof(double));
    double* hch = reinterpret_cast<double*>(hcr.first);
    const double* hcd = reinterpret_cast<const double*>(hcr.second);
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      BandJob& b = J[i];
      int k0 = 0;
      if (b.j > b.pc) {
        k0 = b.j - b.pc;
        if (b.defl[k0 - 1] == 0) {
          b.put(b.v[k0 - 1]);
          b.v[k0 - 1] = nullptr;
        }
      }
      // T(k, j) = T(j, k): k = j - pc is z; k = j + 1 + q - pc (0 <= q < nk) is the q-th
      // pending-block overlap of phase 2, still on the device; other entries are on the host
      MAxpy a{};
      a.y = w[i];
      a.sgn = -1.0;
      a.scale = 1.0;
      a.n = b.n;
      int m = 0;
      b.hcoef.clear();
      std::vector<int> hslot;
      for (int k = k0; k < b.j; ++k) {
        if (m >= kBM) fail("batched band Lanczos: band too wide");
        a.x[m] = b.v[k];
        const int q = k + b.pc - b.j - 1;
        if (k == b.j - b.pc) {
          hslot.push_back(m);
          b.hcoef.push_back(b.z);
        } else if (q >= 0 && q < b.nk) {
          a.c[m] = b.scd + kBM + q;
        } else {
          hslot.push_back(m);
          b.hcoef.push_back(b.T(b.j, k));
        }
        ++m;
      }
      a.m = m;
      if (!b.hcoef.empty()) {
        std::copy(b.hcoef.begin(), b.hcoef.end(), hch + (size_t)i * kBM);
        for (size_t u = 0; u < hslot.size(); ++u) a.c[hslot[u]] = hcd + (size_t)i * kBM + u;
      }
      if (m > 0) TA.add(a, b.n);
    }
    run_maxpy();
    // ---- phase 5: deflated vectors, one at a time (w -= <v_k, w> v_k) ----
    int maxI = 0;
    for (int i = 0; i < nj; ++i)
      if (proceed[i]) maxI = std::max(maxI, (int)J[i].I.size());
    for (int u = 0; u < maxI; ++u) {
      TD.clear();
      TA.clear();
      for (int i = 0; i < nj; ++i) {
        if (!proceed[i] || u >= (int)J[i].I.size()) continue;
        BandJob& b = J[i];
        MDot d{};
        d.x = b.v[b.I[u]];
        d.y[0] = w[i];
        d.m = 1;
        d.n = b.n;
        d.out = b.scd + 2 * kBM + u;
        TD.add(d, b.n);
        MAxpy a{};
        a.y = w[i];
        a.x[0] = b.v[b.I[u]];
        a.c[0] = b.scd + 2 * kBM + u;
        a.sgn = -1.0;
        a.scale = 1.0;
        a.m = 1;
        a.n = b.n;
        TA.add(a, b.n);
      }
      run_mdot();
      run_maxpy();
    }
    // ---- phase 6: T(j, j) = <v_j, w>; w -= T(j, j) v_j ----
    TD.clear();
    TA.clear();
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      BandJob& b = J[i];
      if ((int)b.I.size() > kBM) fail("batched band Lanczos: too many deflations");
      MDot d{};
      d.x = b.v[b.j];
      d.y[0] = w[i];
      d.m = 1;
      d.n = b.n;
      d.out = b.scd + 3 * kBM;
      TD.add(d, b.n);
      MAxpy a{};
      a.y = w[i];
      a.x[0] = b.v[b.j];
      a.c[0] = b.scd + 3 * kBM;
      a.sgn = -1.0;
      a.scale = 1.0;
      a.m = 1;
      a.n = b.n;
      TA.add(a, b.n);
    }
    run_mdot();
    run_maxpy();
    flush_launches();
    for (int i = 0; i < nj; ++i) {
      if (!proceed[i]) continue;
      BandJob& b = J[i];
      b.v[b.j + b.pc] = w[i];
      b.prev_j = b.j;
      b.prev_pc = b.pc;
      b.prev_nk = b.nk;
      b.prev_nI = (int)b.I.size();
      b.check_pending = (b.j >= 4 * b.p0 && b.j % b.p0 == 0);
      ++b.j;
    }
    // (the next phase 1 reads this step's device scalars, completes T and runs the test)
  }
  QED_HIP(hipStreamSynchronize(stream()));
  if (stats) stats->band_steps = steps;
  // projected matrices: eigen-decomposition and Q = P U (host, in parallel)
  parallel_for(nj, [&](int i) {
    BandJob& b = J[i];
    const int jj = b.j;
    if (b.n == 0 || b.p0 == 0 || jj == 0) return;
    std::vector<double> H((size_t)jj * jj);
    for (int c = 0; c < jj; ++c)
      for (int r = 0; r < jj; ++r) H[(size_t)c * jj + r] = b.T(r, c);
    evals[i].assign(jj, 0.0);
    sym_eigen(jj, H.data(), evals[i].data(), true);
    Q[i].assign((size_t)b.p0 * jj, 0.0);
    for (int c = 0; c < jj; ++c)
      for (int m = 0; m < jj; ++m) {
        const double u = H[(size_t)c * jj + m];
        if (u == 0.0) continue;
        for (int k = 0; k < b.p0; ++k) Q[i][(size_t)c * b.p0 + k] += b.P[(size_t)m * b.p0 + k] * u;
      }
  });
}

BatchStats& last_batch_stats() {
  static BatchStats s;
  return s;
}

}  // namespace qed
