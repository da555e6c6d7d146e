// Krylov solvers on device vectors: BLAS-1 with wave64 shuffle reductions, Lanczos for
// the ground state (eigenvalue pass + eigenvector replay), the scalar Lanczos of the
// continued-fraction Green function and Freund's band Lanczos. The recurrence scalars
// (alpha, beta, the projected band matrix) and their small eigenproblems stay in FP64 on
// the host; every O(D) operation runs on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/qed.h"
#include "qed_linalg.hpp"
#include "qed_sector.hpp"

namespace qed {

namespace {

constexpr int RB = 256;
constexpr int kMaxMulti = 64;
constexpr uint32_t kMaxBlocks = 2048;

struct PtrList {
  const double* p[kMaxMulti];
};
struct MutPtrList {
  double* p[kMaxMulti];
};

uint32_t red_blocks(uint64_t n) {
  uint64_t b = (n + (uint64_t)RB * 8 - 1) / ((uint64_t)RB * 8);
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), kMaxBlocks);
}

// scratch for block partials and final scalars (library calls are serialized)
DevBuf& red_scratch() {
  static DevBuf b((size_t)kMaxBlocks * (kMaxMulti + 1) * sizeof(double) + 4096);
  return b;
}
std::vector<double>& host_scratch() {
  static std::vector<double> h;
  return h;
}

__global__ __launch_bounds__(RB) void k_multi_dot(const double* __restrict__ x, PtrList ys, int m,
                                                  uint64_t n, double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  double acc[8];
  for (int base = 0; base < m; base += 8) {
    int cnt = min(8, m - base);
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
      double xv = x[i];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < cnt) acc[q] += xv * ys.p[base + q][i];
    }
    for (int q = 0; q < cnt; ++q) {
      double s = block_sum(acc[q], sh);
      if (threadIdx.x == 0) part[(size_t)(base + q) * gridDim.x + blockIdx.x] = s;
    }
  }
}

__global__ void k_finish(const double* __restrict__ part, int nb, int m, double* __restrict__ out) {
  __shared__ double sh[RB / 64];
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) s += part[(size_t)j * nb + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = s;
  }
}

__global__ void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void k_zero(double* __restrict__ x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

__global__ void k_scal(double a, double* __restrict__ x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

__global__ void k_scal2(double ar, double* __restrict__ r, double aq, double* __restrict__ q,
                        uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    r[i] *= ar;
    q[i] *= aq;
  }
}

// y += a x, partial sums of y^2
__global__ __launch_bounds__(RB) void k_axpy_nrm(double a, const double* __restrict__ x,
                                                 double* __restrict__ y, uint64_t n,
                                                 double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
    double v = y[i] + a * x[i];
    y[i] = v;
    acc += v * v;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---- fused Lanczos step (device scalars) ---------------------------------------------
// sc: [0] 1/beta_{k-1}, [1] -beta_{k-1}/beta_{k-2} (the product's scalars), [2] alpha_k,
// [3] beta_k, [4] the axpy coefficient -alpha_k / beta_{k-1}, [5] beta_{k-1}
// alpha_k = sc[0] * sum(part); the axpy coefficient follows
__global__ void k_lzf_alpha(const double* __restrict__ part, int nb, double* __restrict__ sc) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double a = s * sc[0];
    sc[2] = a;
    sc[4] = -a * sc[0];
  }
}

// beta_k = sqrt(sum(part)); the next product's scalars
__global__ void k_lzf_beta(const double* __restrict__ part, int nb, double* __restrict__ sc) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const double b = sqrt(s);
    sc[3] = b;
    sc[1] = -b * sc[0];  // -beta_k / beta_{k-1}
    sc[0] = b != 0.0 ? 1.0 / b : 0.0;
    sc[5] = b;
  }
}

// y += sc[4] x, partial sums of y^2
__global__ __launch_bounds__(RB) void k_lzf_axpy_nrm(const double* __restrict__ sc, const double* __restrict__ x,
                                                     double* __restrict__ y, uint64_t n,
                                                     double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  const double a = sc[4];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
    double v = y[i] + a * x[i];
    y[i] = v;
    acc += v * v;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// eigenvector replay step: q -= a r; (r, q) <- (q / b, -b r); psi += c r_new
__global__ void k_replay(double a, double b, double c, double* __restrict__ r,
                         double* __restrict__ q, double* __restrict__ psi, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double rv = r[i];
    double qv = q[i] - a * rv;
    double rn = qv / b;
    r[i] = rn;
    q[i] = -b * rv;
    psi[i] += c * rn;
  }
}

// continued-fraction Lanczos step: (psi, q) <- (q / b, -b psi)
__global__ void k_swapscale(double b, double* __restrict__ psi, double* __restrict__ q, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double t = psi[i];
    psi[i] = q[i] / b;
    q[i] = -b * t;
  }
}

// y -= sum_k c[k] xs[k]
__global__ void k_multi_axpy(PtrList xs, const double* __restrict__ c, int m, double* __restrict__ y,
                             uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double v = y[i];
    for (int k = 0; k < m; ++k) v -= c[k] * xs.p[k][i];
    y[i] = v;
  }
}

// ys[k] -= c[k] x
__global__ void k_multi_axpy_rev(MutPtrList ys, const double* __restrict__ c, int m,
                                 const double* __restrict__ x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double xv = x[i];
    for (int k = 0; k < m; ++k) ys.p[k][i] -= c[k] * xv;
  }
}

// y += 0.5 x with 16-byte accesses: the read-modify-write stream the H.v is compared with
__global__ __launch_bounds__(256) void k_stream_rmw(const double2* __restrict__ x, double2* __restrict__ y,
                                                   uint64_t n2) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * 256) {
    const double2 a = x[i];
    double2 b = y[i];
    b.x += 0.5 * a.x;
    b.y += 0.5 * a.y;
    y[i] = b;
  }
}

uint32_t stream_blocks(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), (uint64_t)device_cu_count() * 16);
}

}  // namespace

// lowest eigenvalue of the Lanczos tridiagonal and its normalized eigenvector
// (QL eigenvalues, then inverse iteration for the vector)
void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, double* e0,
                    std::vector<double>& vec) {
  int n = (int)a.size();
  vec.assign(n, 1.0);
  if (n == 1) {
    *e0 = a[0];
    return;
  }
  std::vector<double> d(a), e(b.begin(), b.begin() + (n - 1));
  tridiag_eigen(n, d.data(), e.data(), nullptr);
  double lam = d[0];
  *e0 = lam;
  // inverse iteration on (T - lam) with partial pivoting
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = std::max(scale, std::fabs(a[i]) + (i < n - 1 ? std::fabs(b[i]) : 0.0));
  double shift = lam - 1e-14 * std::max(scale, 1.0);
  for (int it = 0; it < 3; ++it) {
    // solve (T - shift I) x = vec; Gaussian elimination on the tridiagonal with pivoting
    std::vector<double> dl(n - 1), dm(n), du(n - 1), du2(n, 0.0), rhs(vec);
    for (int i = 0; i < n; ++i) dm[i] = a[i] - shift;
    for (int i = 0; i < n - 1; ++i) dl[i] = du[i] = b[i];
    std::vector<int> swapped(n, 0);
    for (int i = 0; i < n - 1; ++i) {
      if (std::fabs(dm[i]) >= std::fabs(dl[i])) {
        if (dm[i] == 0.0) dm[i] = 1e-300;
        double f = dl[i] / dm[i];
        dm[i + 1] -= f * du[i];
        rhs[i + 1] -= f * rhs[i];
        if (i < n - 2) du2[i] = 0.0;
      } else {
        double f = dm[i] / dl[i];
        dm[i] = dl[i];
        double t = dm[i + 1];
        dm[i + 1] = du[i] - f * t;
        if (i < n - 2) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du2[i];
        }
        du[i] = t;
        std::swap(rhs[i], rhs[i + 1]);
        rhs[i + 1] -= f * rhs[i];
      }
    }
    if (dm[n - 1] == 0.0) dm[n - 1] = 1e-300;
    std::vector<double> x(n);
    x[n - 1] = rhs[n - 1] / dm[n - 1];
    if (n >= 2) x[n - 2] = (rhs[n - 2] - du[n - 2] * x[n - 1]) / dm[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (rhs[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dm[i];
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; ++i) vec[i] = x[i] / nrm;
  }
  // sign convention: first component positive (sign does not affect the Ritz residual)
  if (vec[0] < 0)
    for (double& v : vec) v = -v;
}

namespace {

}  // namespace

std::recursive_mutex& lib_mutex() {
  static std::recursive_mutex m;
  return m;
}

void multi_dot(const double* x, const double* const* ys, int m, uint64_t n, double* out, double* work) {
  if (m > kMaxMulti) fail("multi_dot: too many vectors");
  if (m == 0) return;
  if (n == 0) {
    std::fill(out, out + m, 0.0);
    return;
  }
  PtrList l{};
  for (int k = 0; k < m; ++k) l.p[k] = ys[k];
  uint32_t nb = red_blocks(n);
  double* part = work ? work : red_scratch().as<double>();
  double* fin = part + (size_t)kMaxBlocks * kMaxMulti;
  k_multi_dot<<<nb, RB, 0, stream()>>>(x, l, m, n, part);
  k_finish<<<1, RB, 0, stream()>>>(part, nb, m, fin);
  QED_HIP(hipMemcpyAsync(out, fin, m * sizeof(double), hipMemcpyDeviceToHost, stream()));
  QED_HIP(hipStreamSynchronize(stream()));
}

double dot(const double* x, const double* y, uint64_t n, double* work) {
  double r;
  multi_dot(x, &y, 1, n, &r, work);
  return r;
}

double nrm2(const double* x, uint64_t n) { return std::sqrt(dot(x, x, n)); }

void axpy(double a, const double* x, double* y, uint64_t n) {
  if (n) k_axpy<<<stream_blocks(n), 256, 0, stream()>>>(a, x, y, n);
}
void scal(double a, double* x, uint64_t n) {
  if (n) k_scal<<<stream_blocks(n), 256, 0, stream()>>>(a, x, n);
}
void scal2(double ar, double* r, double aq, double* q, uint64_t n) {
  if (n) k_scal2<<<stream_blocks(n), 256, 0, stream()>>>(ar, r, aq, q, n);
}

double axpy_nrm2(double a, const double* x, double* y, uint64_t n, double* work) {
  uint32_t nb = red_blocks(n);
  double* part = work ? work : red_scratch().as<double>();
  double* fin = part + (size_t)kMaxBlocks * kMaxMulti;
  k_axpy_nrm<<<nb, RB, 0, stream()>>>(a, x, y, n, part);
  k_finish<<<1, RB, 0, stream()>>>(part, nb, 1, fin);
  double s;
  QED_HIP(hipMemcpyAsync(&s, fin, sizeof(double), hipMemcpyDeviceToHost, stream()));
  QED_HIP(hipStreamSynchronize(stream()));
  return std::sqrt(s);
}

void multi_axpy(const double* c, const double* const* xs, int m, double* y, uint64_t n) {
  if (m == 0 || n == 0) return;
  if (m > kMaxMulti) fail("multi_axpy: too many vectors");
  PtrList l{};
  for (int k = 0; k < m; ++k) l.p[k] = xs[k];
  double* cd = red_scratch().as<double>() + (size_t)kMaxBlocks * kMaxMulti + 64;
  QED_HIP(hipMemcpyAsync(cd, c, m * sizeof(double), hipMemcpyHostToDevice, stream()));
  k_multi_axpy<<<stream_blocks(n), 256, 0, stream()>>>(l, cd, m, y, n);
  QED_HIP(hipStreamSynchronize(stream()));
}

void multi_axpy_rev(const double* c, const double* x, double* const* ys, int m, uint64_t n) {
  if (m == 0 || n == 0) return;
  if (m > kMaxMulti) fail("multi_axpy_rev: too many vectors");
  MutPtrList l{};
  for (int k = 0; k < m; ++k) l.p[k] = ys[k];
  double* cd = red_scratch().as<double>() + (size_t)kMaxBlocks * kMaxMulti + 64;
  QED_HIP(hipMemcpyAsync(cd, c, m * sizeof(double), hipMemcpyHostToDevice, stream()));
  k_multi_axpy_rev<<<stream_blocks(n), 256, 0, stream()>>>(l, cd, m, x, n);
  QED_HIP(hipStreamSynchronize(stream()));
}

// The eigenvalue pass of the Lanczos ground state, one fused step per iteration where the
// sector takes the fused product: with w_k = beta_{k-1} v_k kept unnormalized,
//   y = H v_k - beta_{k-1} v_{k-1} = (H w_k) / beta_{k-1} - (beta_{k-1} / beta_{k-2}) w_{k-1}
// is one pass of the product kernel (which also sums <w_k, y>), and
//   w_{k+1} = y - (alpha_k / beta_{k-1}) w_k,  beta_k = |w_{k+1}|
// one more streaming pass: 24 D bytes beyond the product instead of 72 D for the separate
// scale, dot and axpy passes, and one scalar download per step.
bool lanczos_eigen_fused(qed_sector& s, double accur, int max_iter, const double* psi, std::vector<double>& alpha,
                         std::vector<double>& beta, std::vector<double>& evec, double& e0,
                         const LanczosWork* ws) {
  const uint64_t n = s.D;
  if (ws && s.dd > kFusedPartDoubles) return false;
  DevBuf rb, qb, sc, part;
  double *r, *q, *scd, *pd;
  if (ws) {
    r = ws->r;
    q = ws->q;
    scd = ws->sc;
    pd = ws->part;
  } else {
    rb.alloc(n * 8);
    qb.alloc(n * 8);
    sc.alloc(8 * sizeof(double));
    part.alloc((size_t)std::max<uint32_t>(s.dd, kMaxBlocks) * 8);
    r = rb.as<double>();
    q = qb.as<double>();
    scd = sc.as<double>();
    pd = part.as<double>();
  }
  QED_HIP(hipMemcpyAsync(r, psi, n * 8, hipMemcpyDeviceToDevice, stream()));
  if (ws) k_zero<<<stream_blocks(n), 256, 0, stream()>>>(q, n);
  else QED_HIP(hipMemsetAsync(q, 0, n * 8, stream()));
  const double init[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
  QED_HIP(hipMemcpyAsync(scd, init, sizeof(init), hipMemcpyHostToDevice, stream()));
  double ritz = 2.0 * accur;
  double ab[2];
  const uint32_t nb = red_blocks(n);
  while (ritz > accur) {
    if (!hv_lanczos(s, r, q, scd, pd)) {
      if (alpha.empty()) return false;
      fail("fused Lanczos product unavailable after the first step");
    }
    k_lzf_alpha<<<1, RB, 0, stream()>>>(pd, (int)s.dd, scd);
    k_lzf_axpy_nrm<<<nb, RB, 0, stream()>>>(scd, r, q, n, pd);
    k_lzf_beta<<<1, RB, 0, stream()>>>(pd, (int)nb, scd);
    QED_HIP(hipMemcpyAsync(ab, scd + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    std::swap(r, q);
    alpha.push_back(ab[0]);
    beta.push_back(ab[1]);
    tridiag_lowest(alpha, beta, &e0, evec);
    ritz = std::fabs(evec.back() * beta.back());
    if ((int)alpha.size() > max_iter)
      fail("Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations");
  }
  return true;
}

void lanczos_gs(qed_sector& s, double accur, int max_iter, double* psi, double* E0, int* niter,
                std::vector<double>* alpha_out, std::vector<double>* beta_out, const LanczosWork* ws) {
  const uint64_t n = s.D;
  // with a workspace: the two recurrence vectors and the reduction partials are the
  // caller's, zero fills are kernels, and the product is one that allocates nothing (the
  // split-halves form builds its tables on first use, so the default row kernel instead)
  const int var = ws && s.hv_variant >= 20 ? 12 : -1;
  std::vector<double> alpha, beta, evec;
  double e0 = 0.0;
  DevBuf rb, qb;
  double *r = nullptr, *q = nullptr;
  auto vectors = [&] {
    if (ws) {
      r = ws->r;
      q = ws->q;
    } else if (!rb.p) {
      rb.alloc(n * 8);
      qb.alloc(n * 8);
      r = rb.as<double>();
      q = qb.as<double>();
    }
  };
  auto reset = [&] {
    vectors();
    QED_HIP(hipMemcpyAsync(r, psi, n * 8, hipMemcpyDeviceToDevice, stream()));
    if (ws) k_zero<<<stream_blocks(n), 256, 0, stream()>>>(q, n);
    else QED_HIP(hipMemsetAsync(q, 0, n * 8, stream()));
  };
  if (var >= 0 || !lanczos_eigen_fused(s, accur, max_iter, psi, alpha, beta, evec, e0, ws)) {
    reset();
    double ritz = 2.0 * accur;
    while (ritz > accur) {
      if (!beta.empty()) scal2(1.0 / beta.back(), r, -beta.back(), q, n);
      hv(s, r, q, var);
      std::swap(r, q);
      double a = dot(q, r, n, ws ? ws->red : nullptr);
      alpha.push_back(a);
      beta.push_back(axpy_nrm2(-a, q, r, n, ws ? ws->red : nullptr));
      tridiag_lowest(alpha, beta, &e0, evec);
      ritz = std::fabs(evec.back() * beta.back());
      if ((int)alpha.size() > max_iter)
        fail("Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations");
    }
  }
  *E0 = e0;
  *niter = (int)beta.size();
  if (alpha_out) *alpha_out = alpha;
  if (beta_out) *beta_out = beta;
  // eigenvector replay from the saved start vector
  reset();
  scal(evec[0], psi, n);
  for (size_t j = 1; j < evec.size(); ++j) {
    hv(s, r, q, var);
    k_replay<<<stream_blocks(n), 256, 0, stream()>>>(alpha[j - 1], beta[j - 1], evec[j], r, q, psi, n);
  }
  QED_HIP(hipStreamSynchronize(stream()));
}

// ---------------------------------------------------------------------------------------
// band Lanczos (Freund), device vectors, following src_ed/Hamiltonian/Lanczos.hpp:214-395

struct VecPool {
  uint64_t n;
  std::vector<std::unique_ptr<DevBuf>> owned;
  std::vector<double*> free_list;
  explicit VecPool(uint64_t n_) : n(n_) {}
  double* get() {
    if (!free_list.empty()) {
      double* p = free_list.back();
      free_list.pop_back();
      return p;
    }
    owned.emplace_back(new DevBuf(n * sizeof(double)));
    return owned.back()->as<double>();
  }
  void put(double* p) {
    if (p) free_list.push_back(p);
  }
};

int band_lanczos(qed_sector& s, const double* phi, int p0, int M0, int max_iter_BL, double accur_deflation,
                 double accur_band_lanczos, std::vector<double>& evals, std::vector<double>& Q, double min_gap) {
  const uint64_t n = s.D;
  if (n == 0 || p0 == 0) return 0;
  if (M0 <= 0) M0 = (int)(14 * p0 * std::log(1.0 * n));
  if (M0 <= 1) fail("band Lanczos: too few iterations requested");
  int pc = p0;
  if (M0 > max_iter_BL) M0 = max_iter_BL;
  int M0p = (M0 / p0 + 1) * p0;
  M0 = M0p - pc;
  VecPool pool(n);
  std::vector<double*> v(M0p + 1, nullptr);
  std::vector<const double*> ph(p0);
  for (int i = 0; i < p0; ++i) ph[i] = phi + (uint64_t)i * n;
  for (int i = 0; i < p0; ++i) {
    v[i] = pool.get();
    QED_HIP(hipMemcpyAsync(v[i], ph[i], n * 8, hipMemcpyDeviceToDevice, stream()));
  }
  std::vector<double> t((size_t)M0 * M0, 0.0), P((size_t)p0 * M0, 0.0);
  auto T = [&](int i, int j) -> double& { return t[(size_t)j * M0 + i]; };
  std::vector<int> defl(M0, 0), I;
  double ev_old = 1e12;
  int j = 0;
  std::vector<double> tmp(kMaxMulti);
  while (j < M0) {
    // norm of v[j] and its overlaps with the start vectors in one pass
    std::vector<const double*> ys(p0 + 1);
    ys[0] = v[j];
    for (int k = 0; k < p0; ++k) ys[k + 1] = ph[k];
    std::vector<double> dd(p0 + 1);
    multi_dot(v[j], ys.data(), p0 + 1, n, dd.data());
    double z = std::sqrt(std::max(dd[0], 0.0));
    if (z < accur_deflation) {
      if (j >= pc) {
        I.push_back(j - pc);
        defl[j - pc] = 1;
      }
      pc--;
      if (pc == 0) break;
      pool.put(v[j]);
      for (int k = j; k < j + pc; ++k) v[k] = v[k + 1];
      v[j + pc] = nullptr;
      continue;
    }
    scal(1.0 / z, v[j], n);
    if (j >= pc) T(j, j - pc) = z;
    for (int k = 0; k < p0; ++k) P[(size_t)j * p0 + k] = dd[k + 1] / z;
    // orthogonalize the pending block against v[j]
    int nk = 0;
    std::vector<const double*> blk;
    std::vector<double*> blk_mut;
    for (int k = j + 1; k < j + pc; ++k) {
      blk.push_back(v[k]);
      blk_mut.push_back(v[k]);
      ++nk;
    }
    if (nk) {
      std::vector<double> zz(nk);
      multi_dot(v[j], blk.data(), nk, n, zz.data());
      multi_axpy_rev(zz.data(), v[j], blk_mut.data(), nk, n);
      for (int q = 0; q < nk; ++q) {
        int k = j + 1 + q;
        if (k >= pc) T(j, k - pc) = zz[q];
      }
    }
    double* w = pool.get();
    QED_HIP(hipMemsetAsync(w, 0, n * 8, stream()));
    hv(s, v[j], w);
    int k0 = 0;
    if (j > pc) {
      k0 = j - pc;
      if (defl[k0 - 1] == 0) {
        pool.put(v[k0 - 1]);
        v[k0 - 1] = nullptr;
      }
    }
    {
      std::vector<const double*> xs;
      std::vector<double> cs;
      for (int k = k0; k < j; ++k) {
        T(k, j) = T(j, k);
        xs.push_back(v[k]);
        cs.push_back(T(k, j));
      }
      multi_axpy(cs.data(), xs.data(), (int)xs.size(), w, n);
    }
    for (int k : I) {
      double ww = dot(v[k], w, n);
      axpy(-ww, v[k], w, n);
      T(k, j) = ww;
    }
    double ww = dot(v[j], w, n);
    axpy(-ww, v[j], w, n);
    T(j, j) = ww;
    for (int k : I) T(j, k) = T(k, j);
    v[j + pc] = w;
    if (j >= 4 * p0 && j % p0 == 0) {
      std::vector<double> Hpr((size_t)j * j), evs(j);
      for (int c = 0; c < j; ++c)
        for (int r = 0; r < j; ++r) Hpr[(size_t)c * j + r] = T(r, c);
      sym_eigen(j, Hpr.data(), evs.data(), false);
      double test = std::fabs(evs[0] - ev_old);
      ev_old = evs[0];
      // no_degenerate_BL (src_ed/Hamiltonian/Lanczos.hpp:368-370)
      if (min_gap > 0.0 && j > 1 && evs[1] - evs[0] < min_gap)
        fail("band Lanczos: the gap between the first two eigenvalues is smaller than " + std::to_string(min_gap));
      if (test < accur_band_lanczos) break;
    }
    ++j;
  }
  QED_HIP(hipStreamSynchronize(stream()));
  if (j == 0) return 0;
  std::vector<double> Hpr((size_t)j * j);
  for (int c = 0; c < j; ++c)
    for (int r = 0; r < j; ++r) Hpr[(size_t)c * j + r] = T(r, c);
  evals.assign(j, 0.0);
  sym_eigen(j, Hpr.data(), evals.data(), true);
  // Q = P U  (p0 x j)
  Q.assign((size_t)p0 * j, 0.0);
  for (int c = 0; c < j; ++c)
    for (int m = 0; m < j; ++m) {
      double u = Hpr[(size_t)c * j + m];
      if (u == 0.0) continue;
      for (int k = 0; k < p0; ++k) Q[(size_t)c * p0 + k] += P[(size_t)m * p0 + k] * u;
    }
  return j;
}

void lanczos_green(qed_sector& s, const double* psi_in, double accur_cf, int max_iter,
                   std::vector<double>& a, std::vector<double>& b) {
  const uint64_t n = s.D;
  DevBuf pb(n * 8), qb(n * 8);
  double* psi = pb.as<double>();
  double* q = qb.as<double>();
  QED_HIP(hipMemcpyAsync(psi, psi_in, n * 8, hipMemcpyDeviceToDevice, stream()));
  QED_HIP(hipMemsetAsync(q, 0, n * 8, stream()));
  a.clear();
  b.assign(1, 1.0);
  double prod = 1.0;
  while (prod > accur_cf && std::fabs(b.back()) > accur_cf && (int)a.size() < max_iter) {
    if (!a.empty()) k_swapscale<<<stream_blocks(n), 256, 0, stream()>>>(b.back(), psi, q, n);
    hv(s, psi, q);
    double z = dot(psi, q, n);
    a.push_back(z);
    double zz = axpy_nrm2(-z, psi, q, n);
    b.push_back(zz);
    prod *= std::fabs(zz);
  }
}


// ---------------------------------------------------------------------------------------
// Davidson-Liu for the K lowest states (src_ed/Hamiltonian/Davidson.hpp:38-230)

namespace {

// out_k[i] = sum_p C[p + m k] in_p[i], k < K (outputs distinct from inputs)
constexpr int kDavMax = 24;  // basis vectors of the Davidson subspace (3 K, K <= 8)

// out[k][i] = sum_p C[p + m k] in[p][i] for k < nout, with every input of index i read
// before any output of it is written: the outputs may be some of the inputs
__global__ __launch_bounds__(256) void k_dav_transform(PtrList in, int m, MutPtrList out, int nout,
                                                       const double* __restrict__ C, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double v[kDavMax];
#pragma unroll
    for (int p = 0; p < kDavMax; ++p) v[p] = p < m ? in.p[p][i] : 0.0;
    for (int k = 0; k < nout; ++k) {
      double acc = 0.0;
#pragma unroll
      for (int p = 0; p < kDavMax; ++p)
        if (p < m) acc += C[p + m * k] * v[p];
      out.p[k][i] = acc;
    }
  }
}

// the diagonal of H as the factorized format reports it (the interaction part)
__device__ __forceinline__ double diag_interaction(const HvView& v, uint64_t I) {
  const uint32_t iu = (uint32_t)(I % v.du), id = (uint32_t)(I / v.du);
  const uint32_t su = v.bup[iu], sd = v.bdn[id];
  if (v.hubbard_uniform) return v.hub_U * (double)__popc(su & sd & v.hub_sites);
  double e = 0.0;
  const uint64_t w = (uint64_t)sd | ((uint64_t)su << 32);
  for (int k = 0; k < v.n_int; ++k)
    if ((w & v.int_mask[k]) == v.int_mask[k]) e += v.int_val[k];
  return e;
}

// the diagonal Davidson reads: the factorized format's (interaction part, on the fly) or a
// stored one (the symmetry-adapted and mixed-number sectors: the true diagonal of H, as the
// reference's other formats report it)
struct DavDiag {
  HvView v;
  const double* d;  // stored diagonal, or null
  __device__ __forceinline__ double operator()(uint64_t I) const { return d ? d[I] : diag_interaction(v, I); }
};

// correction vector of one unconverged root: c = (H x - mu x) / (mu - d + tiny)
__global__ void k_dav_correction(DavDiag dg, const double* __restrict__ hx, const double* __restrict__ x, double mu,
                                 double tiny, double* __restrict__ c, uint64_t n) {
  for (uint64_t I = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n;
       I += (uint64_t)gridDim.x * blockDim.x)
    c[I] = (hx[I] - mu * x[I]) / (-dg(I) + mu + tiny);
}

// the order of find_lowest (src_util/vector_num.hpp:242-264): ascending values, and among
// equal values the later index first
__device__ __forceinline__ bool dav_before(double va, uint64_t ia, double vb, uint64_t ib) {
  return va < vb || (va == vb && ia > ib);
}

// block partials of the first diagonal entry after (pv, pi) in that order (all entries
// when have_prev is 0); a block with no candidate reports +inf
__global__ __launch_bounds__(256) void k_dav_next_lowest(DavDiag dg, uint64_t n, int have_prev, double pv, uint64_t pi,
                                                         double* __restrict__ part_v, uint64_t* __restrict__ part_i) {
  __shared__ double sv[256];
  __shared__ uint64_t si[256];
  double bv = INFINITY;
  uint64_t bi = 0;
  for (uint64_t I = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n;
       I += (uint64_t)gridDim.x * blockDim.x) {
    const double d = dg(I);
    if (have_prev && !dav_before(pv, pi, d, I)) continue;
    if (dav_before(d, I, bv, bi)) {
      bv = d;
      bi = I;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && dav_before(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + w];
      si[threadIdx.x] = si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_v[blockIdx.x] = sv[0];
    part_i[blockIdx.x] = si[0];
  }
}

// the K lowest diagonal entries in find_lowest's order, one device pass per entry
std::vector<uint64_t> lowest_diagonal(const qed_sector& s, const DavDiag& dg, int K) {
  const uint64_t n = s.D;
  const uint32_t nb = (uint32_t)std::min<uint64_t>(kMaxBlocks, std::max<uint64_t>(1, (n + 255) / 256));
  DevBuf pv(nb * sizeof(double)), pi(nb * sizeof(uint64_t));
  std::vector<double> hv_(nb);
  std::vector<uint64_t> hi_(nb);
  std::vector<uint64_t> out;
  double lv = 0.0;
  uint64_t li = 0;
  for (int k = 0; k < K; ++k) {
    k_dav_next_lowest<<<nb, 256, 0, stream()>>>(dg, n, k > 0, lv, li, pv.as<double>(), pi.as<uint64_t>());
    QED_HIP(hipMemcpyAsync(hv_.data(), pv.p, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipMemcpyAsync(hi_.data(), pi.p, nb * sizeof(uint64_t), hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    double bv = INFINITY;
    uint64_t bi = 0;
    for (uint32_t b = 0; b < nb; ++b)
      if (std::isfinite(hv_[b]) && (bv == INFINITY || hv_[b] < bv || (hv_[b] == bv && hi_[b] > bi))) {
        bv = hv_[b];
        bi = hi_[b];
      }
    if (!std::isfinite(bv)) fail("Davidson: fewer diagonal entries than requested states");
    out.push_back(bi);
    lv = bv;
    li = bi;
  }
  return out;
}

// out = C-weighted combinations of `in` (see k_dav_transform); C is m x nout column-major
void dav_transform(const std::vector<double*>& in, const std::vector<double>& C, const std::vector<double*>& out,
                   uint64_t n) {
  const int m = (int)in.size(), nout = (int)out.size();
  if (m > kDavMax || nout > kDavMax) fail("Davidson: subspace too large");
  PtrList li{};
  MutPtrList lo{};
  for (int p = 0; p < m; ++p) li.p[p] = in[p];
  for (int k = 0; k < nout; ++k) lo.p[k] = out[k];
  double* cd = red_scratch().as<double>() + (size_t)kMaxBlocks * kMaxMulti + 64;
  QED_HIP(hipMemcpyAsync(cd, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice, stream()));
  k_dav_transform<<<stream_blocks(n), 256, 0, stream()>>>(li, m, lo, nout, cd, n);
  QED_HIP(hipStreamSynchronize(stream()));
}

}  // namespace

// Davidson-Liu for the K lowest states, following src_ed/Hamiltonian/Davidson.hpp:38-230
// (start vectors on the K lowest diagonal entries, at most 3K basis vectors, the new basis
// made of the Ritz vectors, the basis vectors of the unconverged roots and their
// preconditioned residuals, Gram-Schmidt), with a memory layout for large sectors: the
// products H b are not kept (each is formed in a scratch vector, reduced against the basis
// and dropped), the residuals H x - mu x are formed from a fresh product of the Ritz vector,
// and the basis is transformed in place. Device memory: the K output vectors (the first K
// basis slots) plus 2K further slots and two scratch vectors, 3K + 2 vectors in all (14 for
// K = 4) instead of the reference's ~9K. The diagonal (preconditioner and start vectors) is
// recomputed per entry on factorized sectors; symmetry-adapted and mixed-number sectors
// store it (one more vector), since their diagonal gathers over the orbit's hops.
void davidson(qed_sector& s, int K, double accur, int max_iter, std::vector<double>& val, double* xout,
              int* niter, bool verb) {
  const uint64_t n = s.D;
  const double tiny = 1e-14;
  if ((uint64_t)K > n) K = (int)n;
  if (K < 1) fail("Davidson: no state requested");
  const int J = (int)std::min<uint64_t>(3 * (uint64_t)K, n);
  if (J > kDavMax) fail("Davidson: at most " + std::to_string(kDavMax / 3) + " states");
  val.assign(K, 0.0);
  {
    size_t fr = 0, tot = 0;
    QED_HIP(hipMemGetInfo(&fr, &tot));
    const double need = (double)(J - K + 2 + (s.factorized() ? 0 : 1)) * n * 8;
    if (need > (double)fr)
      fail("Davidson: " + std::to_string(J - K + 2) + " work vectors of the sector need " +
           std::to_string(need / 1e9) + " GB of device memory, " + std::to_string(fr / 1e9) + " GB free");
  }
  std::vector<DevBuf> own;
  std::vector<double*> slot(J);
  for (int k = 0; k < K; ++k) slot[k] = xout + (uint64_t)k * n;
  for (int k = K; k < J; ++k) {
    own.emplace_back(n * 8);
    slot[k] = own.back().as<double>();
  }
  DevBuf tb(n * 8), ub(n * 8);
  double* t = tb.as<double>();
  double* u = ub.as<double>();
  const uint32_t sb = stream_blocks(n);

  if (verb)
    printf("\nDavidson method for %d vectors in sector (%d, %d), dim = %llu\n", K, s.nup, s.ndn,
           (unsigned long long)n);
  DevBuf dstore;
  DavDiag dg{s.factorized() ? view(s) : HvView{}, nullptr};
  if (!s.factorized()) {
    dstore.alloc(n * 8);
    k_zero<<<sb, 256, 0, stream()>>>(dstore.as<double>(), n);
    sector_diagonal(s, dstore.as<double>());
    dg.d = dstore.as<double>();
  }
  const std::vector<uint64_t> low = lowest_diagonal(s, dg, K);
  const double one = 1.0;
  for (int k = 0; k < K; ++k) {
    k_zero<<<sb, 256, 0, stream()>>>(slot[k], n);
    QED_HIP(hipMemcpyAsync(slot[k] + low[k], &one, 8, hipMemcpyHostToDevice, stream()));
  }
  int m = K;
  // bhb(i, j) = (H b_j)[low_i]
  std::vector<double> bhb((size_t)m * m);
  for (int j = 0; j < m; ++j) {
    k_zero<<<sb, 256, 0, stream()>>>(u, n);
    hv(s, slot[j], u);
    for (int i = 0; i < m; ++i)
      QED_HIP(hipMemcpyAsync(&bhb[(size_t)j * m + i], u + low[i], 8, hipMemcpyDeviceToHost, stream()));
  }
  QED_HIP(hipStreamSynchronize(stream()));

  int iter;
  for (iter = 0; iter < max_iter; ++iter) {
    std::vector<double> a(bhb), mu(m);
    sym_eigen(m, a.data(), mu.data(), true);  // column k of a: eigenvector k
    std::vector<double*> basis(slot.begin(), slot.begin() + m);
    // residual norms |H x_k - mu_k x_k| with x_k = sum_p a_pk b_p formed in t
    std::vector<int> unconv;
    double res_last = 0.0;
    for (int k = 0; k < K; ++k) {
      std::vector<double> ck(a.begin() + (size_t)k * m, a.begin() + (size_t)(k + 1) * m);
      dav_transform(basis, ck, {t}, n);
      k_zero<<<sb, 256, 0, stream()>>>(u, n);
      hv(s, t, u);
      res_last = axpy_nrm2(-mu[k], t, u, n);
      if (res_last > accur) unconv.push_back(k);
    }
    if (verb && iter % 10 == 0) {
      printf("--> iteration %d, evalue = %.12g,\tresidual = %.12g\n", iter, mu[0], res_last);
      fflush(stdout);
    }
    for (int k = 0; k < K; ++k) val[k] = mu[k];
    const int Q = (int)unconv.size();
    // the basis in place: slots 0..K-1 <- Ritz vectors, then the unconverged roots' own
    // basis vectors
    std::vector<double> C((size_t)m * (K + Q), 0.0);
    for (int k = 0; k < K; ++k)
      for (int p = 0; p < m; ++p) C[p + (size_t)m * k] = a[(size_t)k * m + p];
    for (int q = 0; q < Q; ++q) C[unconv[q] + (size_t)m * (K + q)] = 1.0;
    dav_transform(basis, C, std::vector<double*>(slot.begin(), slot.begin() + K + Q), n);
    if (Q == 0) break;
    // preconditioned residuals of the unconverged roots
    for (int q = 0; q < Q; ++q) {
      const int k = unconv[q];
      k_zero<<<sb, 256, 0, stream()>>>(u, n);
      hv(s, slot[k], u);
      k_dav_correction<<<sb, 256, 0, stream()>>>(dg, u, slot[k], mu[k], tiny, slot[K + Q + q], n);
    }
    // Gram-Schmidt in the reference's order (vectors below tiny are removed)
    std::vector<double*> b(slot.begin(), slot.begin() + K + 2 * Q);
    for (size_t i = 0; i < b.size(); ++i) {
      const double z = nrm2(b[i], n);
      if (z < tiny) {
        std::swap(b[i], b.back());
        b.pop_back();
        --i;
        continue;
      }
      scal(1.0 / z, b[i], n);
      const int cnt = (int)b.size() - (int)(i + 1);
      if (cnt > 0) {
        std::vector<const double*> ys(b.begin() + i + 1, b.end());
        std::vector<double> zz(cnt);
        multi_dot(b[i], ys.data(), cnt, n, zz.data());
        std::vector<double*> ym(b.begin() + i + 1, b.end());
        multi_axpy_rev(zz.data(), b[i], ym.data(), cnt, n);
      }
    }
    // keep the slots in basis order (a removed vector's slot moves to the end)
    {
      std::vector<double*> rest;
      for (double* p : slot)
        if (std::find(b.begin(), b.end(), p) == b.end()) rest.push_back(p);
      for (size_t i = 0; i < b.size(); ++i) slot[i] = b[i];
      for (size_t i = 0; i < rest.size(); ++i) slot[b.size() + i] = rest[i];
    }
    if ((int)b.size() == K) break;
    m = (int)b.size();
    // bhb(i, k) = <b_i, H b_k> for i >= k (the reference's pairing), one product at a time
    bhb.assign((size_t)m * m, 0.0);
    for (int k = 0; k < m; ++k) {
      k_zero<<<sb, 256, 0, stream()>>>(u, n);
      hv(s, slot[k], u);
      std::vector<const double*> ys(slot.begin() + k, slot.begin() + m);
      std::vector<double> zz(m - k);
      multi_dot(u, ys.data(), m - k, n, zz.data());
      for (int i = k; i < m; ++i) {
        bhb[(size_t)k * m + i] = zz[i - k];
        bhb[(size_t)i * m + k] = zz[i - k];
      }
    }
  }
  if (iter == max_iter) fail("Davidson method failed to converge!");
  *niter = iter;
  // the states are in the first K slots, which are the output vectors unless Gram-Schmidt
  // removed one of them: then exchange contents (through t) until each is in place
  for (int k = 0; k < K; ++k) {
    double* want = xout + (uint64_t)k * n;
    if (slot[k] == want) continue;
    const int j = (int)(std::find(slot.begin(), slot.end(), want) - slot.begin());
    QED_HIP(hipMemcpyAsync(t, want, n * 8, hipMemcpyDeviceToDevice, stream()));
    QED_HIP(hipMemcpyAsync(want, slot[k], n * 8, hipMemcpyDeviceToDevice, stream()));
    QED_HIP(hipMemcpyAsync(slot[k], t, n * 8, hipMemcpyDeviceToDevice, stream()));
    std::swap(slot[k], slot[j]);
  }
  QED_HIP(hipStreamSynchronize(stream()));
}

}  // namespace qed

using namespace qed;

// ---- the allocation-free Lanczos ground state --------------------------------------------
namespace {

// Device memory of qed_lanczos_gs_noalloc, one copy per GPU for the life of the process:
// the two recurrence vectors of a sector of at most kNoAllocMaxDim amplitudes, the BLAS-1
// block partials, the fused step's scalars and row partials (34.7 MB in all).
constexpr uint64_t kNoAllocMaxDim = 1ull << 21;
constexpr size_t kNoAllocDoubles = 2 * kNoAllocMaxDim + kRedDoubles + 8 + kFusedPartDoubles;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current GPU, looked up on the first call there
// (the name is a host-side stand-in and must not be passed to a kernel). Caller holds the
// library mutex.
double* noalloc_scratch() {
  int dev = 0;
  QED_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail("qed_lanczos_gs_noalloc: device index out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    QED_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail("qed_lanczos_gs_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int qed_lanczos_gs(qed_sector* s, double accur, int max_iter, double* psi, double* E0, int* niter,
                   double* alpha, double* beta) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<double> a, b;
    lanczos_gs(*s, accur, max_iter, psi, E0, niter, &a, &b);
    if (alpha) std::copy(a.begin(), a.end(), alpha);
    if (beta) std::copy(b.begin(), b.end(), beta);
  });
}

int qed_lanczos_gs_noalloc(qed_sector* s, double accur, int max_iter, double* psi, double* E0, int* niter,
                           double* alpha, double* beta) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (s->D > kNoAllocMaxDim)
      fail("qed_lanczos_gs_noalloc: sector dimension " + std::to_string(s->D) + " exceeds the scratch limit of " +
           std::to_string(kNoAllocMaxDim));
    double* base = noalloc_scratch();
    const LanczosWork ws{base, base + kNoAllocMaxDim, base + 2 * kNoAllocMaxDim,
                         base + 2 * kNoAllocMaxDim + kRedDoubles, base + 2 * kNoAllocMaxDim + kRedDoubles + 8};
    std::vector<double> a, b;
    lanczos_gs(*s, accur, max_iter, psi, E0, niter, &a, &b, &ws);
    if (alpha) std::copy(a.begin(), a.end(), alpha);
    if (beta) std::copy(b.begin(), b.end(), beta);
  });
}

// Measurement: average milliseconds per Lanczos step of the eigenvalue pass from psi
// (normalized), `steps` steps each (no stopping test), HIP events on the library stream:
// ms[0] the fused step (product with the scale and dot folded in, then axpy + norm),
// ms[1] the separate passes (scale, product, dot, axpy + norm) with their host round
// trips. ms[0] = -1 when the sector does not take the fused product.
int qed_lanczos_steps_timed(qed_sector* s, const double* psi, int steps, double* ms) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    const uint64_t n = s->D;
    DevBuf rb(n * 8), qb(n * 8), sc(8 * sizeof(double)), part((size_t)std::max<uint32_t>(s->dd, kMaxBlocks) * 8);
    double* r = rb.as<double>();
    double* q = qb.as<double>();
    double* scd = sc.as<double>();
    double* pd = part.as<double>();
    hipEvent_t e0, e1;
    QED_HIP(hipEventCreate(&e0));
    QED_HIP(hipEventCreate(&e1));
    const uint32_t nb = red_blocks(n);
    auto reset = [&] {
      QED_HIP(hipMemcpyAsync(r, psi, n * 8, hipMemcpyDeviceToDevice, stream()));
      QED_HIP(hipMemsetAsync(q, 0, n * 8, stream()));
      const double init[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
      QED_HIP(hipMemcpyAsync(scd, init, sizeof(init), hipMemcpyHostToDevice, stream()));
      QED_HIP(hipStreamSynchronize(stream()));
    };
    ms[0] = -1.0;
    reset();
    if (hv_lanczos(*s, r, q, scd, pd)) {
      reset();
      QED_HIP(hipEventRecord(e0, stream()));
      double ab[2];
      for (int k = 0; k < steps; ++k) {
        hv_lanczos(*s, r, q, scd, pd);
        k_lzf_alpha<<<1, RB, 0, stream()>>>(pd, (int)s->dd, scd);
        k_lzf_axpy_nrm<<<nb, RB, 0, stream()>>>(scd, r, q, n, pd);
        k_lzf_beta<<<1, RB, 0, stream()>>>(pd, (int)nb, scd);
        QED_HIP(hipMemcpyAsync(ab, scd + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
        QED_HIP(hipStreamSynchronize(stream()));
        std::swap(r, q);
      }
      QED_HIP(hipEventRecord(e1, stream()));
      QED_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      QED_HIP(hipEventElapsedTime(&t, e0, e1));
      ms[0] = t / std::max(1, steps);
    }
    reset();
    QED_HIP(hipEventRecord(e0, stream()));
    double bprev = 0.0;
    for (int k = 0; k < steps; ++k) {
      if (k > 0) scal2(1.0 / bprev, r, -bprev, q, n);
      hv(*s, r, q);
      std::swap(r, q);
      const double a = dot(q, r, n);
      bprev = axpy_nrm2(-a, q, r, n);
    }
    QED_HIP(hipEventRecord(e1, stream()));
    QED_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QED_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[1] = t / std::max(1, steps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

int qed_lanczos_green(qed_sector* s, const double* psi, double accur_cf, int max_iter, double* a,
                      double* b, int* n) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<double> av, bv;
    lanczos_green(*s, psi, accur_cf, max_iter, av, bv);
    std::copy(av.begin(), av.end(), a);
    std::copy(bv.begin(), bv.end(), b);
    *n = (int)av.size();
  });
}

int qed_band_lanczos(qed_sector* s, const double* phi, int p, int M0, int max_iter_BL,
                     double accur_deflation, double accur_band_lanczos, double* evals, double* Q,
                     int cap, int* M) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<double> e, q;
    int m = band_lanczos(*s, phi, p, M0, max_iter_BL, accur_deflation, accur_band_lanczos, e, q);
    if (m > cap) fail("band Lanczos: output capacity too small");
    *M = m;
    std::copy(e.begin(), e.end(), evals);
    std::copy(q.begin(), q.end(), Q);
  });
}

int qed_davidson(qed_sector* s, int k, double accur, int max_iter, double* evals, double* psi,
                 int* niter) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<double> e;
    davidson(*s, k, accur, max_iter, e, psi, niter, verbose_ed());
    std::copy(e.begin(), e.end(), evals);
  });
}

// Achievable read-modify-write bandwidth: reps launches of y += 0.5 x over n doubles (n
// even, 16-byte aligned vectors), timed with events; *ms = milliseconds per launch.
int qed_stream_rmw_timed(const double* x, double* y, uint64_t n, int reps, double* ms) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (n % 2) fail("stream: n must be even");
    const uint64_t n2 = n / 2;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n2 + 255) / 256, (uint64_t)device_cu_count() * 64);
    auto run = [&] {
      k_stream_rmw<<<grid, 256, 0, stream()>>>(reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), n2);
    };
    run();
    hipEvent_t e0, e1;
    QED_HIP(hipEventCreate(&e0));
    QED_HIP(hipEventCreate(&e1));
    QED_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < std::max(1, reps); ++r) run();
    QED_HIP(hipEventRecord(e1, stream()));
    QED_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QED_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    QED_HIP(hipGetLastError());
    *ms = (double)t / std::max(1, reps);
  });
}

int qed_blas_dot(const double* x, const double* y, uint64_t n, double* out) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { *out = dot(x, y, n); });
}

int qed_util_tridiag_eigen(int n, double* d, const double* e, double* z) {
  return guard([&] { tridiag_eigen(n, d, e, z); });
}

int qed_util_sym_eigen(int n, double* a, double* w, int vectors) {
  return guard([&] { sym_eigen(n, a, w, vectors != 0); });
}

}  // extern "C"
