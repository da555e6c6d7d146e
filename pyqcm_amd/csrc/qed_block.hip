// Block Lanczos on p device vectors and the matrix continued fraction of the Green
// function built from it (GF_method 'M').
//
// The recurrence H Q_j = Q_{j-1} B_{j-1}^T + Q_j A_j + Q_{j+1} B_j runs on blocks of p
// vectors of one sector: p H.v products, then one pass each for the p x p overlaps
// (Gram kernel: every vector of both blocks read once per group of 4 columns), the
// three-term update and, in the polar form, the change of basis Q_{j+1} = Z B_j^-1. The
// p x p blocks and the lowest eigenvalue of the block-tridiagonal projection (banded
// inverse iteration, dense below a few hundred rows) stay on the host in FP64.
//
// The continued fraction G(z) = W^T F_0 W, F_M = 0, F_j = (z - A_j - B_j^T F_{j+1} B_j)^-1
// is evaluated on the device: one workgroup per frequency, the p x p complex blocks in LDS,
// Gauss-Jordan inversion with partial pivoting.
#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/qed.h"
#include "qed_linalg.hpp"
#include "qed_sector.hpp"

namespace qed {

namespace {

constexpr int kMaxBlock = 24;  // block size limit (LDS of the continued-fraction kernel)
constexpr int kLG = 4;         // Gram columns per workgroup row
constexpr int RB = 256;
constexpr uint32_t kMaxParts = 1024;

struct Blk {
  const double* p[kMaxBlock];
};
struct MutBlk {
  double* p[kMaxBlock];
};

uint32_t part_blocks(uint64_t n) {
  uint64_t b = (n + (uint64_t)RB * 8 - 1) / ((uint64_t)RB * 8);
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), kMaxParts);
}
uint32_t stream_grid(uint64_t n) {
  uint64_t b = (n + RB - 1) / RB;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), (uint64_t)device_cu_count() * 16);
}

// part[(k + p l) nb + b] = partial <X_k, Y_l> over this workgroup's elements, for k < p
// and the kLG columns l = kLG blockIdx.y + c
template <int PM>
__global__ __launch_bounds__(RB) void k_block_gram(Blk X, Blk Y, int p, uint64_t n, double* __restrict__ part) {
  __shared__ double sh[RB / 64];
  const int l0 = blockIdx.y * kLG;
  double acc[PM][kLG];
#pragma unroll
  for (int k = 0; k < PM; ++k)
#pragma unroll
    for (int c = 0; c < kLG; ++c) acc[k][c] = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
    double y[kLG];
#pragma unroll
    for (int c = 0; c < kLG; ++c) y[c] = (l0 + c < p) ? Y.p[l0 + c][i] : 0.0;
#pragma unroll
    for (int k = 0; k < PM; ++k)
      if (k < p) {
        const double x = X.p[k][i];
#pragma unroll
        for (int c = 0; c < kLG; ++c) acc[k][c] += x * y[c];
      }
  }
#pragma unroll
  for (int k = 0; k < PM; ++k)
#pragma unroll
    for (int c = 0; c < kLG; ++c)
      if (k < p && l0 + c < p) {
        const double s = block_sum(acc[k][c], sh);
        if (threadIdx.x == 0) part[(size_t)(k + p * (l0 + c)) * gridDim.x + blockIdx.x] = s;
      }
}

// out[j] = sum_b part[j nb + b], one workgroup per entry
__global__ __launch_bounds__(RB) void k_block_finish(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[RB / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) s += part[(size_t)blockIdx.x * nb + i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Z_l -= sum_k C1(k, l) Q_k, then -= sum_k C2(k, l) P_k (when np > 0), in the order of the
// sequential updates (C1, C2 column-major p x p)
template <int PM>
__global__ __launch_bounds__(RB) void k_block_update(Blk Q, Blk P, MutBlk Z, int p, int np,
                                                      const double* __restrict__ C1,
                                                      const double* __restrict__ C2, uint64_t n) {
  __shared__ double c1[PM * PM], c2[PM * PM];
  for (int t = threadIdx.x; t < p * p; t += RB) {
    c1[t] = C1[t];
    c2[t] = np ? C2[t] : 0.0;
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
    double q[PM], r[PM];
#pragma unroll
    for (int k = 0; k < PM; ++k) {
      q[k] = (k < p) ? Q.p[k][i] : 0.0;
      r[k] = (k < np) ? P.p[k][i] : 0.0;
    }
#pragma unroll
    for (int l = 0; l < PM; ++l)
      if (l < p) {
        double v = Z.p[l][i];
#pragma unroll
        for (int k = 0; k < PM; ++k)
          if (k < p) v -= c1[k + p * l] * q[k];
#pragma unroll
        for (int k = 0; k < PM; ++k)
          if (k < np) v -= c2[k + p * l] * r[k];
        Z.p[l][i] = v;
      }
  }
}

// O_l = sum_k C(k, l) Z_k (outputs distinct from inputs)
template <int PM>
__global__ __launch_bounds__(RB) void k_block_combine(Blk Z, MutBlk O, int p, const double* __restrict__ C,
                                                       uint64_t n) {
  __shared__ double c[PM * PM];
  for (int t = threadIdx.x; t < p * p; t += RB) c[t] = C[t];
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
    double z[PM];
#pragma unroll
    for (int k = 0; k < PM; ++k) z[k] = (k < p) ? Z.p[k][i] : 0.0;
#pragma unroll
    for (int l = 0; l < PM; ++l)
      if (l < p) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < PM; ++k)
          if (k < p) v += c[k + p * l] * z[k];
        O.p[l][i] = v;
      }
  }
}

// device scratch: Gram partials, results and coefficient blocks (calls are serialized
// by the library mutex)
DevBuf& bscratch() {
  static DevBuf b(((size_t)kMaxParts * kMaxBlock * kMaxBlock + 4 * kMaxBlock * kMaxBlock) * sizeof(double));
  return b;
}

template <class F>
void by_width(int p, F&& f) {
  if (p <= 4)
    f(std::integral_constant<int, 4>());
  else if (p <= 8)
    f(std::integral_constant<int, 8>());
  else if (p <= 16)
    f(std::integral_constant<int, 16>());
  else
    f(std::integral_constant<int, kMaxBlock>());
}

// G(k, l) = <X_k, Y_l> (column-major p x p, host)
void block_gram(const std::vector<double*>& X, const std::vector<double*>& Y, int p, uint64_t n,
                std::vector<double>& G) {
  Blk bx{}, by{};
  for (int k = 0; k < p; ++k) {
    bx.p[k] = X[k];
    by.p[k] = Y[k];
  }
  const uint32_t nb = part_blocks(n);
  double* part = bscratch().as<double>();
  double* out = part + (size_t)kMaxParts * kMaxBlock * kMaxBlock;
  dim3 grid(nb, (p + kLG - 1) / kLG);
  by_width(p, [&](auto W) { k_block_gram<decltype(W)::value><<<grid, RB, 0, stream()>>>(bx, by, p, n, part); });
  k_block_finish<<<p * p, RB, 0, stream()>>>(part, nb, out);
  G.resize((size_t)p * p);
  QED_HIP(hipMemcpyAsync(G.data(), out, (size_t)p * p * sizeof(double), hipMemcpyDeviceToHost, stream()));
  QED_HIP(hipStreamSynchronize(stream()));
}

void block_update(const std::vector<double*>& Q, const std::vector<double*>* P, const std::vector<double*>& Z,
                  int p, const std::vector<double>& C1, const std::vector<double>* C2, uint64_t n) {
  Blk bq{}, bp{};
  MutBlk bz{};
  for (int k = 0; k < p; ++k) {
    bq.p[k] = Q[k];
    bz.p[k] = Z[k];
    if (P) bp.p[k] = (*P)[k];
  }
  double* c = bscratch().as<double>() + (size_t)kMaxParts * kMaxBlock * kMaxBlock + kMaxBlock * kMaxBlock;
  QED_HIP(hipMemcpyAsync(c, C1.data(), (size_t)p * p * sizeof(double), hipMemcpyHostToDevice, stream()));
  if (C2)
    QED_HIP(hipMemcpyAsync(c + kMaxBlock * kMaxBlock, C2->data(), (size_t)p * p * sizeof(double),
                           hipMemcpyHostToDevice, stream()));
  const int np = P ? p : 0;
  by_width(p, [&](auto W) {
    k_block_update<decltype(W)::value><<<stream_grid(n), RB, 0, stream()>>>(bq, bp, bz, p, np, c,
                                                                            c + kMaxBlock * kMaxBlock, n);
  });
  QED_HIP(hipStreamSynchronize(stream()));
}

void block_combine(const std::vector<double*>& Z, const std::vector<double*>& O, int p, const std::vector<double>& C,
                   uint64_t n) {
  Blk bz{};
  MutBlk bo{};
  for (int k = 0; k < p; ++k) {
    bz.p[k] = Z[k];
    bo.p[k] = O[k];
  }
  double* c = bscratch().as<double>() + (size_t)kMaxParts * kMaxBlock * kMaxBlock + 3 * kMaxBlock * kMaxBlock;
  QED_HIP(hipMemcpyAsync(c, C.data(), (size_t)p * p * sizeof(double), hipMemcpyHostToDevice, stream()));
  by_width(p, [&](auto W) {
    k_block_combine<decltype(W)::value><<<stream_grid(n), RB, 0, stream()>>>(bz, bo, p, c, n);
  });
  QED_HIP(hipStreamSynchronize(stream()));
}

// lowest eigenvalue of the block-tridiagonal projection with floors 0..j (diagonal blocks
// A, sub-diagonal blocks B_jj at rows (jj+1) p, columns jj p); half-bandwidth 2p - 1
struct BlockTridiag {
  const std::vector<std::vector<double>>& A;
  const std::vector<std::vector<double>>& B;
  int p;
  double operator()(int r, int c) const {
    const int br = r / p, bc = c / p, ir = r % p, ic = c % p;
    if (br == bc) return A[br][ir + (size_t)p * ic];
    if (br == bc + 1) return B[bc][ir + (size_t)p * ic];
    if (bc == br + 1) return B[br][ic + (size_t)p * ir];
    return 0.0;
  }
};

}  // namespace

// Block Lanczos (src_ed/Hamiltonian/Lanczos.hpp:426-603 with QR blocks, :604-779 with
// polar blocks). q0: p device vectors (rows of p x D), orthonormalized again here as the
// reference does. A, B receive the floors (p x p column-major each).
int block_lanczos(qed_sector& s, const double* q0, int p, int M0, int max_iter_BL, double accur_deflation,
                  double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                  std::vector<std::vector<double>>& B) {
  return block_lanczos_op(
      s.D, [&](const double* x, double* y) { hv(s, x, y); }, q0, p, M0, max_iter_BL, accur_deflation,
      accur_band_lanczos, qr, A, B);
}

int block_lanczos_op(uint64_t n, const BlockOperator& op, const double* q0, int p, int M0, int max_iter_BL,
                     double accur_deflation, double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                     std::vector<std::vector<double>>& B) {
  A.clear();
  B.clear();
  if (n == 0 || p == 0) return 0;
  if (p > kMaxBlock) fail("block Lanczos: block size above " + std::to_string(kMaxBlock));
  if (accur_band_lanczos < 0.0) accur_band_lanczos = 0.0;
  if (M0 <= 0) M0 = (int)(14 * p * std::log(1.0 * n));
  if (M0 > max_iter_BL) M0 = max_iter_BL;
  const int nblk = qr ? 3 : 4;
  DevBuf store((size_t)nblk * p * n * sizeof(double));
  std::vector<std::vector<double*>> blk(nblk, std::vector<double*>(p));
  for (int b = 0; b < nblk; ++b)
    for (int k = 0; k < p; ++k) blk[b][k] = store.as<double>() + ((size_t)b * p + k) * n;
  std::vector<double*> Qp = blk[0], Qc = blk[1], Z = blk[2], Qn = qr ? blk[2] : blk[3];
  QED_HIP(hipMemcpyAsync(Qc[0], q0, (size_t)p * n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  for (int l = 0; l < p; ++l) {
    for (int k = 0; k < l; ++k) axpy(-dot(Qc[k], Qc[l], n), Qc[k], Qc[l], n);
    const double nrm = nrm2(Qc[l], n);
    if (nrm < accur_deflation)
      fail("blockLanczos: starting block is rank-deficient at column " + std::to_string(l));
    scal(1.0 / nrm, Qc[l], n);
  }
  double ev_old = 1e12, ev_old2 = 1e12;
  std::vector<double> evec;
  std::vector<double> G;
  for (int j = 0; j < M0; ++j) {
    for (int k = 0; k < p; ++k) {
      QED_HIP(hipMemsetAsync(Z[k], 0, n * sizeof(double), stream()));
      op(Qc[k], Z[k]);
    }
    block_gram(Qc, Z, p, n, G);  // G(k, l) = <Q_k, Z_l>
    std::vector<double> Aj((size_t)p * p);
    for (int l = 0; l < p; ++l)
      for (int k = 0; k <= l; ++k) Aj[k + (size_t)p * l] = Aj[l + (size_t)p * k] = G[k + (size_t)p * l];
    A.push_back(Aj);
    if (j > 0) {
      // Z_l -= sum_k B_{j-1}(l, k) P_k: coefficient (k, l) of the update is B_{j-1}(l, k)
      const std::vector<double>& Bp = B.back();
      std::vector<double> C2((size_t)p * p);
      for (int l = 0; l < p; ++l)
        for (int k = 0; k < p; ++k) C2[k + (size_t)p * l] = Bp[l + (size_t)p * k];
      block_update(Qc, &Qp, Z, p, Aj, &C2, n);
    } else {
      block_update(Qc, nullptr, Z, p, Aj, nullptr, n);
    }
    std::vector<double> Bj((size_t)p * p, 0.0);
    bool singular = false;
    if (qr) {
      // modified Gram-Schmidt of the residual block: Z = Q_{j+1} B_j
      for (int l = 0; l < p; ++l) {
        for (int k = 0; k < l; ++k) {
          const double z = dot(Z[k], Z[l], n);
          Bj[k + (size_t)p * l] = z;
          axpy(-z, Z[k], Z[l], n);
        }
        const double nrm = nrm2(Z[l], n);
        if (nrm < accur_deflation) {
          for (int ll = l; ll < p; ++ll) QED_HIP(hipMemsetAsync(Z[ll], 0, n * sizeof(double), stream()));
          singular = true;
          break;
        }
        Bj[l + (size_t)p * l] = nrm;
        scal(1.0 / nrm, Z[l], n);
      }
      B.push_back(Bj);
    } else {
      // polar decomposition: Gram = V D V^T, B_j = V D^1/2 V^T, Q_{j+1} = Z V D^-1/2 V^T
      block_gram(Z, Z, p, n, G);
      std::vector<double> Gs((size_t)p * p), D(p);
      for (int l = 0; l < p; ++l)
        for (int k = 0; k <= l; ++k) Gs[k + (size_t)p * l] = Gs[l + (size_t)p * k] = G[k + (size_t)p * l];
      sym_eigen(p, Gs.data(), D.data(), true);
      for (int m = 0; m < p; ++m)
        if (D[m] < accur_deflation * accur_deflation) {
          singular = true;
          break;
        }
      std::vector<double> Binv((size_t)p * p, 0.0);
      for (int c = 0; c < p; ++c)
        for (int r = 0; r < p; ++r) {
          double sb = 0.0, si = 0.0;
          for (int m = 0; m < p; ++m) {
            const double vv = Gs[r + (size_t)p * m] * Gs[c + (size_t)p * m];
            sb += vv * std::sqrt(std::max(D[m], 0.0));
            if (!singular) si += vv / std::sqrt(D[m]);
          }
          Bj[r + (size_t)p * c] = sb;
          Binv[r + (size_t)p * c] = si;
        }
      B.push_back(Bj);
      if (singular) break;
      block_combine(Z, Qn, p, Binv, n);
    }
    if ((j + 1) >= 2 * p && (j + 1) % p == 0) {
      const int sz = (j + 1) * p;
      BlockTridiag T{A, B, p};
      double e0 = 0.0;
      bool ok = false;
      if (sz > 384 && ev_old2 < 1e11) {
        // shifted inverse iteration on the band from the previous eigenvector; the shift
        // lies below the previous value by more than its last decrease
        const double drop = std::fabs(ev_old2 - ev_old);
        const double sigma = ev_old - std::max(4.0 * drop, 1e-9 * std::max(1.0, std::fabs(ev_old)));
        ok = band_lowest(T, sz, 2 * p - 1, sigma, 3, evec, e0);
      }
      if (!ok) {
        std::vector<double> H((size_t)sz * sz), e(sz);
        for (int c = 0; c < sz; ++c)
          for (int r = 0; r < sz; ++r) H[(size_t)c * sz + r] = T(r, c);
        sym_eigen(sz, H.data(), e.data(), true);
        e0 = e[0];
        evec.assign(H.begin(), H.begin() + sz);
      }
      const double delta = std::fabs(e0 - ev_old);
      ev_old2 = ev_old;
      ev_old = e0;
      if (delta < accur_band_lanczos) break;
    }
    if (singular) break;
    if (qr) {
      std::swap(Qp, Qc);  // Q_prev <- Q_cur, Q_cur <- Z, Z <- old Q_prev
      std::swap(Qc, Z);
      Qn = Z;
    } else {
      std::vector<double*> t = Qp;
      Qp = Qc;
      Qc = Qn;
      Qn = t;
    }
  }
  QED_HIP(hipStreamSynchronize(stream()));
  return (int)A.size();
}

namespace {

// y += diag(e) x
__global__ void k_diag_axpy(uint64_t n, const double* __restrict__ e, const double* __restrict__ x,
                            double* __restrict__ y) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += e[i] * x[i];
}

// y += T x for the direct sum of block-tridiagonal matrices: part k has floors [f0_k, f1_k)
// of p x p blocks A_j (diagonal) and B_j (sub-diagonal B_{j-1}, super-diagonal B_j^T),
// column-major; floors of one part are consecutive in A, B and in x (one thread per row)
__global__ void k_blocktri_axpy(int p, int nf, const int* __restrict__ part_end, const double* __restrict__ A,
                                const double* __restrict__ B, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf * p) return;
  const int j = i / p, r = i % p;
  int f0 = 0, f1 = part_end[0];
  if (j >= f1) {
    f0 = f1;
    f1 = part_end[1];
  }
  const size_t pp = (size_t)p * p;
  double acc = 0.0;
  for (int c = 0; c < p; ++c) acc += A[j * pp + r + (size_t)p * c] * x[(size_t)j * p + c];
  if (j > f0)
    for (int c = 0; c < p; ++c) acc += B[(j - 1) * pp + r + (size_t)p * c] * x[(size_t)(j - 1) * p + c];
  if (j + 1 < f1)
    for (int c = 0; c < p; ++c) acc += B[j * pp + c + (size_t)p * r] * x[(size_t)(j + 1) * p + c];
  y[i] += acc;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// one workgroup per frequency: backward recursion of the matrix continued fraction with
// pa x pa blocks, then G = W^T F_0 W (W: pa x q column-major), added (or its transpose)
// to out[z] (q x q complex column-major, interleaved)
__global__ __launch_bounds__(RB) void k_mcf_eval(int pa, int q, int M, const double* __restrict__ A,
                                                 const double* __restrict__ B, const double* __restrict__ W,
                                                 const double* __restrict__ z, int transpose,
                                                 double* __restrict__ out) {
  extern __shared__ double smem[];
  double2* F = reinterpret_cast<double2*>(smem);          // pa x pa
  double2* T = F + pa * pa;                                // pa x max(pa, q)
  double2* Aug = T + pa * kMaxBlock;                       // pa x 2 pa, column-major
  double* Bm = reinterpret_cast<double*>(Aug + 2 * pa * pa);  // pa x pa
  __shared__ int piv;
  __shared__ double2 pval;
  const int tid = threadIdx.x;
  const double2 zz = make_double2(z[2 * blockIdx.x], z[2 * blockIdx.x + 1]);
  const int pp = pa * pa;
  for (int t = tid; t < pp; t += RB) F[t] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = M - 1; j >= 0; --j) {
    const double* Aj = A + (size_t)j * pp;
    const bool tail = j < M - 1;
    if (tail) {
      for (int t = tid; t < pp; t += RB) Bm[t] = B[(size_t)j * pp + t];
      __syncthreads();
      for (int t = tid; t < pp; t += RB) {  // T = F B
        const int r = t % pa, c = t / pa;
        double2 s = make_double2(0.0, 0.0);
        for (int k = 0; k < pa; ++k) {
          const double b = Bm[k + pa * c];
          s.x += F[r + pa * k].x * b;
          s.y += F[r + pa * k].y * b;
        }
        T[t] = s;
      }
      __syncthreads();
    }
    for (int t = tid; t < 2 * pp; t += RB) {  // [z - A_j - B^T T | 1]
      const int r = t % pa, c = t / pa;
      double2 v;
      if (c < pa) {
        v = make_double2((r == c ? zz.x : 0.0) - Aj[r + pa * c], r == c ? zz.y : 0.0);
        if (tail)
          for (int k = 0; k < pa; ++k) {
            const double b = Bm[k + pa * r];
            v.x -= b * T[k + pa * c].x;
            v.y -= b * T[k + pa * c].y;
          }
      } else {
        v = make_double2(r == c - pa ? 1.0 : 0.0, 0.0);
      }
      Aug[t] = v;
    }
    __syncthreads();
    // Gauss-Jordan with partial pivoting (largest modulus in the column)
    for (int c = 0; c < pa; ++c) {
      if (tid == 0) {
        int best = c;
        double bm = -1.0;
        for (int r = c; r < pa; ++r) {
          const double2 v = Aug[r + pa * c];
          const double m = v.x * v.x + v.y * v.y;
          if (m > bm) {
            bm = m;
            best = r;
          }
        }
        piv = best;
      }
      __syncthreads();
      const int pr = piv;
      if (pr != c)
        for (int t = tid; t < 2 * pa; t += RB) {
          const double2 a = Aug[c + pa * t];
          Aug[c + pa * t] = Aug[pr + pa * t];
          Aug[pr + pa * t] = a;
        }
      __syncthreads();
      if (tid == 0) {
        const double2 d = Aug[c + pa * c];
        const double den = d.x * d.x + d.y * d.y;
        pval = make_double2(d.x / den, -d.y / den);
      }
      __syncthreads();
      const double2 inv = pval;
      for (int t = tid; t < 2 * pa; t += RB) Aug[c + pa * t] = cmul(Aug[c + pa * t], inv);
      __syncthreads();
      for (int r = tid; r < pa; r += RB) T[r] = Aug[r + pa * c];  // elimination factors
      __syncthreads();
      for (int t = tid; t < 2 * pp; t += RB) {
        const int r = t % pa, cc = t / pa;
        if (r != c) {
          const double2 f = T[r];
          const double2 a = cmul(f, Aug[c + pa * cc]);
          Aug[t].x -= a.x;
          Aug[t].y -= a.y;
        }
      }
      __syncthreads();
    }
    for (int t = tid; t < pp; t += RB) F[t] = Aug[pp + t];
    __syncthreads();
  }
  // T = F W (pa x q), then G = W^T T
  for (int t = tid; t < pa * q; t += RB) {
    const int r = t % pa, c = t / pa;
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < pa; ++k) {
      const double w = W[k + pa * c];
      s.x += F[r + pa * k].x * w;
      s.y += F[r + pa * k].y * w;
    }
    T[t] = s;
  }
  __syncthreads();
  double* o = out + (size_t)2 * blockIdx.x * q * q;
  for (int t = tid; t < q * q; t += RB) {
    const int r = t % q, c = t / q;
    double2 s = make_double2(0.0, 0.0);
    for (int k = 0; k < pa; ++k) {
      const double w = W[k + pa * r];
      s.x += w * T[k + pa * c].x;
      s.y += w * T[k + pa * c].y;
    }
    const int d = transpose ? (c + q * r) : t;
    o[2 * d] += s.x;
    o[2 * d + 1] += s.y;
  }
}

}  // namespace

// G[z] += W^T F_0(z) W (or its transpose) for nz frequencies; A, B: M floors pa x pa
// (host), W: pa x q (host), G: host, q x q complex column-major per frequency
// Block Lanczos on diag(e) (dimension n) from p start vectors given on the host (p x n,
// vector-major): the matrix continued fraction of a Lehmann form (Q_matrix_to_mcf,
// src_ed/matrix_continued_fraction.hpp:275-314)
int block_lanczos_diag(const std::vector<double>& e, const std::vector<double>& phi, int p, int M0, int max_iter_BL,
                       double accur_deflation, double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                       std::vector<std::vector<double>>& B) {
  const uint64_t n = e.size();
  DevBuf de(n * 8), dq((size_t)p * n * 8);
  QED_HIP(hipMemcpyAsync(de.p, e.data(), n * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dq.p, phi.data(), (size_t)p * n * 8, hipMemcpyHostToDevice, stream()));
  const double* ed = de.as<double>();
  return block_lanczos_op(
      n,
      [&](const double* x, double* y) {
        k_diag_axpy<<<(unsigned)((n + 255) / 256), 256, 0, stream()>>>(n, ed, x, y);
        QED_HIP(hipGetLastError());
      },
      dq.as<double>(), p, M0, max_iter_BL, accur_deflation, accur_band_lanczos, qr, A, B);
}

// Block Lanczos on the direct sum of two block-tridiagonal matrices (floors Me and Mh of
// p x p blocks, column-major, consecutive) from p start vectors given on the host (p x N,
// N = (Me + Mh) p): the combined matrix continued fraction (combine_via_lanczos,
// src_ed/matrix_continued_fraction.hpp:442-492)
int block_lanczos_blocktri(int p, int Me, int Mh, const std::vector<double>& Ablk, const std::vector<double>& Bblk,
                           const std::vector<double>& phi, int M0, int max_iter_BL, double accur_deflation,
                           double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                           std::vector<std::vector<double>>& B) {
  const int nf = Me + Mh;
  const uint64_t n = (uint64_t)nf * p;
  const size_t pp = (size_t)p * p;
  if (Ablk.size() != nf * pp || Bblk.size() != nf * pp || phi.size() != (size_t)p * n)
    fail("combined block Lanczos: block arrays do not match the floors");
  DevBuf dA(nf * pp * 8), dB(nf * pp * 8), dq((size_t)p * n * 8), dend(2 * sizeof(int));
  const int ends[2] = {Me, nf};
  QED_HIP(hipMemcpyAsync(dA.p, Ablk.data(), nf * pp * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dB.p, Bblk.data(), nf * pp * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dq.p, phi.data(), (size_t)p * n * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dend.p, ends, sizeof ends, hipMemcpyHostToDevice, stream()));
  const double* a = dA.as<double>();
  const double* b = dB.as<double>();
  const int* pe = dend.as<int>();
  return block_lanczos_op(
      n,
      [&](const double* x, double* y) {
        k_blocktri_axpy<<<(unsigned)((n + 255) / 256), 256, 0, stream()>>>(p, nf, pe, a, b, x, y);
        QED_HIP(hipGetLastError());
      },
      dq.as<double>(), p, M0, max_iter_BL, accur_deflation, accur_band_lanczos, qr, A, B);
}

void mcf_eval(int pa, int q, int M, const double* A, const double* B, const double* W, int nz, const double* z,
              bool transpose, double* G) {
  if (M == 0 || nz == 0 || pa == 0) return;
  if (pa > kMaxBlock || q > kMaxBlock) fail("matrix continued fraction: block size above " + std::to_string(kMaxBlock));
  const size_t pp = (size_t)pa * pa;
  DevBuf dA(M * pp * 8), dB(M * pp * 8), dW((size_t)pa * q * 8), dz((size_t)nz * 16), dG((size_t)nz * q * q * 16);
  QED_HIP(hipMemcpyAsync(dA.p, A, M * pp * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dB.p, B, M * pp * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dW.p, W, (size_t)pa * q * 8, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dz.p, z, (size_t)nz * 16, hipMemcpyHostToDevice, stream()));
  QED_HIP(hipMemcpyAsync(dG.p, G, (size_t)nz * q * q * 16, hipMemcpyHostToDevice, stream()));
  const size_t lds = (pp + (size_t)pa * kMaxBlock + 2 * pp) * sizeof(double2) + pp * sizeof(double);
  k_mcf_eval<<<nz, RB, lds, stream()>>>(pa, q, M, dA.as<double>(), dB.as<double>(), dW.as<double>(),
                                        dz.as<double>(), transpose ? 1 : 0, dG.as<double>());
  QED_HIP(hipGetLastError());
  QED_HIP(hipMemcpyAsync(G, dG.p, (size_t)nz * q * q * 16, hipMemcpyDeviceToHost, stream()));
  QED_HIP(hipStreamSynchronize(stream()));
}

}  // namespace qed

using namespace qed;

extern "C" {

int qed_block_lanczos(qed_sector* s, const double* q0, int p, int M0, int max_iter_BL, double accur_deflation,
                      double accur_band_lanczos, int qr, double* A, double* B, int cap, int* M) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<std::vector<double>> a, b;
    const int m = block_lanczos(*s, q0, p, M0, max_iter_BL, accur_deflation, accur_band_lanczos, qr != 0, a, b);
    if (m > cap) fail("block Lanczos: output capacity too small");
    *M = m;
    const size_t pp = (size_t)p * p;
    for (int j = 0; j < m; ++j) {
      std::copy(a[j].begin(), a[j].end(), A + j * pp);
      std::copy(b[j].begin(), b[j].end(), B + j * pp);
    }
  });
}

int qed_mcf_eval(int p, int q, int M, const double* A, const double* B, const double* W, int nz, const double* z,
                 int transpose, double* G) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { mcf_eval(p, q, M, A, B, W, nz, z, transpose != 0, G); });
}

}  // extern "C"
