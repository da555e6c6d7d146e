// Green-function representations and the lattice k-sums on the device.
//
//  * Lehmann form: G_ab(z) = sum_m Q_am Q_bm / (z - e_m), batched over frequencies
//    (the reference's VDVH contraction, src_util/VDVH_kernel.cpp:28-48).
//  * Continued fractions evaluated bottom-up (src_ed/continued_fraction.cpp:52-57).
//  * CPT / VCA / CDMFT k-sums: for every (w, k) the n x n complex matrix
//    M = Gc(w)^-1 - V(k) is inverted (G_cpt = M^-1, src_qcm/CPT.cpp:76-86) or factorized
//    (log|det(V Gc - 1)| = log|det Gc| + log|det M|, src_qcm/lattice_model_instance.cpp:624-643).
//    n lanes hold one matrix, lane j owns column j in registers; the pivot column is
//    broadcast with shuffles. Several matrices share a wave (n <= 16 -> >= 4 per wave).
#include <cmath>
#include <complex>
#include <vector>

#include "../../include/qed.h"
#include "qed_common.hpp"
#include "qed_linalg.hpp"

namespace qed {

namespace {

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cinv(cplx a) {
  double d = a.re * a.re + a.im * a.im;
  return {a.re / d, -a.im / d};
}
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ cplx shfl(cplx v, int src) {
  return {__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
}

// ---- Lehmann G(z) --------------------------------------------------------------------
// one workgroup per frequency; thread t accumulates element (b, a) = (t % L, t / L)
__global__ __launch_bounds__(256) void k_qmatrix_gf(int L, int M, const double* __restrict__ e,
                                                    const double* __restrict__ Q,
                                                    const double* __restrict__ z,
                                                    double* __restrict__ G) {
  __shared__ double ur[256], ui[256];
  __shared__ double qs[256 * 16];
  const int iz = blockIdx.x;
  const double zr = z[2 * iz], zi = z[2 * iz + 1];
  const int t = threadIdx.x;
  const int b = t % L, a = t / L;
  const bool act = t < L * L;
  double gr = 0.0, gi = 0.0;
  for (int m0 = 0; m0 < M; m0 += 256) {
    const int mc = min(256, M - m0);
    __syncthreads();
    if (t < mc) {
      double dr = zr - e[m0 + t], di = zi;
      double d = dr * dr + di * di;
      ur[t] = dr / d;
      ui[t] = -di / d;
    }
    for (int i = t; i < mc * L && L <= 16; i += 256) qs[i] = Q[(size_t)m0 * L + i];
    __syncthreads();
    if (act) {
      for (int m = 0; m < mc; ++m) {
        double w = qs[m * L + a] * qs[m * L + b];
        gr += w * ur[m];
        gi += w * ui[m];
      }
    }
  }
  if (act) {
    G[((size_t)iz * L * L + (size_t)b + (size_t)a * L) * 2] += gr;
    G[((size_t)iz * L * L + (size_t)b + (size_t)a * L) * 2 + 1] += gi;
  }
}

__global__ void k_cf_eval(int n, const int* __restrict__ off, const int* __restrict__ depth,
                          const double* __restrict__ a, const double* __restrict__ b, int nz,
                          const double* __restrict__ z, double* __restrict__ out) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n * nz) return;
  int f = id % n, iz = id / n;
  cplx zz{z[2 * iz], z[2 * iz + 1]};
  cplx G{0.0, 0.0};
  const double* aa = a + off[f];
  const double* bb = b + off[f];
  for (int i = depth[f] - 1; i >= 0; --i) {
    cplx den = csub(cplx{zz.re - aa[i], zz.im}, G);
    cplx inv = cinv(den);
    G = {bb[i] * inv.re, bb[i] * inv.im};
  }
  out[2 * (size_t)id] = G.re;
  out[2 * (size_t)id + 1] = G.im;
}

// ---- batched small complex matrices ---------------------------------------------------

struct KsumArgs {
  int n, nw, nterm, nk, mode;
  const double* ginv;   // nw x n x n complex (column-major per matrix): Gc(w)^-1
  const double* V0;     // n x n complex
  const double* Vt;     // nterm x n x n complex
  const double* nv;     // nterm x 3
  const double* k;      // nk x 3
  double* out;
  double* part;         // partial sums (mode 2: per group n x n; sef: per group scalar)
  int kchunk;           // k points per group
};

// Builds column j of M = Ginv(w) - V(k) into reg[0..N)
template <int N>
__device__ __forceinline__ void build_column(const KsumArgs& A, int w, int kk, int j, cplx* col) {
  const double* gi = A.ginv + ((size_t)w * N * N + (size_t)j * N) * 2;
  const double* v0 = A.V0 + (size_t)j * N * 2;
#pragma unroll
  for (int i = 0; i < N; ++i) col[i] = {gi[2 * i] - v0[2 * i], gi[2 * i + 1] - v0[2 * i + 1]};
  const double kx = A.k[3 * kk], ky = A.k[3 * kk + 1], kz = A.k[3 * kk + 2];
  for (int t = 0; t < A.nterm; ++t) {
    const double ph = 2.0 * M_PI * (kx * A.nv[3 * t] + ky * A.nv[3 * t + 1] + kz * A.nv[3 * t + 2]);
    double s, c;
    sincos(ph, &s, &c);
    const double* vt = A.Vt + ((size_t)t * N * N + (size_t)j * N) * 2;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cplx v{vt[2 * i], vt[2 * i + 1]};
      cplx e = cmul(v, cplx{c, s});
      col[i] = csub(col[i], e);
    }
  }
}

// In-place Gauss-Jordan inversion of the matrix whose column j lives in col (lane base+j of
// a group of N lanes). Returns log|det| in every lane.
template <int N>
__device__ __forceinline__ double gj_invert(cplx* col, int base, int j, bool want_inverse) {
  int perm[N];
  double logdet = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // pivot search in column k (held by lane base+k)
    int p = k;
    double best = -1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double m = (i >= k) ? cabs2(col[i]) : -1.0;
      if (m > best) {
        best = m;
        p = i;
      }
    }
    p = __shfl(p, base + k, 64);
    perm[k] = p;
    // swap rows k and p in every column
    cplx rk = col[k], rp = col[k];
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i == p) rp = col[i];
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i == p) col[i] = rk;
    col[k] = rp;
    // broadcast column k after the swap
    cplx ck[N];
#pragma unroll
    for (int i = 0; i < N; ++i) ck[i] = shfl(col[i], base + k);
    const cplx piv = ck[k];
    logdet += 0.5 * log(cabs2(piv));
    if (!want_inverse) {
      // plain elimination below the pivot (determinant only)
      const cplx dinv = cinv(piv);
      const cplx f = cmul(col[k], dinv);
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (i > k) col[i] = csub(col[i], cmul(ck[i], f));
      continue;
    }
    const cplx dinv = cinv(piv);
    // row k scaled; column k replaced by the eliminator
    const cplx rowk = (j == k) ? cplx{1.0, 0.0} : col[k];
    const cplx rs = cmul(rowk, dinv);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == k) {
        col[i] = rs;
      } else {
        cplx base_v = (j == k) ? cplx{0.0, 0.0} : col[i];
        col[i] = csub(base_v, cmul(ck[i], rs));
      }
    }
  }
  if (want_inverse) {
    // undo the row interchanges as column interchanges, last first
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p == k) continue;
      // lanes k and p exchange their columns
      const int src = (j == k) ? p : ((j == p) ? k : j);
#pragma unroll
      for (int i = 0; i < N; ++i) col[i] = shfl(col[i], base + src);
    }
  }
  return logdet;
}

// At least two wavefronts per SIMD: the fully unrolled elimination of a 12- or 16-wide
// matrix otherwise takes over 256 registers and runs one wavefront per SIMD (measured on
// the 3x4 cluster: 30.7 ms at one, 22.4 ms at two with a few spilled registers, 25.3 ms at
// three with more spills; 16 wide on 128^2 k x 512 w: 86.3 ms at one, 59.8 ms at two)
template <int N>
__global__ __launch_bounds__(256, 2) void k_ksum(KsumArgs A) {
  constexpr int GPW = 64 / N;  // groups per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / N, j = lane % N;
  const int base = g * N;
  const bool lane_ok = g < GPW;
  const int gid = (blockIdx.x * 4 + wave) * GPW + g;  // global group id
  const int ngroups_k = (A.nk + A.kchunk - 1) / A.kchunk;
  const int w = gid / ngroups_k, kc = gid % ngroups_k;
  const bool act = lane_ok && w < A.nw;
  cplx acc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] = {0.0, 0.0};
  double sef = 0.0;
  if (act) {
    const int k0 = kc * A.kchunk, k1 = min(A.nk, k0 + A.kchunk);
    for (int kk = k0; kk < k1; ++kk) {
      cplx col[N];
      build_column<N>(A, w, kk, j, col);
      const double ld = gj_invert<N>(col, base, j, A.mode != 3);
      if (A.mode == 0) {
        double* o = A.out + (((size_t)w * A.nk + kk) * N * N + (size_t)j * N) * 2;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          o[2 * i] = col[i].re;
          o[2 * i + 1] = col[i].im;
        }
      } else if (A.mode == 1) {
        // trace: lane j contributes element (j, j)
        cplx d{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < N; ++i)
          if (i == j) d = col[i];
        double tr = d.re, ti = d.im;
        if constexpr ((N & (N - 1)) == 0) {
          for (int o = 1; o < N; o <<= 1) {
            tr += __shfl_xor(tr, o, 64);
            ti += __shfl_xor(ti, o, 64);
          }
        } else {
          // groups of N lanes do not align with xor butterflies: gather the diagonal
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            sr += __shfl(tr, base + i, 64);
            si += __shfl(ti, base + i, 64);
          }
          tr = sr;
          ti = si;
        }
        if (j == 0) {
          A.out[2 * ((size_t)w * A.nk + kk)] = tr;
          A.out[2 * ((size_t)w * A.nk + kk) + 1] = ti;
        }
      } else if (A.mode == 2) {
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] = cadd(acc[i], col[i]);
      } else {
        sef += ld;
      }
    }
    if (A.mode == 2) {
      double* o = A.part + (((size_t)w * ngroups_k + kc) * N * N + (size_t)j * N) * 2;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        o[2 * i] = acc[i].re;
        o[2 * i + 1] = acc[i].im;
      }
    } else if (A.mode == 3 && j == 0) {
      A.part[(size_t)w * ngroups_k + kc] = sef;
    }
  }
}

// sum of partials over k-chunks
__global__ void k_ksum_reduce(int nw, int nparts, int len, const double* __restrict__ part,
                              double scale, double* __restrict__ out) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nw * len) return;
  int w = id / len, q = id % len;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[((size_t)w * nparts + p) * len + q];
  out[id] = s * scale;
}

template <class T>
DevBuf up(const T* p, size_t n) {
  DevBuf b(std::max<size_t>(n, 1) * sizeof(T));
  if (n) QED_HIP(hipMemcpyAsync(b.p, p, n * sizeof(T), hipMemcpyHostToDevice, stream()));
  return b;
}

template <int N>
void launch_ksum(KsumArgs A, int ngroups) {
  constexpr int GPW = 64 / N;
  const int waves = (ngroups + GPW - 1) / GPW;
  const int blocks = (waves + 3) / 4;
  k_ksum<N><<<blocks, 256, 0, stream()>>>(A);
}

void ksum_dispatch(KsumArgs A, int ngroups) {
  switch (A.n) {
    case 1: launch_ksum<1>(A, ngroups); break;
    case 2: launch_ksum<2>(A, ngroups); break;
    case 4: launch_ksum<4>(A, ngroups); break;
    case 5: case 6: case 7: case 8: launch_ksum<8>(A, ngroups); break;
    case 12: launch_ksum<12>(A, ngroups); break;  // 5 groups of 12 lanes per wave
    case 16: launch_ksum<16>(A, ngroups); break;
    default: fail("lattice k-sum: cluster Green function dimension above 16 is not supported");
  }
  QED_HIP(hipGetLastError());
}

// Pads n x n complex matrices to N x N (identity padding for inverse-like uses, zero for V)
std::vector<double> pad_mats(const double* src, int count, int n, int N, bool identity) {
  std::vector<double> out((size_t)count * N * N * 2, 0.0);
  for (int c = 0; c < count; ++c) {
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        double re = 0.0, im = 0.0;
        if (i < n && j < n) {
          re = src[(((size_t)c * n * n) + i + (size_t)j * n) * 2];
          im = src[(((size_t)c * n * n) + i + (size_t)j * n) * 2 + 1];
        } else if (identity && i == j) {
          re = 1.0;
        }
        out[(((size_t)c * N * N) + i + (size_t)j * N) * 2] = re;
        out[(((size_t)c * N * N) + i + (size_t)j * N) * 2 + 1] = im;
      }
  }
  return out;
}

int padded_dim(int n) {
  if (n <= 2) return n;
  if (n <= 4) return 4;
  if (n <= 8) return 8;
  if (n <= 12) return 12;
  if (n <= 16) return 16;
  fail("lattice k-sum: cluster Green function dimension above 16 is not supported");
}

// Gc(w)^-1 for each frequency (host, n <= 16), padded to N
std::vector<double> inverse_gc(const double* Gc, int nw, int n, int N, std::vector<double>* logdet) {
  std::vector<double> out((size_t)nw * N * N * 2, 0.0);
  if (logdet) logdet->assign(nw, 0.0);
  std::vector<std::complex<double>> a((size_t)n * n);
  for (int w = 0; w < nw; ++w) {
    for (size_t q = 0; q < (size_t)n * n; ++q)
      a[q] = {Gc[((size_t)w * n * n + q) * 2], Gc[((size_t)w * n * n + q) * 2 + 1]};
    if (logdet) {
      std::vector<std::complex<double>> b(a);
      (*logdet)[w] = clogabsdet(n, b.data());
    }
    if (!cinverse(n, a.data())) fail("cluster Green function is singular at a requested frequency");
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        std::complex<double> v = (i < n && j < n) ? a[i + (size_t)j * n]
                                                  : std::complex<double>(i == j ? 1.0 : 0.0, 0.0);
        out[(((size_t)w * N * N) + i + (size_t)j * N) * 2] = v.real();
        out[(((size_t)w * N * N) + i + (size_t)j * N) * 2 + 1] = v.imag();
      }
  }
  return out;
}

// device time of the last CPT k-sum (HIP events on the library stream)
double& last_ksum_ms() {
  static double ms = 0.0;
  return ms;
}

struct KsumTimer {
  hipEvent_t a = nullptr, b = nullptr;
  KsumTimer() {
    QED_HIP(hipEventCreate(&a));
    QED_HIP(hipEventCreate(&b));
    QED_HIP(hipEventRecord(a, stream()));
  }
  void stop() {
    QED_HIP(hipEventRecord(b, stream()));
    QED_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    QED_HIP(hipEventElapsedTime(&ms, a, b));
    last_ksum_ms() = ms;
  }
  ~KsumTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

}  // namespace qed

using namespace qed;

extern "C" {

int qed_qmatrix_gf(int L, int M, const double* e, const double* Q, int nz, const double* z, double* G) {
  return guard([&] {
    if (L <= 0 || nz <= 0) return;
    if (L > 16) fail("qmatrix_gf: at most 16 orbitals per block");
    DevBuf de = up(e, (size_t)M), dq = up(Q, (size_t)L * M), dz = up(z, (size_t)2 * nz);
    DevBuf dg = up(G, (size_t)2 * nz * L * L);
    if (M > 0) k_qmatrix_gf<<<nz, 256, 0, stream()>>>(L, M, de.as<double>(), dq.as<double>(), dz.as<double>(), dg.as<double>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(G, dg.p, (size_t)2 * nz * L * L * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_cf_eval(int n, const int* off, const int* depth, const double* a, const double* b, int nz,
                const double* z, double* out) {
  return guard([&] {
    if (n <= 0 || nz <= 0) return;
    size_t tot = 0;
    for (int f = 0; f < n; ++f) tot = std::max<size_t>(tot, (size_t)off[f] + depth[f]);
    DevBuf doff = up(off, n), ddep = up(depth, n), da = up(a, tot), db = up(b, tot), dz = up(z, 2 * (size_t)nz);
    DevBuf dout((size_t)2 * n * nz * 8);
    int tot_threads = n * nz;
    k_cf_eval<<<(tot_threads + 255) / 256, 256, 0, stream()>>>(n, doff.as<int>(), ddep.as<int>(), da.as<double>(),
                                                               db.as<double>(), nz, dz.as<double>(), dout.as<double>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(out, dout.p, (size_t)2 * n * nz * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_ksum_last_ms(double* ms) {
  *ms = last_ksum_ms();
  return 0;
}

int qed_cpt_batch(int n, int nw, const double* Gc, const double* V0, int nterm, const double* Vt,
                  const double* nv, int nk, const double* k, int mode, double* out) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail("cpt_batch: mode must be 0, 1 or 2");
    if (n <= 0 || nw <= 0 || nk <= 0) return;
    const int N = padded_dim(n);
    std::vector<double> ginv = inverse_gc(Gc, nw, n, N, nullptr);
    std::vector<double> v0 = pad_mats(V0, 1, n, N, false);
    std::vector<double> vt = pad_mats(Vt, nterm, n, N, false);
    DevBuf dgi = upload(ginv), dv0 = upload(v0), dvt = upload(vt.empty() ? std::vector<double>{0} : vt);
    DevBuf dnv = up(nv, (size_t)3 * nterm), dk = up(k, (size_t)3 * nk);
    KsumArgs A{};
    A.n = N;
    A.nw = nw;
    A.nterm = nterm;
    A.nk = nk;
    A.mode = mode;
    A.ginv = dgi.as<double>();
    A.V0 = dv0.as<double>();
    A.Vt = dvt.as<double>();
    A.nv = dnv.as<double>();
    A.k = dk.as<double>();
    A.kchunk = (mode == 2) ? std::max(1, nk / 64) : 1;
    const int ngk = (nk + A.kchunk - 1) / A.kchunk;
    const int ngroups = nw * ngk;
    DevBuf dout, dpart;
    size_t out_len;
    if (mode == 0) out_len = (size_t)nw * nk * N * N * 2;
    else if (mode == 1) out_len = (size_t)nw * nk * 2;
    else out_len = (size_t)nw * N * N * 2;
    dout.alloc(out_len * 8);
    if (mode == 2) dpart.alloc((size_t)ngroups * N * N * 2 * 8);
    A.out = dout.as<double>();
    A.part = dpart.as<double>();
    KsumTimer timer;
    ksum_dispatch(A, ngroups);
    if (mode == 2) {
      int len = N * N * 2;
      k_ksum_reduce<<<(nw * len + 255) / 256, 256, 0, stream()>>>(nw, ngk, len, dpart.as<double>(), 1.0 / nk,
                                                                 dout.as<double>());
      QED_HIP(hipGetLastError());
    }
    timer.stop();
    std::vector<double> h(out_len);
    QED_HIP(hipMemcpyAsync(h.data(), dout.p, out_len * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    // unpad
    if (mode == 1) {
      std::copy(h.begin(), h.end(), out);
      // padded identity blocks add (N - n) to the trace of M^-1
      for (size_t q = 0; q < (size_t)nw * nk; ++q) out[2 * q] -= (N - n);
    } else {
      size_t cnt = (mode == 0) ? (size_t)nw * nk : (size_t)nw;
      for (size_t c = 0; c < cnt; ++c)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            out[((c * n * n) + i + (size_t)j * n) * 2] = h[((c * N * N) + i + (size_t)j * N) * 2];
            out[((c * n * n) + i + (size_t)j * n) * 2 + 1] = h[((c * N * N) + i + (size_t)j * N) * 2 + 1];
          }
    }
  });
}

int qed_sef_batch(int n, int nw, const double* Gc, const double* V0, int nterm, const double* Vt,
                  const double* nv, int nk, const double* k, double* out) {
  return guard([&] {
    if (n <= 0 || nw <= 0 || nk <= 0) return;
    const int N = padded_dim(n);
    std::vector<double> logdetG;
    std::vector<double> ginv = inverse_gc(Gc, nw, n, N, &logdetG);
    std::vector<double> v0 = pad_mats(V0, 1, n, N, false);
    std::vector<double> vt = pad_mats(Vt, nterm, n, N, false);
    DevBuf dgi = upload(ginv), dv0 = upload(v0), dvt = upload(vt.empty() ? std::vector<double>{0} : vt);
    DevBuf dnv = up(nv, (size_t)3 * nterm), dk = up(k, (size_t)3 * nk);
    KsumArgs A{};
    A.n = N;
    A.nw = nw;
    A.nterm = nterm;
    A.nk = nk;
    A.mode = 3;
    A.ginv = dgi.as<double>();
    A.V0 = dv0.as<double>();
    A.Vt = dvt.as<double>();
    A.nv = dnv.as<double>();
    A.k = dk.as<double>();
    A.kchunk = std::max(1, nk / 256);
    const int ngk = (nk + A.kchunk - 1) / A.kchunk;
    DevBuf dpart((size_t)nw * ngk * 8), dout((size_t)nw * 8);
    A.part = dpart.as<double>();
    A.out = dout.as<double>();
    ksum_dispatch(A, nw * ngk);
    k_ksum_reduce<<<(nw + 255) / 256, 256, 0, stream()>>>(nw, ngk, 1, dpart.as<double>(), 1.0 / nk,
                                                         dout.as<double>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(out, dout.p, nw * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    // log|det(V G - 1)| = log|det G| + log|det(G^-1 - V)|  (padding blocks contribute log 1 = 0)
    for (int w = 0; w < nw; ++w) out[w] += logdetG[w];
  });
}

}  // extern "C"
