// CDMFT bath fit on the device.
//
//  * Distance function (src_qcm/CPT.cpp:519-555) for many bath parameter sets and hosts:
//      d = scale * sum_w weight[w] || Gamma(i w) + G_host(i w) ||_F^2,
//      Gamma_ac(z) = sum_m tcb[a,m] tcb[c,m] / (z - eb[m])   (diagonalized bath).
//  * Batched adaptive Nelder-Mead: one wavefront per CDMFT instance runs the whole
//    minimization (the reference calls scipy.optimize.minimize with method='Nelder-Mead',
//    adaptive=True, an initial simplex x0 + step e_i, xatol, fatol, maxfev;
//    pyqcm/cdmft.py:1614-1633). The host function of the instance, its simplex and the
//    trial points live in LDS; each distance evaluation is spread over the frequencies.
//
// This file is compiled with -ffp-contract=off: the simplex arithmetic must round like
// numpy's (no fused multiply-adds), and the distance kernel and the fit kernel share one
// wavefront distance function, so a distance evaluated by either is the same double.
// The fit therefore follows, step for step, the host-driven Nelder-Mead that evaluates
// its trial points with qed_cdmft_distance_batch; the simplex is ordered by a stable sort
// (exact ties keep their previous order).
#include <cmath>
#include <vector>

#include "../../include/qed.h"
#include "qed_common.hpp"

namespace qed {
namespace {

// Distance of one parameter set, evaluated by the whole workgroup (one wavefront for the
// distance kernel, four for the fit). G: host, nw matrices n x n complex column-major
// (interleaved re, im); eb: nb bath energies; tcb: n x nb couplings column-major.
// Scratch (per wave): inv 2 nw nb, term nw n n, part nw doubles.
// Stages: 1/(i w - e_m) per (w, m); |Gamma + G|^2 per matrix entry (a + c n); the
// entries of each frequency summed in order and weighted; the frequencies summed in order.
// The host-visible distance and the fit both use this function, so their values agree
// to the last bit.
__device__ double wave_distance(int n, int nb, int nw, const double* w, const double* weight,
                                const double* G, const double* eb, const double* tcb,
                                double scale, double* inv, double* term, double* part) {
  const int lane = threadIdx.x, nth = blockDim.x;
  const int nn = n * n;
  for (int q = lane; q < nw * nb; q += nth) {
    const int iw = q / nb, m = q % nb;
    // 1/(i w - e) = (-e - i w)/(e^2 + w^2)
    const double d = eb[m] * eb[m] + w[iw] * w[iw];
    inv[2 * q] = -eb[m] / d;
    inv[2 * q + 1] = -w[iw] / d;
  }
  __syncthreads();
  for (int q = lane; q < nw * nn; q += nth) {
    const int iw = q / nn, e = q % nn;
    const int a = e % n, c = e / n;
    double gr = G[2 * (size_t)q];
    double gi = G[2 * (size_t)q + 1];
    const double* iv = inv + 2 * iw * nb;
    for (int m = 0; m < nb; ++m) {
      const double tt = tcb[a + m * n] * tcb[c + m * n];
      gr = gr + tt * iv[2 * m];
      gi = gi + tt * iv[2 * m + 1];
    }
    term[q] = gr * gr + gi * gi;
  }
  __syncthreads();
  for (int iw = lane; iw < nw; iw += nth) {
    double s = 0.0;
    for (int e = 0; e < nn; ++e) s = s + term[iw * nn + e];
    part[iw] = s * weight[iw];
  }
  __syncthreads();
  double s = 0.0;
  for (int iw = 0; iw < nw; ++iw) s = s + part[iw];
  __syncthreads();
  return s * scale;
}

inline size_t wave_scratch(int n, int nb, int nw) {
  return (size_t)2 * nw * nb + (size_t)nw * n * n + nw;
}

// one wavefront per parameter set
__global__ __launch_bounds__(64) void k_dist_wave(int n, int nb, int nw, const double* __restrict__ w,
                                                  const double* __restrict__ weight,
                                                  const double* __restrict__ Ghost,
                                                  const int* __restrict__ host_of,
                                                  const double* __restrict__ eb,
                                                  const double* __restrict__ tcb, double scale,
                                                  double* __restrict__ d) {
  extern __shared__ double lds[];
  const int b = blockIdx.x;
  double* inv = lds;
  double* term = inv + 2 * nw * nb;
  double* part = term + nw * n * n;
  const double* G = Ghost + (size_t)host_of[b] * nw * n * n * 2;
  double v = wave_distance(n, nb, nw, w, weight, G, eb + (size_t)b * nb, tcb + (size_t)b * n * nb,
                           scale, inv, term, part);
  if (threadIdx.x == 0) d[b] = v;
}

constexpr int kFitThreads = 256;  // four wavefronts per fit: the distance stages run 4x wider

struct FitArgs {
  int n, nb, nw, nvar;
  const double* w;
  const double* weight;
  const double* Ghost;
  const int* host_of;
  const double* E;  // nvar x nb
  const double* T;  // nvar x (n nb)
  const double* x0;
  double step, xatol, fatol, scale;
  int maxfev;
  double* x;
  double* f;
  int* nfev;
  int* status;
};

// LDS layout (doubles): G (nw n n 2) | sim ((N+1) N) | fsim (N+1) | xbar N | pts (N N) |
// fpts N | eb nb | tcb n nb | distance scratch (inv, term, part) | E (N nb) | T (N n nb) |
// w nw | weight nw
struct Lds {
  double *G, *sim, *fsim, *xbar, *pts, *fpts, *eb, *tcb, *inv, *term, *part, *E, *T, *w, *wt;
};

__device__ Lds carve(double* base, const FitArgs& A) {
  const int N = A.nvar;
  Lds L;
  L.G = base;
  L.sim = L.G + A.nw * A.n * A.n * 2;
  L.fsim = L.sim + (N + 1) * N;
  L.xbar = L.fsim + (N + 1);
  L.pts = L.xbar + N;
  L.fpts = L.pts + N * N;
  L.eb = L.fpts + N;
  L.tcb = L.eb + A.nb;
  L.inv = L.tcb + A.n * A.nb;
  L.term = L.inv + 2 * A.nw * A.nb;
  L.part = L.term + A.nw * A.n * A.n;
  L.E = L.part + A.nw;
  L.T = L.E + A.nvar * A.nb;
  L.w = L.T + (size_t)A.nvar * A.n * A.nb;
  L.wt = L.w + A.nw;
  return L;
}

// distance of the point x (LDS, length nvar), evaluated by the whole wave
__device__ double eval_point(const FitArgs& A, const Lds& L, const double* x) {
  const int lane = threadIdx.x;
  const int nt = A.n * A.nb;
  for (int m = lane; m < A.nb; m += (int)blockDim.x) {
    double acc = 0.0;
    for (int p = 0; p < A.nvar; ++p) acc = acc + x[p] * L.E[p * A.nb + m];
    L.eb[m] = acc;
  }
  for (int q = lane; q < nt; q += (int)blockDim.x) {
    double acc = 0.0;
    for (int p = 0; p < A.nvar; ++p) acc = acc + x[p] * L.T[(size_t)p * nt + q];
    L.tcb[q] = acc;
  }
  __syncthreads();
  return wave_distance(A.n, A.nb, A.nw, L.w, L.wt, L.G, L.eb, L.tcb, A.scale, L.inv, L.term,
                       L.part);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// maximum over the workgroup (every thread gets it); red: blockDim/64 doubles of LDS
__device__ double block_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
  return m;
}

// Simplex vertices stay in fixed LDS rows; ord[0..N] lists the rows by increasing value
// (a stable order: equal values keep their previous relative order, which is what a
// stable argsort of the simplex in its current order gives).
// Re-sorts after every vertex except the one at rank N changed value (lane 0).
__device__ void sort_all(const Lds& L, int* ord, int N) {
  if (threadIdx.x == 0) {
    for (int i = 1; i <= N; ++i) {
      int v = ord[i];
      double fv = L.fsim[v];
      int j = i - 1;
      while (j >= 0 && L.fsim[ord[j]] > fv) {
        ord[j + 1] = ord[j];
        --j;
      }
      ord[j + 1] = v;
    }
  }
  __syncthreads();
}

// The vertex at rank N got a new value: move it to its stable position (lane 0).
__device__ void reinsert_last(const Lds& L, int* ord, int N) {
  if (threadIdx.x == 0) {
    int v = ord[N];
    double fv = L.fsim[v];
    int j = N - 1;
    while (j >= 0 && L.fsim[ord[j]] > fv) {
      ord[j + 1] = ord[j];
      --j;
    }
    ord[j + 1] = v;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kFitThreads) void k_nelder_mead(FitArgs A) {
  extern __shared__ double lds[];
  __shared__ int ord[65];
  __shared__ double red[kFitThreads / 64];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = A.nvar;
  const Lds L = carve(lds, A);
  const double* Gh = A.Ghost + (size_t)A.host_of[b] * A.nw * A.n * A.n * 2;
  for (int i = lane; i < A.nw * A.n * A.n * 2; i += kFitThreads) L.G[i] = Gh[i];
  for (int i = lane; i < N * A.nb; i += kFitThreads) L.E[i] = A.E[i];
  for (int i = lane; i < N * A.n * A.nb; i += kFitThreads) L.T[i] = A.T[i];
  for (int i = lane; i < A.nw; i += kFitThreads) {
    L.w[i] = A.w[i];
    L.wt[i] = A.weight[i];
  }
  const double dim = (double)N;
  const double rho = 1.0, chi = 1.0 + 2.0 / dim, psi = 0.75 - 1.0 / (2.0 * dim),
               sigma = 1.0 - 1.0 / dim;
  const double* x0 = A.x0 + (size_t)b * N;
  for (int i = lane; i < (N + 1) * N; i += kFitThreads) {
    int r = i / N, c = i % N;
    L.sim[i] = x0[c] + ((r == c + 1) ? A.step : 0.0);
  }
  for (int r = lane; r <= N; r += kFitThreads) ord[r] = r;
  __syncthreads();
  for (int r = 0; r <= N; ++r) {
    double fv = eval_point(A, L, L.sim + r * N);
    if (lane == 0) L.fsim[r] = fv;
  }
  __syncthreads();
  int nfev = N + 1;
  sort_all(L, ord, N);
  int status = 0;
  double* xr = L.pts;      // reflection
  double* xo = L.pts + N;  // expansion / contraction
  while (true) {
    if (nfev >= A.maxfev) {
      status = 1;
      break;
    }
    // termination: max |sim[r] - sim[0]| <= xatol and max |f[0] - f[r]| <= fatol
    const double* best = L.sim + ord[0] * N;
    double dx = 0.0, df = 0.0;
    for (int i = lane; i < N * N; i += kFitThreads) {
      int r = i / N + 1, c = i % N;
      dx = fmax(dx, fabs(L.sim[ord[r] * N + c] - best[c]));
    }
    for (int r = 1 + lane; r <= N; r += kFitThreads) df = fmax(df, fabs(L.fsim[ord[0]] - L.fsim[ord[r]]));
    dx = block_max(dx, red);
    df = block_max(df, red);
    if (dx <= A.xatol && df <= A.fatol) break;
    // centroid of the N best vertices, summed in rank order
    double* worst = L.sim + ord[N] * N;
    if (lane < N) {
      double s = L.sim[ord[0] * N + lane];
      for (int r = 1; r < N; ++r) s = s + L.sim[ord[r] * N + lane];
      const double xb = s / dim;
      L.xbar[lane] = xb;
      xr[lane] = (1.0 + rho) * xb - rho * worst[lane];
    }
    __syncthreads();
    const double fxr = eval_point(A, L, xr);
    ++nfev;
    const double f0 = L.fsim[ord[0]], fsecond = L.fsim[ord[N - 1]], fworst = L.fsim[ord[N]];
    if (fxr < f0) {
      if (lane < N) xo[lane] = (1.0 + rho * chi) * L.xbar[lane] - rho * chi * worst[lane];
      __syncthreads();
      const double fxe = eval_point(A, L, xo);
      ++nfev;
      const bool take_e = fxe < fxr;
      if (lane < N) worst[lane] = take_e ? xo[lane] : xr[lane];
      if (lane == 0) L.fsim[ord[N]] = take_e ? fxe : fxr;
      __syncthreads();
      reinsert_last(L, ord, N);
    } else if (fxr < fsecond) {
      if (lane < N) worst[lane] = xr[lane];
      if (lane == 0) L.fsim[ord[N]] = fxr;
      __syncthreads();
      reinsert_last(L, ord, N);
    } else {
      bool shrink = false;
      if (fxr < fworst) {
        if (lane < N) xo[lane] = (1.0 + psi * rho) * L.xbar[lane] - psi * rho * worst[lane];
        __syncthreads();
        const double fxc = eval_point(A, L, xo);
        ++nfev;
        if (fxc <= fxr) {
          if (lane < N) worst[lane] = xo[lane];
          if (lane == 0) L.fsim[ord[N]] = fxc;
          __syncthreads();
          reinsert_last(L, ord, N);
        } else {
          shrink = true;
        }
      } else {
        if (lane < N) xo[lane] = (1.0 - psi) * L.xbar[lane] + psi * worst[lane];
        __syncthreads();
        const double fxcc = eval_point(A, L, xo);
        ++nfev;
        if (fxcc < fworst) {
          if (lane < N) worst[lane] = xo[lane];
          if (lane == 0) L.fsim[ord[N]] = fxcc;
          __syncthreads();
          reinsert_last(L, ord, N);
        } else {
          shrink = true;
        }
      }
      if (shrink) {
        // every vertex but the best moves toward it, then all are re-ranked
        const double* bst = L.sim + ord[0] * N;
        for (int i = lane; i < N * N; i += kFitThreads) {
          int r = i / N + 1, c = i % N;
          double* row = L.sim + ord[r] * N;
          row[c] = bst[c] + sigma * (row[c] - bst[c]);
        }
        __syncthreads();
        for (int r = 1; r <= N; ++r) {
          const double fv = eval_point(A, L, L.sim + ord[r] * N);
          if (lane == 0) L.fsim[ord[r]] = fv;
        }
        nfev += N;
        __syncthreads();
        sort_all(L, ord, N);
      }
    }
  }
  if (lane < N) A.x[(size_t)b * N + lane] = L.sim[ord[0] * N + lane];
  if (lane == 0) {
    double fbest = L.fsim[0];
    for (int r = 1; r <= N; ++r) fbest = fmin(fbest, L.fsim[r]);
    A.f[b] = fbest;
    A.nfev[b] = nfev;
    A.status[b] = status;
  }
}

template <class T>
DevBuf upload(const T* p, size_t n) {
  DevBuf b(std::max<size_t>(n, 1) * sizeof(T));
  if (n) QED_HIP(hipMemcpyAsync(b.p, p, n * sizeof(T), hipMemcpyHostToDevice, stream()));
  return b;
}

void distance(int n, int nb, int nw, const double* w, const double* weight, int nhost,
              const double* Ghost, int B, const int* host_of, const double* eb, const double* tcb,
              double scale, double* d) {
  if (B <= 0) return;
  if (n <= 0 || nb < 0 || nw <= 0 || nhost <= 0) fail("cdmft distance: bad sizes");
  for (int b = 0; b < B; ++b)
    if (host_of[b] < 0 || host_of[b] >= nhost) fail("cdmft distance: host index out of range");
  DevBuf dw = upload(w, nw), dwt = upload(weight, nw);
  DevBuf dg = upload(Ghost, (size_t)nhost * nw * n * n * 2);
  DevBuf dh = upload(host_of, B);
  DevBuf de = upload(eb, (size_t)B * nb), dt = upload(tcb, (size_t)B * n * nb);
  size_t lds = wave_scratch(n, nb, nw) * sizeof(double);
  if (lds > 64 * 1024) fail("cdmft distance: problem too large for one workgroup's LDS");
  DevBuf dd((size_t)B * 8);
  k_dist_wave<<<B, 64, lds, stream()>>>(n, nb, nw, dw.as<double>(), dwt.as<double>(),
                                         dg.as<double>(), dh.as<int>(), de.as<double>(),
                                         dt.as<double>(), scale, dd.as<double>());
  QED_HIP(hipGetLastError());
  QED_HIP(hipMemcpyAsync(d, dd.p, (size_t)B * 8, hipMemcpyDeviceToHost, stream()));
  QED_HIP(hipStreamSynchronize(stream()));
}

}  // namespace
}  // namespace qed

using namespace qed;

extern "C" {

int qed_cdmft_distance_batch(int n, int nb, int nw, const double* w, const double* weight,
                             const double* Ghost, int B, const double* eb, const double* tcb,
                             double scale, double* d) {
  return guard([&] {
    std::vector<int> h(std::max(B, 1), 0);
    distance(n, nb, nw, w, weight, 1, Ghost, B, h.data(), eb, tcb, scale, d);
  });
}

int qed_cdmft_distance_multi(int n, int nb, int nw, const double* w, const double* weight,
                             int nhost, const double* Ghost, int B, const int* host_of,
                             const double* eb, const double* tcb, double scale, double* d) {
  return guard([&] { distance(n, nb, nw, w, weight, nhost, Ghost, B, host_of, eb, tcb, scale, d); });
}

int qed_cdmft_fit_batch(int n, int nb, int nw, const double* w, const double* weight, int nhost,
                        const double* Ghost, int B, const int* host_of, int nvar, const double* E,
                        const double* T, const double* x0, double initial_step, double xatol,
                        double fatol, int maxfev, double scale, double* x, double* f, int* nfev,
                        int* status) {
  return guard([&] {
    if (B <= 0) return;
    if (n <= 0 || nb <= 0 || nw <= 0 || nhost <= 0 || nvar <= 0) fail("cdmft fit: bad sizes");
    if (nvar > 64) fail("cdmft fit: at most 64 variational parameters");
    for (int b = 0; b < B; ++b)
      if (host_of[b] < 0 || host_of[b] >= nhost) fail("cdmft fit: host index out of range");
    const int N = nvar;
    size_t lds_d = (size_t)nw * n * n * 2 + (size_t)(N + 1) * N + (N + 1) + N + (size_t)N * N + N +
                   nb + (size_t)n * nb + wave_scratch(n, nb, nw) + (size_t)N * nb + (size_t)N * n * nb + 2 * nw;
    size_t lds = lds_d * sizeof(double);
    if (lds > 64 * 1024) fail("cdmft fit: problem too large for one workgroup's LDS");
    DevBuf dw = upload(w, nw), dwt = upload(weight, nw);
    DevBuf dg = upload(Ghost, (size_t)nhost * nw * n * n * 2);
    DevBuf dh = upload(host_of, B);
    DevBuf dE = upload(E, (size_t)nvar * nb), dT = upload(T, (size_t)nvar * n * nb);
    DevBuf dx0 = upload(x0, (size_t)B * N);
    DevBuf dx((size_t)B * N * 8), df((size_t)B * 8), dn((size_t)B * 4), ds((size_t)B * 4);
    FitArgs A{n, nb, nw, nvar, dw.as<double>(), dwt.as<double>(), dg.as<double>(), dh.as<int>(),
              dE.as<double>(), dT.as<double>(), dx0.as<double>(), initial_step, xatol, fatol,
              scale, maxfev, dx.as<double>(), df.as<double>(), dn.as<int>(), ds.as<int>()};
    k_nelder_mead<<<B, kFitThreads, lds, stream()>>>(A);
    QED_HIP(hipGetLastError());
    QED_HIP(hipMemcpyAsync(x, dx.p, (size_t)B * N * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipMemcpyAsync(f, df.p, (size_t)B * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipMemcpyAsync(nfev, dn.p, (size_t)B * 4, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipMemcpyAsync(status, ds.p, (size_t)B * 4, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"
