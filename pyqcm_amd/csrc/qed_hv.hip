// Sector construction (device-side ranking by the combinatorial number system) and the
// on-the-fly factorized H.v kernels.
//
// Layout in HBM: a sector vector is the reference's factorized order I = i_up + du * i_dn
// (src_ed/ED_basis.cpp:142-165), i.e. a dd x du row-major FP64 matrix X[i_dn][i_up].
//   up-spin hops act inside one row  (gathers within X[r, :])
//   down-spin hops combine rows      (Y[r, :] += v * X[r', :])
// Hop tables are CSR over half states with packed uint32 entries
// (target rank | value index << 20) and a small table of signed amplitudes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <atomic>

#include "../../include/qed.h"
#include "qed_sector.hpp"

namespace qed {

namespace {

constexpr uint32_t kTargetBits = kEllTargetBits;
constexpr uint32_t kTargetMask = (1u << kTargetBits) - 1u;
constexpr uint32_t kMaxVals = 1u << (32 - kTargetBits);

struct Bond {
  int from, to;  // c†_to c_from
};

__global__ void k_unrank(uint32_t* out, uint32_t n, int L, int N, const uint32_t* binom) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cns_unrank(i, L, N, binom);
}

__global__ void k_rank(const uint32_t* states, uint32_t* ranks, uint64_t n, int N, uint32_t dim,
                       uint32_t fullmask, const uint32_t* binom) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s = states[i];
  ranks[i] = ((s & ~fullmask) == 0 && __popc(s) == N) ? cns_rank(s, binom) : dim;
}

// number of allowed hops per half state
__global__ void k_hop_count(const uint32_t* bin, uint32_t dim, const Bond* bonds, int nb,
                            uint32_t* cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  uint32_t s = bin[i], c = 0;
  for (int b = 0; b < nb; ++b) {
    uint32_t mf = 1u << bonds[b].from, mt = 1u << bonds[b].to;
    c += ((s & mf) && !(s & mt)) ? 1u : 0u;
  }
  cnt[i] = c;
}

// hop entries (target rank by CNS, Jordan-Wigner sign) and the one-body diagonal
__global__ void k_hop_fill(const uint32_t* bin, uint32_t dim, const Bond* bonds, int nb,
                           const uint32_t* ptr, uint32_t* ent, const int* dsite,
                           const double* dval, int nd, double* diag, const uint32_t* binom) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  uint32_t s = bin[i], k = ptr[i];
  for (int b = 0; b < nb; ++b) {
    int f = bonds[b].from, t = bonds[b].to;
    uint32_t mf = 1u << f, mt = 1u << t;
    if (!(s & mf) || (s & mt)) continue;
    int lo = f < t ? f : t, hi = f < t ? t : f;
    uint32_t between = ((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u);
    int neg = __popc(s & between) & 1;
    uint32_t j = cns_rank(s ^ mf ^ mt, binom);
    ent[k++] = j | ((uint32_t)(2 * b + neg) << kTargetBits);
  }
  if (!diag) return;
  double d = 0.0;
  for (int q = 0; q < nd; ++q)
    if (s & (1u << dsite[q])) d += dval[q];
  diag[i] = d;
}

// density-density energy of the factorized basis state (su, sd)
__device__ __forceinline__ double interaction(const HvView& v, uint32_t su, uint32_t sd) {
  return density_density(v, su, sd);
}

// ---------------------------------------------------------------------------------------
// H.v, one workgroup per down row r: X[r, :] staged in LDS for the up-spin gathers, the
// down-spin row combinations read straight from HBM (coalesced along the row).
template <int BS, bool DOWN, bool LDSX>
__global__ __launch_bounds__(BS) void k_hv_rows(HvView v, const double* __restrict__ x,
                                                double* __restrict__ y) {
  extern __shared__ double smem[];
  const uint32_t du = v.du;
  double* vals_s = smem;
  double* xs = smem + v.up_nvals;
  const uint32_t r = blockIdx.x;
  const double* xr = x + (uint64_t)r * du;
  for (uint32_t i = threadIdx.x; i < v.up_nvals; i += BS) vals_s[i] = v.up_vals[i];
  if (LDSX)
    for (uint32_t c = threadIdx.x; c < du; c += BS) xs[c] = xr[c];
  __syncthreads();
  const double* xsrc = LDSX ? xs : xr;
  const uint32_t sd = v.bdn[r];
  const double dg = v.diag_dn[r];
  uint32_t d0 = 0, d1 = 0;
  if (DOWN) {
    d0 = v.dn_ptr[r];
    d1 = v.dn_ptr[r + 1];
  }
  double* yr = y + (uint64_t)r * du;
  for (uint32_t c = threadIdx.x; c < du; c += BS) {
    const uint32_t su = v.bup[c];
    double acc = (dg + v.diag_up[c] + interaction(v, su, sd)) * xsrc[c];
    const uint32_t k0 = v.up_ptr[c], k1 = v.up_ptr[c + 1];
    for (uint32_t k = k0; k < k1; ++k) {
      const uint32_t e = v.up_ent[k];
      acc += vals_s[e >> kTargetBits] * xsrc[e & kTargetMask];
    }
    if (DOWN) {
      for (uint32_t k = d0; k < d1; ++k) {
        const uint32_t e = v.dn_ent[k];
        acc += v.dn_vals[e >> kTargetBits] * x[(uint64_t)(e & kTargetMask) * du + c];
      }
    }
    yr[c] += acc;
  }
}

// H.v down-spin term only, tiled by column chunks of W doubles so that one XCD works on
// one chunk (dd x W x 8 bytes) at a time and serves the ~22 row re-reads from its L2.
// Workgroup w runs on XCD w % 8; the k-th workgroup of an XCD takes chunk
// (k / nrb) * 8 + xcd and row block k % nrb.
template <int W>
__global__ __launch_bounds__(256) void k_hv_down_chunks(HvView v, const double* __restrict__ x,
                                                        double* __restrict__ y, uint32_t nchunks,
                                                        uint32_t rows_per_block, uint32_t nrb) {
  const uint32_t w = blockIdx.x;
  const uint32_t xcd = w & 7u, kk = w >> 3;
  const uint32_t chunk = (kk / nrb) * 8u + xcd;
  const uint32_t rb = kk % nrb;
  if (chunk >= nchunks) return;
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint32_t c = chunk * W + (uint32_t)(lane % W);
  const uint32_t sub = (uint32_t)(lane / W);
  const uint32_t du = v.du, dd = v.dd;
  const uint32_t r0 = rb * rows_per_block;
  for (uint32_t rr = (uint32_t)wid * RPW + sub; rr < rows_per_block; rr += 4u * RPW) {
    const uint32_t r = r0 + rr;
    if (r >= dd) break;
    if (c < du) {
      double acc = 0.0;
      const uint32_t k0 = v.dn_ptr[r], k1 = v.dn_ptr[r + 1];
      for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t e = v.dn_ent[k];
        acc += v.dn_vals[e >> kTargetBits] * x[(uint64_t)(e & kTargetMask) * du + c];
      }
      y[(uint64_t)r * du + c] += acc;
    }
  }
}

// CSR -> ELL, padding with zero-amplitude self hops (value index zv holds 0.0)
__global__ void k_ell_colmajor(const uint32_t* ptr, const uint32_t* ent, uint32_t dim, uint32_t W,
                               uint32_t zv, uint32_t* out) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  uint32_t k0 = ptr[c], n = ptr[c + 1] - k0;
  for (uint32_t k = 0; k < W; ++k)
    out[(uint64_t)k * dim + c] = (k < n) ? ent[k0 + k] : (c | (zv << kTargetBits));
}

__global__ void k_ell_rowmajor(const uint32_t* ptr, const uint32_t* ent, uint32_t dim, uint32_t W,
                               uint32_t zv, uint32_t* out) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= dim) return;
  uint32_t k0 = ptr[r], n = ptr[r + 1] - k0;
  for (uint32_t k = 0; k < W; ++k)
    out[(uint64_t)r * W + k] = (k < n) ? ent[k0 + k] : (r | (zv << kTargetBits));
}

// H.v, one workgroup per down row r. Up-spin hops: ELL entries (coalesced per lane) gather
// from the row staged in LDS. Down-spin hops: the row's neighbour list is uniform across
// the workgroup; 8 neighbour rows are requested before any of them is consumed so that
// each wave keeps 8 row loads in flight.
constexpr int kG = 8;

// Row r of y += H x for one sector: the body shared by the single-sector kernel and the
// batched multi-sector kernel. smem: nvu amplitudes, then (LDSX) the row X[r, :].
template <int BS, bool LDSX>
__device__ __forceinline__ void ell_row(const HvView& v, const uint32_t* __restrict__ up_ell,
                                        uint32_t Wu, const uint32_t* __restrict__ dn_ell,
                                        uint32_t Wd, const double* __restrict__ x,
                                        double* __restrict__ y, uint32_t r, double* smem) {
  const uint32_t du = v.du;
  const uint32_t nvu = v.up_nvals;
  double* vals_s = smem;
  double* xs = smem + nvu;
  const double* xr = x + (uint64_t)r * du;
  for (uint32_t i = threadIdx.x; i < nvu; i += BS) vals_s[i] = v.up_vals[i];
  if (LDSX)
    for (uint32_t c = threadIdx.x; c < du; c += BS) xs[c] = xr[c];
  __syncthreads();
  const double* xsrc = LDSX ? xs : xr;
  const uint32_t sd = v.bdn[r];
  const double dg = v.diag_dn[r];
  const uint32_t dcnt = v.dn_ptr[r + 1] - v.dn_ptr[r];
  const uint32_t dgroups = (dcnt + kG - 1) / kG * kG;
  const uint32_t* dl = dn_ell + (uint64_t)r * Wd;
  double* yr = y + (uint64_t)r * du;
  for (uint32_t c = threadIdx.x; c < du; c += BS) {
    const uint32_t su = v.bup[c];
    double acc = (dg + v.diag_up[c] + interaction(v, su, sd)) * xsrc[c];
    double acc2 = 0.0;
    for (uint32_t g = 0; g < Wu; g += kG) {
      uint32_t e[kG];
#pragma unroll
      for (int q = 0; q < kG; ++q) e[q] = up_ell[(uint64_t)(g + q) * du + c];
#pragma unroll
      for (int q = 0; q < kG; ++q) acc += vals_s[e[q] >> kTargetBits] * xsrc[e[q] & kTargetMask];
    }
    for (uint32_t g = 0; g < dgroups; g += kG) {
      uint32_t e[kG];
      double xv[kG];
#pragma unroll
      for (int q = 0; q < kG; ++q) e[q] = dl[g + q];
#pragma unroll
      for (int q = 0; q < kG; ++q) xv[q] = x[(uint64_t)(e[q] & kTargetMask) * du + c];
#pragma unroll
      for (int q = 0; q < kG; ++q) acc2 += v.dn_vals[e[q] >> kTargetBits] * xv[q];
    }
    yr[c] += acc + acc2;
  }
}

// H.v, one workgroup per down row r. Up-spin hops: ELL entries (coalesced per lane) gather
// from the row staged in LDS. Down-spin hops: the row's neighbour list is uniform across
// the workgroup; 8 neighbour rows are requested before any of them is consumed so that
// each wave keeps 8 row loads in flight.
template <int BS, bool LDSX>
__global__ __launch_bounds__(BS) void k_hv_ell(HvView v, const uint32_t* __restrict__ up_ell,
                                               uint32_t Wu, const uint32_t* __restrict__ dn_ell,
                                               uint32_t Wd, const double* __restrict__ x,
                                               double* __restrict__ y) {
  extern __shared__ double smem[];
  ell_row<BS, LDSX>(v, up_ell, Wu, dn_ell, Wd, x, y, blockIdx.x, smem);
}

// Batched H.v over many sectors ("jobs") in one persistent launch: the rows of all jobs
// form one tile list (job j owns tiles row0[j] .. row0[j] + dd_j - 1); a fixed grid walks
// it with a stride. Jobs whose active flag is 0 are skipped. Parity selects which of the
// job's two vector pairs is (x, y), so that an alternating recurrence needs only two job
// tables.
template <int BS>
__global__ __launch_bounds__(BS) void k_hv_ell_batch(const HvJob* __restrict__ jobs, int njobs,
                                                     uint32_t ntiles, int parity,
                                                     const int* __restrict__ active,
                                                     const uint32_t* __restrict__ tile_job) {
  extern __shared__ double smem[];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int lo = (int)tile_job[t];  // job owning tile t (row0[lo] <= t < row0[lo + 1])
    const HvJob& J = jobs[lo];
    if (active[lo]) {
      ell_row<BS, true>(J.v, J.up_ell, J.Wu, J.dn_ell, J.Wd, J.x[parity], J.y[parity], t - J.row0,
                        smem);
    }
    __syncthreads();
  }
}

// Two-pass tiled form of the batched product for sectors with short rows (the impurity
// problems of a CDMFT batch: up to a few thousand states per spin). The row kernel above
// re-reads the whole up-spin hop table for every row and about twenty neighbour rows per
// row for the down-spin hops, both from L2, for every row it produces. Here
//   pass A (row tiles): R consecutive rows of X in LDS; each thread holds the hop-table
//     entries of one column and applies them to the R rows (table traffic / R), with the
//     diagonal: y[r, :] += (diag + up hops) x[r, :];
//   pass B (column tiles): the W-column panel X[:, c0:c0+W] of all rows in LDS; the
//     down-spin hops are gathers inside it: y[:, c0:c0+W] += (down hops) x[:, c0:c0+W].
// Each pass streams X once and y once; the re-reads are LDS reads.
__device__ __forceinline__ int find_job(const HvJob* __restrict__ jobs, int njobs, uint32_t t, bool pass_b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((pass_b ? jobs[mid].b0 : jobs[mid].a0) <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int R>
__global__ __launch_bounds__(256) void k_hv_tileA_batch(const HvJob* __restrict__ jobs, int njobs,
                                                        uint32_t ntiles, int parity,
                                                        const int* __restrict__ active) {
  extern __shared__ double smem[];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int j = find_job(jobs, njobs, t, false);
    const HvJob& J = jobs[j];
    if (active[j]) {
      const HvView& v = J.v;
      const uint32_t du = v.du, dd = v.dd, nvu = v.up_nvals;
      const uint32_t r0 = (t - J.a0) * R;
      const uint32_t nr = min((uint32_t)R, dd - r0);
      double* vals = smem;
      double* xs = smem + ((nvu + 1u) & ~1u);
      const double* __restrict__ x = J.x[parity];
      double* __restrict__ y = J.y[parity];
      for (uint32_t i = threadIdx.x; i < nvu; i += 256) vals[i] = v.up_vals[i];
      const double* xr = x + (uint64_t)r0 * du;
      const uint32_t na = nr * du;
      for (uint32_t i0 = 0; i0 < na; i0 += 256 * kG) {  // kG loads in flight per thread
        double t[kG];
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const uint32_t i = i0 + threadIdx.x + 256u * k;
          t[k] = i < na ? xr[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const uint32_t i = i0 + threadIdx.x + 256u * k;
          if (i < na) xs[i] = t[k];
        }
      }
      __syncthreads();
      for (uint32_t c = threadIdx.x; c < du; c += 256) {
        const uint32_t su = v.bup[c];
        const double dgu = v.diag_up[c];
        double acc[R];
#pragma unroll
        for (int q = 0; q < R; ++q)
          acc[q] = (q < (int)nr) ? (v.diag_dn[r0 + q] + dgu + interaction(v, su, v.bdn[r0 + q])) * xs[q * du + c]
                                 : 0.0;
        for (uint32_t g = 0; g < J.Wu; g += kG) {
          uint32_t e[kG];
#pragma unroll
          for (int k = 0; k < kG; ++k) e[k] = J.up_ell[(uint64_t)(g + k) * du + c];
#pragma unroll
          for (int k = 0; k < kG; ++k) {
            const double cf = vals[e[k] >> kTargetBits];
            const uint32_t tg = e[k] & kTargetMask;
#pragma unroll
            for (int q = 0; q < R; ++q)
              if (q < (int)nr) acc[q] += cf * xs[q * du + tg];
          }
        }
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (q < (int)nr) y[(uint64_t)(r0 + q) * du + c] += acc[q];
      }
    }
    __syncthreads();
  }
}

template <int W>
__global__ __launch_bounds__(256) void k_hv_tileB_batch(const HvJob* __restrict__ jobs, int njobs,
                                                        uint32_t ntiles, int parity,
                                                        const int* __restrict__ active) {
  extern __shared__ double smem[];
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int j = find_job(jobs, njobs, t, true);
    const HvJob& J = jobs[j];
    if (active[j]) {
      const HvView& v = J.v;
      const uint32_t du = v.du, dd = v.dd;
      const uint32_t c0 = (t - J.b0) * W;
      const uint32_t nc = min((uint32_t)W, du - c0);
      const double* __restrict__ x = J.x[parity];
      double* __restrict__ y = J.y[parity];
      double* xs = smem;  // dd x W
      const uint32_t nb = dd * W;
      for (uint32_t i0 = 0; i0 < nb; i0 += 256 * kG) {  // kG loads in flight per thread
        double t[kG];
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const uint32_t i = i0 + threadIdx.x + 256u * k;
          const uint32_t r = i / W, cc = i % W;
          t[k] = (i < nb && cc < nc) ? x[(uint64_t)r * du + c0 + cc] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const uint32_t i = i0 + threadIdx.x + 256u * k;
          if (i < nb) xs[i] = t[k];
        }
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < dd * W; i += 256) {
        const uint32_t r = i / W, cc = i % W;
        const uint32_t* dl = J.dn_ell + (uint64_t)r * J.Wd;
        double acc = 0.0;
        for (uint32_t g = 0; g < J.Wd; g += kG) {
          uint32_t e[kG];
#pragma unroll
          for (int k = 0; k < kG; ++k) e[k] = dl[g + k];
#pragma unroll
          for (int k = 0; k < kG; ++k) acc += v.dn_vals[e[k] >> kTargetBits] * xs[(e[k] & kTargetMask) * W + cc];
        }
        if (cc < nc) y[(uint64_t)r * du + c0 + cc] += acc;
      }
    }
    __syncthreads();
  }
}

// Same product, two adjacent columns per thread: 16-byte loads of the down-spin neighbour
// rows and of the ELL entry pairs (needs an even up dimension).
// LZ (one Lanczos step fused into the product): y <- lz[0] (H x) + lz[1] y, and the
// workgroup's part of <x, y_new> in lz_part[row] (the scalars are device values, so that
// a step needs no host round trip before it starts).
template <int BS, bool DOWN = true, bool LZ = false>
__global__ __launch_bounds__(BS) void k_hv_ell2(HvView v, const uint32_t* __restrict__ up_ell,
                                                uint32_t Wu, const uint32_t* __restrict__ dn_ell,
                                                uint32_t Wd, const double* __restrict__ x,
                                                double* __restrict__ y,
                                                const uint32_t* __restrict__ order = nullptr,
                                                const double* __restrict__ lz = nullptr,
                                                double* __restrict__ lz_part = nullptr) {
  extern __shared__ double smem[];
  const uint32_t du = v.du;
  const uint32_t nvu = v.up_nvals;
  double* vals_s = smem;
  double* xs = smem + ((nvu + 1u) & ~1u);  // 16-byte aligned row
  const uint32_t r = order ? order[blockIdx.x] : blockIdx.x;
  const double* xr = x + (uint64_t)r * du;
  const uint32_t dh = du / 2;
  for (uint32_t i = threadIdx.x; i < nvu; i += BS) vals_s[i] = v.up_vals[i];
  for (uint32_t c = threadIdx.x; c < dh; c += BS)
    reinterpret_cast<double2*>(xs)[c] = reinterpret_cast<const double2*>(xr)[c];
  __syncthreads();
  const uint32_t sd = v.bdn[r];
  const double dg = v.diag_dn[r];
  const uint32_t dcnt = v.dn_ptr[r + 1] - v.dn_ptr[r];
  const uint32_t dgroups = (dcnt + kG - 1) / kG * kG;
  const uint32_t* dl = dn_ell + (uint64_t)r * Wd;
  double2* yr = reinterpret_cast<double2*>(y + (uint64_t)r * du);
  double dotp = 0.0;
  for (uint32_t c2 = threadIdx.x; c2 < dh; c2 += BS) {
    const uint32_t c = 2 * c2;
    double a0 = (dg + v.diag_up[c] + interaction(v, v.bup[c], sd)) * xs[c];
    double a1 = (dg + v.diag_up[c + 1] + interaction(v, v.bup[c + 1], sd)) * xs[c + 1];
    for (uint32_t g = 0; g < Wu; g += kG) {
      uint2 e[kG];
#pragma unroll
      for (int q = 0; q < kG; ++q)
        e[q] = *reinterpret_cast<const uint2*>(up_ell + (uint64_t)(g + q) * du + c);
#pragma unroll
      for (int q = 0; q < kG; ++q) {
        a0 += vals_s[e[q].x >> kTargetBits] * xs[e[q].x & kTargetMask];
        a1 += vals_s[e[q].y >> kTargetBits] * xs[e[q].y & kTargetMask];
      }
    }
    double b0 = 0.0, b1 = 0.0;
    for (uint32_t g = 0; DOWN && g < dgroups; g += kG) {
      uint32_t e[kG];
      double2 xv[kG];
#pragma unroll
      for (int q = 0; q < kG; ++q) e[q] = dl[g + q];
#pragma unroll
      for (int q = 0; q < kG; ++q)
        xv[q] = *reinterpret_cast<const double2*>(x + (uint64_t)(e[q] & kTargetMask) * du + c);
#pragma unroll
      for (int q = 0; q < kG; ++q) {
        const double cf = v.dn_vals[e[q] >> kTargetBits];
        b0 += cf * xv[q].x;
        b1 += cf * xv[q].y;
      }
    }
    double2 yo = yr[c2];
    if (LZ) {
      yo.x = lz[0] * (a0 + b0) + lz[1] * yo.x;
      yo.y = lz[0] * (a1 + b1) + lz[1] * yo.y;
      dotp += xs[c] * yo.x + xs[c + 1] * yo.y;
    } else {
      yo.x += a0 + b0;
      yo.y += a1 + b1;
    }
    yr[c2] = yo;
  }
  if (LZ) {
    const double t = block_sum(dotp, smem + ((nvu + 1u) & ~1u) + du);
    if (threadIdx.x == 0) lz_part[r] = t;
  }
}

__global__ void k_diag_interaction(HvView v, double* d, uint64_t D) {
  uint64_t I = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= D) return;
  uint32_t iu = (uint32_t)(I % v.du), id = (uint32_t)(I / v.du);
  d[I] += interaction(v, v.bup[iu], v.bdn[id]);
}

// y[T] += z * sign * x[S(T)] over the target sector (creation) or the target of a
// destruction, both as gathers: c_a maps source basis states injectively.
__global__ void k_apply_c(const uint32_t* tu, const uint32_t* td, uint32_t tdu, uint64_t DT,
                          int L, int Nsu, int Nsd, int site, int down, int pm, double z,
                          const double* __restrict__ x, uint32_t sdu, double* __restrict__ y,
                          const uint32_t* binom) {
  uint64_t T = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (T >= DT) return;
  uint32_t iu = (uint32_t)(T % tdu), id = (uint32_t)(T / tdu);
  uint32_t su = tu[iu], sd = td[id];
  uint32_t m = 1u << site;
  // source state: creation -> target had the orbital occupied, source = target minus it;
  //               destruction -> source = target plus the orbital.
  uint32_t ssu = su, ssd = sd;
  if (pm > 0) {
    if (down) { if (!(sd & m)) return; ssd = sd ^ m; }
    else      { if (!(su & m)) return; ssu = su ^ m; }
  } else {
    if (down) { if (sd & m) return; ssd = sd | m; }
    else      { if (su & m) return; ssu = su | m; }
  }
  // parity of c_a acting on the state that holds orbital a (the target for creation,
  // the source for destruction): same-spin orbitals below a, plus all up orbitals for a
  // down orbital (up orbitals are ordered first)
  uint32_t hu = (pm > 0) ? su : ssu, hd = (pm > 0) ? sd : ssd;
  int par = down ? (__popc(hd & (m - 1u)) + __popc(hu)) : __popc(hu & (m - 1u));
  double sgn = (par & 1) ? -1.0 : 1.0;
  uint32_t ju = cns_rank(ssu, binom), jd = cns_rank(ssd, binom);
  (void)Nsu;
  (void)Nsd;
  y[T] += z * sgn * x[(uint64_t)jd * sdu + ju];
}

struct HalfTables {
  DevBuf bin, ptr, ent, vals, diag;
  uint32_t dim = 0, nnz = 0, nvals = 0, maxdeg = 0;
};

HalfTables build_half_tables(int L, int N, const std::vector<Bond>& bonds,
                             const std::vector<double>& bval, const std::vector<int>& dsite,
                             const std::vector<double>& dval) {
  HalfTables h;
  const auto& B = binom_host();
  if (N < 0 || N > L) return h;
  h.dim = B[L * 33 + N];
  if (h.dim == 0) return h;
  if (h.dim > kTargetMask) fail("half basis too large for the packed hop tables");
  if (2 * bonds.size() > kMaxVals) fail("too many distinct hopping amplitudes in one sector");
  hipStream_t st = stream();
  const uint32_t* bd = binom_dev();
  h.bin.alloc(h.dim * sizeof(uint32_t));
  k_unrank<<<(h.dim + 255) / 256, 256, 0, st>>>(h.bin.as<uint32_t>(), h.dim, L, N, bd);
  DevBuf bonds_d = upload(bonds.empty() ? std::vector<Bond>{{0, 0}} : bonds);
  DevBuf cnt(h.dim * sizeof(uint32_t));
  k_hop_count<<<(h.dim + 255) / 256, 256, 0, st>>>(h.bin.as<uint32_t>(), h.dim,
                                                   bonds_d.as<Bond>(), (int)bonds.size(),
                                                   cnt.as<uint32_t>());
  std::vector<uint32_t> c(h.dim), ptr(h.dim + 1, 0);
  QED_HIP(hipMemcpyAsync(c.data(), cnt.p, h.dim * 4, hipMemcpyDeviceToHost, st));
  QED_HIP(hipStreamSynchronize(st));
  for (uint32_t i = 0; i < h.dim; ++i) {
    ptr[i + 1] = ptr[i] + c[i];
    h.maxdeg = std::max(h.maxdeg, c[i]);
  }
  h.nnz = ptr[h.dim];
  h.ptr = upload(ptr);
  h.ent.alloc(std::max<uint32_t>(h.nnz, 1) * sizeof(uint32_t));
  std::vector<double> vals(2 * bonds.size() + 1, 0.0);
  for (size_t b = 0; b < bonds.size(); ++b) {
    vals[2 * b] = bval[b];
    vals[2 * b + 1] = -bval[b];
  }
  h.nvals = (uint32_t)vals.size();
  h.vals = upload(vals);
  DevBuf ds = upload(dsite.empty() ? std::vector<int>{0} : dsite);
  DevBuf dv = upload(dval.empty() ? std::vector<double>{0.0} : dval);
  h.diag.alloc(h.dim * sizeof(double));
  k_hop_fill<<<(h.dim + 255) / 256, 256, 0, st>>>(
      h.bin.as<uint32_t>(), h.dim, bonds_d.as<Bond>(), (int)bonds.size(), h.ptr.as<uint32_t>(),
      h.ent.as<uint32_t>(), ds.as<int>(), dv.as<double>(), (int)dsite.size(), h.diag.as<double>(),
      bd);
  QED_HIP(hipStreamSynchronize(st));
  return h;
}

// Structure of one half space for a bond list: ascending basis, CSR hop lists and their
// ELL copy (up spin column-major, down spin row-major). It does not depend on the
// amplitudes, so sectors of the same model with other parameter values share it.
struct HalfStruct {
  DevBuf bin, ptr, ent, ell;
  std::vector<uint32_t> bin_h;
  uint32_t dim = 0, nnz = 0, maxdeg = 0, w = 0, nvals = 0;
};

std::shared_ptr<HalfStruct> half_struct(int L, int N, const std::vector<Bond>& bonds, bool colmajor) {
  static std::mutex mu;
  static std::map<std::vector<int>, std::shared_ptr<HalfStruct>> cache;
  std::vector<int> key{L, N, colmajor ? 1 : 0};
  for (const Bond& b : bonds) {
    key.push_back(b.from);
    key.push_back(b.to);
  }
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  auto h = std::make_shared<HalfStruct>();
  const auto& B = binom_host();
  if (N >= 0 && N <= L) h->dim = B[L * 33 + N];
  h->nvals = (uint32_t)(2 * bonds.size() + 1);
  if (h->dim > 0) {
    if (h->dim > kTargetMask) fail("half basis too large for the packed hop tables");
    if (2 * bonds.size() > kMaxVals) fail("too many distinct hopping amplitudes in one sector");
    hipStream_t st = stream();
    const uint32_t* bd = binom_dev();
    h->bin.alloc(h->dim * sizeof(uint32_t));
    k_unrank<<<(h->dim + 255) / 256, 256, 0, st>>>(h->bin.as<uint32_t>(), h->dim, L, N, bd);
    DevBuf bonds_d = upload(bonds.empty() ? std::vector<Bond>{{0, 0}} : bonds);
    DevBuf cnt(h->dim * sizeof(uint32_t));
    k_hop_count<<<(h->dim + 255) / 256, 256, 0, st>>>(h->bin.as<uint32_t>(), h->dim, bonds_d.as<Bond>(),
                                                      (int)bonds.size(), cnt.as<uint32_t>());
    std::vector<uint32_t> c(h->dim), ptr(h->dim + 1, 0);
    h->bin_h.resize(h->dim);
    QED_HIP(hipMemcpyAsync(c.data(), cnt.p, h->dim * 4, hipMemcpyDeviceToHost, st));
    QED_HIP(hipMemcpyAsync(h->bin_h.data(), h->bin.p, h->dim * 4, hipMemcpyDeviceToHost, st));
    QED_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < h->dim; ++i) {
      ptr[i + 1] = ptr[i] + c[i];
      h->maxdeg = std::max(h->maxdeg, c[i]);
    }
    h->nnz = ptr[h->dim];
    h->ptr = upload(ptr);
    h->ent.alloc(std::max<uint32_t>(h->nnz, 1) * sizeof(uint32_t));
    k_hop_fill<<<(h->dim + 255) / 256, 256, 0, st>>>(h->bin.as<uint32_t>(), h->dim, bonds_d.as<Bond>(),
                                                     (int)bonds.size(), h->ptr.as<uint32_t>(),
                                                     h->ent.as<uint32_t>(), nullptr, nullptr, 0, nullptr, bd);
    // ELL copy, padded to multiples of 8 entries with zero-amplitude self hops
    h->w = (h->maxdeg + 7) / 8 * 8;
    h->ell.alloc(std::max<uint64_t>((uint64_t)h->w * h->dim, 1) * 4);
    if (colmajor)
      k_ell_colmajor<<<(h->dim + 255) / 256, 256, 0, st>>>(h->ptr.as<uint32_t>(), h->ent.as<uint32_t>(), h->dim,
                                                           h->w, h->nvals - 1, h->ell.as<uint32_t>());
    else
      k_ell_rowmajor<<<(h->dim + 255) / 256, 256, 0, st>>>(h->ptr.as<uint32_t>(), h->ent.as<uint32_t>(), h->dim,
                                                           h->w, h->nvals - 1, h->ell.as<uint32_t>());
    QED_HIP(hipGetLastError());
    QED_HIP(hipStreamSynchronize(st));
  }
  if (cache.size() > 512) cache.clear();
  cache[key] = h;
  return h;
}

// one-body diagonal of every half state (same order of additions as the device fill)
void half_diag(const HalfStruct& h, const std::vector<int>& dsite, const std::vector<double>& dval, double* d) {
  for (uint32_t i = 0; i < h.dim; ++i) {
    const uint32_t s = h.bin_h[i];
    double a = 0.0;
    for (size_t q = 0; q < dsite.size(); ++q)
      if (s & (1u << dsite[q])) a += dval[q];
    d[i] = a;
  }
}

}  // namespace

HvView view(const qed_sector& s) {
  HvView v;
  v.bup = s.bup.as<uint32_t>();
  v.bdn = s.bdn.as<uint32_t>();
  v.up_ptr = s.up_ptr.as<uint32_t>();
  v.up_ent = s.up_ent.as<uint32_t>();
  v.dn_ptr = s.dn_ptr.as<uint32_t>();
  v.dn_ent = s.dn_ent.as<uint32_t>();
  v.up_vals = s.up_vals.as<double>();
  v.dn_vals = s.dn_vals.as<double>();
  v.diag_up = s.diag_up.as<double>();
  v.diag_dn = s.diag_dn.as<double>();
  v.int_mask = s.int_mask_d.as<uint64_t>();
  v.int_val = s.int_val_d.as<double>();
  v.du = s.du;
  v.dd = s.dd;
  v.up_nvals = s.up_nvals;
  v.n_int = (int)s.int_mask.size();
  v.hubbard_uniform = s.hubbard_uniform ? 1 : 0;
  v.hub_U = s.hub_U;
  v.hub_sites = s.hub_sites;
  return v;
}

static size_t lds_budget() { return 160 * 1024 - 1024; }

static void set_lds_attributes() {
  static std::once_flag once;
  std::call_once(once, [] {
    const int bytes = (int)lds_budget();
    (void)hipFuncSetAttribute((const void*)k_hv_rows<1024, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_rows<1024, false, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_rows<256, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_rows<256, false, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell<1024, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell<256, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell2<512>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell2<512, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell2<512, true, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);


  });
}

// Kernel variants of y += H x:
//   0  one workgroup per row, CSR hop lists (straightforward form, kept as a baseline)
//   1  two passes: up-spin + diagonal per row, then the down-spin term by L2-sized column chunks
//   4  one workgroup per row, ELL tables, 8-deep load groups, one column per thread
//   5  as 4 with two columns per thread and 16-byte loads (default; falls back to 4 when the
//      up dimension is odd or small, or the row does not fit in LDS)
//   12 variant 5 visiting the rows in the tiled order below (default)
//   20 the split-halves passes (qed_hvsplit.hip; 21, 22: one pass alone, for timing)
// Rows of the down-spin half basis in 2D tiles: a state splits into its low and high
// halves of orbitals (a, b); rows are grouped by the particle number of b, then visited
// in T x T tiles of (rank of b, rank of a). Hops inside one half change only a or only b.
std::vector<uint32_t> tiled_row_order(const qed_sector& s, int T) {
  std::vector<uint32_t> st(s.dd);
  QED_HIP(hipMemcpy(st.data(), s.bdn.p, s.dd * sizeof(uint32_t), hipMemcpyDeviceToHost));
  const int half = s.L / 2;
  const uint32_t lo_mask = (1u << half) - 1u;
  std::map<uint32_t, uint32_t> arank, brank;  // ascending order within one particle number
  std::map<int, std::vector<uint32_t>> as, bs;
  for (uint32_t v : st) {
    as[__builtin_popcount(v & lo_mask)].push_back(v & lo_mask);
    bs[__builtin_popcount(v >> half)].push_back(v >> half);
  }
  for (auto& kv : as) {
    std::sort(kv.second.begin(), kv.second.end());
    kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
    for (size_t i = 0; i < kv.second.size(); ++i) arank[kv.second[i]] = (uint32_t)i;
  }
  for (auto& kv : bs) {
    std::sort(kv.second.begin(), kv.second.end());
    kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
    for (size_t i = 0; i < kv.second.size(); ++i) brank[kv.second[i]] = (uint32_t)i;
  }
  struct Key {
    uint32_t nb, bt, at, br, ar, row;
  };
  std::vector<Key> k(s.dd);
  for (uint32_t r = 0; r < s.dd; ++r) {
    const uint32_t a = st[r] & lo_mask, b = st[r] >> half;
    const uint32_t ar = arank[a], br = brank[b];
    k[r] = {(uint32_t)__builtin_popcount(b), br / (uint32_t)T, ar / (uint32_t)T, br, ar, r};
  }
  std::sort(k.begin(), k.end(), [](const Key& p, const Key& q) {
    return std::tie(p.nb, p.bt, p.at, p.br, p.ar) < std::tie(q.nb, q.bt, q.at, q.br, q.ar);
  });
  std::vector<uint32_t> ord(s.dd);
  for (uint32_t i = 0; i < s.dd; ++i) ord[i] = k[i].row;
  return ord;
}

// the tiled row order of the default row kernel, built once per sector (at creation for
// the sectors that take that kernel, so that later products allocate nothing)
void ensure_row_order(const qed_sector& s) {
  if (s.row_order.p) return;
  std::vector<uint32_t> ord = tiled_row_order(s, 16);
  s.row_order.alloc(ord.size() * sizeof(uint32_t));
  QED_HIP(hipMemcpy(s.row_order.p, ord.data(), ord.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
}

void hv(const qed_sector& s, const double* x, double* y, int variant) {
  if (s.D == 0) return;
  if (s.mixed) {
    hv_mixed(s, x, y);
    return;
  }
  if (s.general) {
    hv_general(s, x, y, variant);
    return;
  }
  set_lds_attributes();
  if (variant < 0) variant = s.hv_variant;
  if (variant < 0) variant = 12;
  if (variant >= 20 && variant <= 22) {
    // split-halves passes (20 both; 21 row pass only, 22 tile pass only: timing splits)
    if (hv_split(s, x, y, variant == 20 ? 3 : variant == 21 ? 1 : 2)) return;
    variant = 12;
  }
  HvView v = view(s);
  hipStream_t st = stream();
  const uint32_t* ue = s.up_ell.as<uint32_t>();
  const uint32_t* de = s.dn_ell.as<uint32_t>();
  if (variant == 12) {
    // the row kernel visiting the rows in 2D tiles over the two halves of the down-spin
    // orbitals (neighbour rows of a tile share either half); measured on the 16-site
    // sector: tile side 16 gives 5.01 ms against 5.19 ms in rank order, identical results
    const size_t lds2 = ((size_t)s.du + s.up_nvals + 1) * sizeof(double);
    if (s.du % 2 == 0 && s.du >= 4096 && lds2 <= lds_budget()) {
      ensure_row_order(s);
      k_hv_ell2<512><<<s.dd, 512, lds2, st>>>(v, ue, s.up_w, de, s.dn_w, x, y, s.row_order.as<uint32_t>());
      QED_HIP(hipGetLastError());
      return;
    }
    variant = 5;
  }
  if (variant == 5) {
    const size_t lds2 = ((size_t)s.du + s.up_nvals + 1) * sizeof(double);
    if (s.du % 2 == 0 && s.du >= 4096 && lds2 <= lds_budget()) {
      k_hv_ell2<512><<<s.dd, 512, lds2, st>>>(v, ue, s.up_w, de, s.dn_w, x, y);
      QED_HIP(hipGetLastError());
      return;
    }
    variant = 4;
  }
  if (variant == 4) {
    const size_t lds = ((size_t)s.du + s.up_nvals) * sizeof(double);
    const bool fits = lds <= lds_budget();
    const size_t sh = fits ? lds : s.up_nvals * sizeof(double);
    if (s.du >= 4096) {
      if (fits) k_hv_ell<1024, true><<<s.dd, 1024, sh, st>>>(v, ue, s.up_w, de, s.dn_w, x, y);
      else k_hv_ell<1024, false><<<s.dd, 1024, sh, st>>>(v, ue, s.up_w, de, s.dn_w, x, y);
    } else {
      k_hv_ell<256, true><<<s.dd, 256, sh, st>>>(v, ue, s.up_w, de, s.dn_w, x, y);
    }
    QED_HIP(hipGetLastError());
    return;
  }
  if (variant != 0 && variant != 1) fail("unknown H.v kernel variant " + std::to_string(variant));
  const size_t lds_x = ((size_t)s.du + s.up_nvals) * sizeof(double);
  const bool fits = lds_x <= lds_budget();
  const size_t shm = fits ? lds_x : s.up_nvals * sizeof(double);
  const bool down = variant == 0;
  if (s.du >= 4096) {
    if (fits) {
      if (down) k_hv_rows<1024, true, true><<<s.dd, 1024, shm, st>>>(v, x, y);
      else k_hv_rows<1024, false, true><<<s.dd, 1024, shm, st>>>(v, x, y);
    } else {
      if (down) k_hv_rows<1024, true, false><<<s.dd, 1024, shm, st>>>(v, x, y);
      else k_hv_rows<1024, false, false><<<s.dd, 1024, shm, st>>>(v, x, y);
    }
  } else {
    if (down) k_hv_rows<256, true, true><<<s.dd, 256, shm, st>>>(v, x, y);
    else k_hv_rows<256, false, true><<<s.dd, 256, shm, st>>>(v, x, y);
  }
  if (variant == 1) {
    constexpr int W = 32;
    const uint32_t nchunks = (s.du + W - 1) / W;
    const uint32_t rpb = 64;
    const uint32_t nrb = (s.dd + rpb - 1) / rpb;
    const uint64_t groups = ((nchunks + 7) / 8) * 8ull * nrb;
    k_hv_down_chunks<W><<<(uint32_t)groups, 256, 0, st>>>(v, x, y, nchunks, rpb, nrb);
  }
  QED_HIP(hipGetLastError());
}


bool hv_lanczos(const qed_sector& s, const double* x, double* y, const double* lz, double* lz_part) {
  if (s.D == 0 || s.hv_variant >= 0 || !s.factorized()) return false;
  // the default row kernel (tiled row order); the LZ epilogue needs LDS for the row, the
  // amplitudes and eight doubles of reduction scratch
  const size_t lds2 = ((size_t)s.du + s.up_nvals + 1 + 16) * sizeof(double);
  if (s.du % 2 != 0 || s.du < 4096 || lds2 > lds_budget()) return false;
  set_lds_attributes();
  ensure_row_order(s);
  HvView v = view(s);
  k_hv_ell2<512, true, true><<<s.dd, 512, lds2, stream()>>>(v, s.up_ell.as<uint32_t>(), s.up_w,
                                                          s.dn_ell.as<uint32_t>(), s.dn_w, x, y,
                                                          s.row_order.as<uint32_t>(), lz, lz_part);
  QED_HIP(hipGetLastError());
  return true;
}

// ---- batched launch with autotuned (block size, grid) ------------------------------------
namespace {
struct TuneKey {
  std::vector<uint64_t> sig;
  bool operator<(const TuneKey& o) const { return sig < o.sig; }
};
std::map<TuneKey, HvBatchConfig>& tune_cache() {
  static std::map<TuneKey, HvBatchConfig> c;
  return c;
}

template <int R>
void launch_tileA(const HvBatch& B, uint32_t grid, int parity, const int* active, hipStream_t st) {
  k_hv_tileA_batch<R><<<grid, 256, B.lds_a, st>>>(B.jobs, B.njobs, B.ntiles_a, parity, active);
}
template <int W>
void launch_tileB(const HvBatch& B, uint32_t grid, int parity, const int* active, hipStream_t st) {
  k_hv_tileB_batch<W><<<grid, 256, B.lds_b, st>>>(B.jobs, B.njobs, B.ntiles_b, parity, active);
}

void launch_batch(const HvBatch& B, const HvBatchConfig& cfg, int parity, const int* active,
                  hipStream_t st) {
  if (cfg.mode == 1) {
    const uint32_t ga = cfg.grid ? std::min<uint32_t>(cfg.grid, B.ntiles_a) : B.ntiles_a;
    const uint32_t gb = cfg.grid ? std::min<uint32_t>(cfg.grid, B.ntiles_b) : B.ntiles_b;
    switch (B.R) {
      case 16: launch_tileA<16>(B, ga, parity, active, st); break;
      case 8: launch_tileA<8>(B, ga, parity, active, st); break;
      case 4: launch_tileA<4>(B, ga, parity, active, st); break;
      default: launch_tileA<2>(B, ga, parity, active, st); break;
    }
    switch (B.W) {
      case 32: launch_tileB<32>(B, gb, parity, active, st); break;
      case 16: launch_tileB<16>(B, gb, parity, active, st); break;
      case 8: launch_tileB<8>(B, gb, parity, active, st); break;
      default: launch_tileB<4>(B, gb, parity, active, st); break;
    }
    return;
  }
  const uint32_t grid = cfg.grid ? std::min<uint32_t>(cfg.grid, B.ntiles) : B.ntiles;
  const uint32_t* tj = B.tile_job.as<uint32_t>();
  if (cfg.block == 1024)
    k_hv_ell_batch<1024><<<grid, 1024, B.lds, st>>>(B.jobs, B.njobs, B.ntiles, parity, active, tj);
  else if (cfg.block == 512)
    k_hv_ell_batch<512><<<grid, 512, B.lds, st>>>(B.jobs, B.njobs, B.ntiles, parity, active, tj);
  else
    k_hv_ell_batch<256><<<grid, 256, B.lds, st>>>(B.jobs, B.njobs, B.ntiles, parity, active, tj);
}
}  // namespace

void hv_batch_prepare(HvBatch& B, const std::vector<const qed_sector*>& S) {
  std::vector<HvJob> h(S.size());
  uint32_t row = 0;
  size_t lds = 0;
  for (size_t j = 0; j < S.size(); ++j) {
    const qed_sector& s = *S[j];
    HvJob& J = h[j];
    J.v = view(s);
    J.up_ell = s.up_ell.as<uint32_t>();
    J.dn_ell = s.dn_ell.as<uint32_t>();
    J.Wu = s.up_w;
    J.Wd = s.dn_w;
    J.row0 = row;
    J.x[0] = J.x[1] = nullptr;
    J.y[0] = J.y[1] = nullptr;
    row += s.dd;
    lds = std::max(lds, ((size_t)s.du + s.up_nvals) * sizeof(double));
  }
  if (lds > lds_budget()) fail("batched H.v: a sector row does not fit in LDS");
  // tiled two-pass form: R rows (pass A) and W-column panels (pass B) within 32 KB of LDS
  constexpr size_t kTileLds = 32 * 1024;  // several workgroups per CU (measured: 64 KB tiles are slower)
  int R = 16, W = 32;
  for (const qed_sector* sp : S) {
    const qed_sector& s = *sp;
    while (R >= 2 && ((size_t)R * s.du + s.up_nvals + 1) * sizeof(double) > kTileLds) R /= 2;
    while (W >= 4 && (size_t)W * s.dd * sizeof(double) > kTileLds) W /= 2;
  }
  B.tiled = R >= 2 && W >= 4 && !S.empty();
  B.R = R;
  B.W = W;
  B.lds_a = B.lds_b = 0;
  uint32_t ta = 0, tb = 0;
  for (size_t j = 0; j < S.size(); ++j) {
    const qed_sector& s = *S[j];
    h[j].a0 = ta;
    h[j].b0 = tb;
    if (B.tiled) {
      ta += (s.dd + R - 1) / R;
      tb += (s.du + W - 1) / W;
      B.lds_a = std::max(B.lds_a, ((size_t)R * s.du + s.up_nvals + 1) * sizeof(double));
      B.lds_b = std::max(B.lds_b, (size_t)W * s.dd * sizeof(double));
    }
  }
  B.ntiles_a = ta;
  B.ntiles_b = tb;
  B.host = h;
  B.njobs = (int)S.size();
  B.ntiles = row;
  B.lds = lds;
  B.dev.alloc(std::max<size_t>(1, h.size()) * sizeof(HvJob));
  B.jobs = B.dev.as<HvJob>();
  // tile -> job map of the row kernel (the job mix is fixed for the batch's lifetime)
  std::vector<uint32_t> tj(std::max<uint32_t>(1, row), 0);
  for (size_t j = 0; j < S.size(); ++j)
    for (uint32_t r = 0; r < S[j]->dd; ++r) tj[h[j].row0 + r] = (uint32_t)j;
  B.tile_job = upload(tj);
  // autotuner key: the shape class of the batch (the launch shape that wins depends on
  // the row lengths and on the amount of work, not on the exact sector mix, so batches
  // of one CDMFT run share their tuning): largest half dimensions and table widths, the
  // tiling, and the total rows and amplitudes to within a factor of two
  uint64_t tot = 0;
  uint32_t mdu = 0, mdd = 0, mw = 0;
  for (const qed_sector* s : S) {
    tot += s->D;
    mdu = std::max(mdu, s->du);
    mdd = std::max(mdd, s->dd);
    mw = std::max(mw, std::max(s->up_w, s->dn_w));
  }
  auto log2u = [](uint64_t v) {
    int b = 0;
    while (v > 1) {
      v >>= 1;
      ++b;
    }
    return (uint64_t)b;
  };
  B.sig.assign({mdu, mdd, mw, (uint64_t)B.R, (uint64_t)B.W, log2u(row), log2u(tot)});
}

void hv_batch_upload(HvBatch& B) {
  QED_HIP(hipMemcpyAsync(B.jobs, B.host.data(), B.host.size() * sizeof(HvJob), hipMemcpyHostToDevice,
                         stream()));
}

HvBatchConfig hv_batch_tune(const HvBatch& B, const int* active, bool force) {
  static std::once_flag once;
  std::call_once(once, [] {
    const int bytes = (int)lds_budget();
    (void)hipFuncSetAttribute((const void*)k_hv_ell_batch<256>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell_batch<512>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute((const void*)k_hv_ell_batch<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  });
  TuneKey key{B.sig};
  auto it = tune_cache().find(key);
  if (it != tune_cache().end() && !force) return it->second;
  // candidates: block size x persistent grid (multiples of the CU count) or one tile per block
  const int cus = device_cu_count();
  std::vector<HvBatchConfig> cand;
  for (int bs : {256, 512, 1024})
    for (int mult : {0, 2, 4, 8}) cand.push_back({bs, mult ? (uint32_t)(cus * mult) : 0u, 0.0f, 0});
  if (B.tiled)
    for (int mult : {0, 4, 8}) cand.push_back({256, mult ? (uint32_t)(cus * mult) : 0u, 0.0f, 1});
  hipEvent_t e0, e1;
  QED_HIP(hipEventCreate(&e0));
  QED_HIP(hipEventCreate(&e1));
  HvBatchConfig best = cand[0];
  best.ms = 1e30f;
  for (auto& c : cand) {
    launch_batch(B, c, 0, active, stream());  // warm-up (y accumulates; callers pass scratch y)
    QED_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < 3; ++r) launch_batch(B, c, 0, active, stream());
    QED_HIP(hipEventRecord(e1, stream()));
    QED_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    QED_HIP(hipEventElapsedTime(&ms, e0, e1));
    c.ms = ms / 3.0f;
    if (c.ms < best.ms) best = c;
  }
  QED_HIP(hipEventDestroy(e0));
  QED_HIP(hipEventDestroy(e1));
  QED_HIP(hipGetLastError());
  tune_cache()[key] = best;
  return best;
}

void hv_batch(const HvBatch& B, const HvBatchConfig& cfg, int parity, const int* active) {
  launch_batch(B, cfg, parity, active, stream());
  QED_HIP(hipGetLastError());
}

}  // namespace qed

using namespace qed;

namespace {

// start-vector state shared by every sector, as the reference keeps one process-wide
// normal_distribution (src_ed/Hamiltonian/Lanczos.cpp:7)
// The stream is std::normal_distribution<double>(0, 1) (libstdc++: Marsaglia's polar
// method, the second deviate of each pair kept for the next call) over a fresh
// std::default_random_engine(seed) (minstd_rand0) through
// std::generate_canonical<double, 53> (two engine draws per uniform). It is restated here
// so that it can be generated in parallel: attempt t of the polar method uses engine
// draws 4t+1 .. 4t+4, reachable by a jump of the linear congruential engine, so blocks of
// attempts are counted and then filled independently. The values are identical to the
// sequential stream.
struct NormalState {
  bool saved_available = false;
  double saved = 0.0;
};
NormalState g_normal;
std::mutex g_normal_mutex;

constexpr uint64_t kLcgA = 16807, kLcgM = 2147483647;

uint64_t lcg_pow(uint64_t k) {  // a^k mod m
  uint64_t a = kLcgA, r = 1;
  while (k) {
    if (k & 1) r = r * a % kLcgM;
    a = a * a % kLcgM;
    k >>= 1;
  }
  return r;
}

// std::generate_canonical<double, 53> over minstd_rand0 (min 1, max m - 1): two draws
double canonical2(uint64_t x1, uint64_t x2) {
  const long double r = static_cast<long double>(kLcgM - 1) - static_cast<long double>(1) + 1.0L;
  double sum = 0.0, tmp = 1.0;
  sum += double(x1 - 1) * tmp;
  tmp *= r;
  sum += double(x2 - 1) * tmp;
  tmp *= r;
  double ret = sum / tmp;
  if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
  return ret;
}

// one polar-method attempt starting at engine state st (draws st*a .. st*a^4)
inline bool polar_attempt(uint64_t& st, double& y_out, double& x_out) {
  const uint64_t x1 = st * kLcgA % kLcgM, x2 = x1 * kLcgA % kLcgM, x3 = x2 * kLcgA % kLcgM,
                 x4 = x3 * kLcgA % kLcgM;
  st = x4;
  const double x = 2.0 * canonical2(x1, x2) - 1.0;
  const double y = 2.0 * canonical2(x3, x4) - 1.0;
  const double r2 = x * x + y * y;
  if (r2 > 1.0 || r2 == 0.0) return false;
  const double mult = std::sqrt(-2 * std::log(r2) / r2);
  x_out = x * mult;
  y_out = y * mult;
  return true;
}

// next n values of the stream (value = 0.0 + (deviate * 1.0 + 0.0), as drawn with mean
// 0 and standard deviation 1) for a fresh engine seeded with `seed`
void normal_stream(unsigned seed, uint64_t n, double* out, int threads) {
  uint64_t s0 = seed % kLcgM;
  if (s0 == 0) s0 = 1;
  auto emit = [](double v) { return 0.0 + (v * 1.0 + 0.0); };
  uint64_t pos = 0;
  if (n == 0) return;
  if (g_normal.saved_available) {
    out[pos++] = emit(g_normal.saved);
    g_normal.saved_available = false;
  }
  const uint64_t need = n - pos;
  if (need == 0) return;
  const uint64_t pairs = (need + 1) / 2;
  const uint64_t K = 1 << 15;  // attempts per block
  uint64_t nblk = (uint64_t)(pairs * 1.2732395447351628 * 1.02 / K) + 2;  // 4/pi acceptance
  std::vector<uint64_t> cnt;
  const int nt = std::max(1, threads);
  auto run = [&](uint64_t b0, uint64_t b1, auto&& body) {
    std::vector<std::thread> th;
    std::atomic<uint64_t> next{b0};
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&] {
        for (uint64_t b = next++; b < b1; b = next++) body(b);
      });
    for (auto& t : th) t.join();
  };
  // pass 1: accepted attempts per block (extend until there are enough)
  uint64_t have = 0, counted = 0;
  while (true) {
    cnt.resize(nblk, 0);
    run(counted, nblk, [&](uint64_t b) {
      uint64_t st = s0 * lcg_pow(4 * K * b) % kLcgM;
      uint64_t c = 0;
      double y, x;
      for (uint64_t t = 0; t < K; ++t) c += polar_attempt(st, y, x);
      cnt[b] = c;
    });
    for (uint64_t b = counted; b < nblk; ++b) have += cnt[b];
    counted = nblk;
    if (have >= pairs) break;
    nblk += nblk / 8 + 2;
  }
  // pass 2: block b fills pairs [first[b], first[b] + cnt[b])
  std::vector<uint64_t> first(nblk + 1, 0);
  for (uint64_t b = 0; b < nblk; ++b) first[b + 1] = first[b] + cnt[b];
  uint64_t lastblk = 0;
  while (first[lastblk + 1] < pairs) ++lastblk;
  double saved = 0.0;
  run(0, lastblk + 1, [&](uint64_t b) {
    uint64_t st = s0 * lcg_pow(4 * K * b) % kLcgM;
    uint64_t p = first[b];
    double y, x;
    for (uint64_t t = 0; t < K && p < pairs; ++t) {
      if (!polar_attempt(st, y, x)) continue;
      const uint64_t i = 2 * p;
      out[pos + i] = emit(y);
      if (i + 1 < need) out[pos + i + 1] = emit(x);
      else saved = x;  // odd count: the pair's second deviate stays for the next call
      ++p;
    }
  });
  if (need % 2 == 1) {
    g_normal.saved = saved;
    g_normal.saved_available = true;
  }
}

}  // namespace

extern "C" {

int qed_halfbasis(int L, int N, uint32_t* out, uint32_t* dim) {
  return guard([&] {
    if (L < 0 || L > 32) fail("number of orbitals out of range");
    const auto& B = binom_host();
    uint32_t n = (N < 0 || N > L) ? 0 : B[L * 33 + N];
    *dim = n;
    if (!out || n == 0) return;
    DevBuf d(n * sizeof(uint32_t));
    k_unrank<<<(n + 255) / 256, 256, 0, stream()>>>(d.as<uint32_t>(), n, L, N, binom_dev());
    QED_HIP(hipMemcpyAsync(out, d.p, n * 4, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_halfbasis_rank(int L, int N, const uint32_t* states, uint64_t n, uint32_t* ranks) {
  return guard([&] {
    if (L < 0 || L > 32) fail("number of orbitals out of range");
    const auto& B = binom_host();
    uint32_t dim = (N < 0 || N > L) ? 0 : B[L * 33 + N];
    if (n == 0) return;
    DevBuf s(n * 4), r(n * 4);
    QED_HIP(hipMemcpyAsync(s.p, states, n * 4, hipMemcpyHostToDevice, stream()));
    uint32_t full = L >= 32 ? 0xffffffffu : ((1u << L) - 1u);
    k_rank<<<(uint32_t)((n + 255) / 256), 256, 0, stream()>>>(s.as<uint32_t>(), r.as<uint32_t>(), n,
                                                              N, dim, full, binom_dev());
    QED_HIP(hipMemcpyAsync(ranks, r.p, n * 4, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_sector_create(int L, int n_up, int n_dn, int n_ops, const int* types, const double* coefs,
                      const int* nels, const int* er, const int* ec, const double* ev,
                      qed_sector** out) {
  return guard([&] {
    *out = nullptr;
    if (L <= 0 || L > 32) fail("number of orbitals out of range");
    auto s = std::make_unique<qed_sector>();
    s->L = L;
    s->nup = n_up;
    s->ndn = n_dn;
    // per-spin bond lists from the Hermitian-completed element lists
    std::vector<Bond> bu, bd;
    std::vector<double> vu, vd;
    std::vector<int> dsu, dsd;
    std::vector<double> dvu, dvd;
    int off = 0;
    for (int k = 0; k < n_ops; ++k) {
      const double z = coefs[k];
      for (int i = 0; i < nels[k]; ++i) {
        int r = er[off + i], c = ec[off + i];
        double v = ev[off + i];
        if (r < 0 || c < 0 || r >= 2 * L || c >= 2 * L) fail("operator element index out of range");
        if (types[k] == 0) {
          if (r > c) continue;  // only r <= c is read (one_body_operator ctor)
          if ((r < L) != (c < L)) fail("spin-flip one-body terms are not factorizable");
          if (std::fabs(v) < 1.0e-12) continue;
          bool dn = r >= L;
          int rs = dn ? r - L : r, cs = dn ? c - L : c;
          if (r == c) {
            (dn ? dsd : dsu).push_back(rs);
            (dn ? dvd : dvu).push_back(v * z);
            auto& ds = dn ? s->dsite_dn : s->dsite_up;
            ds.resize(L, 0.0);
            ds[rs] += v * z;
          } else {
            auto& B = dn ? bd : bu;
            auto& V = dn ? vd : vu;
            B.push_back({cs, rs});
            V.push_back(v * z);
            B.push_back({rs, cs});
            V.push_back(v * z);
            (dn ? s->pd_i : s->pu_i).push_back(rs);
            (dn ? s->pd_j : s->pu_j).push_back(cs);
            (dn ? s->pd_t : s->pu_t).push_back(v * z);
          }
        } else if (types[k] == 1) {
          int a = std::min(r, c), b = std::max(r, c);
          uint64_t m = (a >= L ? (uint64_t(1) << (a - L + 32)) : (uint64_t(1) << a)) |
                       (b >= L ? (uint64_t(1) << (b - L + 32)) : (uint64_t(1) << b));
          s->int_mask.push_back(m);
          s->int_val.push_back(v * z);
        } else {
          fail("unsupported operator type in the factorized format");
        }
      }
      off += nels[k];
    }
    // uniform Hubbard detection: elements (i, i+L) with one common value, each site once
    s->hubbard_uniform = true;
    s->hub_sites = 0;
    s->hub_U = s->int_val.empty() ? 0.0 : s->int_val[0];
    for (size_t q = 0; q < s->int_mask.size(); ++q) {
      uint64_t m = s->int_mask[q];
      uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
      if (lo != hi || __builtin_popcount(lo) != 1 || (s->hub_sites & lo) || s->int_val[q] != s->hub_U) {
        s->hubbard_uniform = false;
        break;
      }
      s->hub_sites |= lo;
    }
    if (!s->hubbard_uniform) s->hub_sites = 0;
    auto hu = half_struct(L, n_up, bu, true);
    auto hd = half_struct(L, n_dn, bd, false);
    s->keep_up = hu;
    s->keep_dn = hd;
    s->du = hu->dim;
    s->dd = hd->dim;
    s->D = (uint64_t)s->du * s->dd;
    s->bup = DevBuf::view(hu->bin.p, hu->bin.bytes);
    s->bdn = DevBuf::view(hd->bin.p, hd->bin.bytes);
    s->up_ptr = DevBuf::view(hu->ptr.p, hu->ptr.bytes);
    s->dn_ptr = DevBuf::view(hd->ptr.p, hd->ptr.bytes);
    s->up_ent = DevBuf::view(hu->ent.p, hu->ent.bytes);
    s->dn_ent = DevBuf::view(hd->ent.p, hd->ent.bytes);
    s->up_ell = DevBuf::view(hu->ell.p, hu->ell.bytes);
    s->dn_ell = DevBuf::view(hd->ell.p, hd->ell.bytes);
    s->up_nnz = hu->nnz;
    s->dn_nnz = hd->nnz;
    s->up_nvals = hu->nvals;
    s->dn_nvals = hd->nvals;
    s->up_maxdeg = hu->maxdeg;
    s->dn_maxdeg = hd->maxdeg;
    s->up_w = hu->w;
    s->dn_w = hd->w;
    // this sector's amplitudes in one allocation: signed hop values (with a trailing
    // zero for the ELL padding), one-body diagonals, interaction values and masks
    const size_t nvu = hu->nvals, nvd = hd->nvals;
    const size_t nint = std::max<size_t>(1, s->int_val.size());
    std::vector<double> hbuf(nvu + nvd + s->du + s->dd + nint, 0.0);
    double* vup = hbuf.data();
    double* vdn = vup + nvu;
    double* dgu = vdn + nvd;
    double* dgd = dgu + s->du;
    double* iv = dgd + s->dd;
    for (size_t b = 0; b < bu.size(); ++b) {
      vup[2 * b] = vu[b];
      vup[2 * b + 1] = -vu[b];
    }
    for (size_t b = 0; b < bd.size(); ++b) {
      vdn[2 * b] = vd[b];
      vdn[2 * b + 1] = -vd[b];
    }
    half_diag(*hu, dsu, dvu, dgu);
    half_diag(*hd, dsd, dvd, dgd);
    for (size_t q = 0; q < s->int_val.size(); ++q) iv[q] = s->int_val[q];
    std::vector<uint64_t> masks(nint, 0);
    for (size_t q = 0; q < s->int_mask.size(); ++q) masks[q] = s->int_mask[q];
    const size_t nd = hbuf.size();
    s->valbuf.alloc((nd + nint) * 8);
    QED_HIP(hipMemcpyAsync(s->valbuf.p, hbuf.data(), nd * 8, hipMemcpyHostToDevice, stream()));
    QED_HIP(hipMemcpyAsync(s->valbuf.as<double>() + nd, masks.data(), nint * 8, hipMemcpyHostToDevice, stream()));
    double* base = s->valbuf.as<double>();
    s->up_vals = DevBuf::view(base, nvu * 8);
    s->dn_vals = DevBuf::view(base + nvu, nvd * 8);
    s->diag_up = DevBuf::view(base + nvu + nvd, (size_t)s->du * 8);
    s->diag_dn = DevBuf::view(base + nvu + nvd + s->du, (size_t)s->dd * 8);
    s->int_val_d = DevBuf::view(base + nvu + nvd + s->du + s->dd, nint * 8);
    s->int_mask_d = DevBuf::view(base + nd, nint * 8);
    QED_HIP(hipStreamSynchronize(stream()));
    if (s->du % 2 == 0 && s->du >= 4096 && ((size_t)s->du + s->up_nvals + 1) * sizeof(double) <= lds_budget())
      ensure_row_order(*s);
    *out = s.release();
  });
}

int qed_sector_destroy(qed_sector* s) {
  return guard([&] {
    QED_HIP(hipStreamSynchronize(stream()));
    delete s;
  });
}

int qed_sector_dims(const qed_sector* s, uint64_t* dim, uint32_t* du, uint32_t* dd) {
  return guard([&] {
    if (dim) *dim = s->D;
    if (du) *du = s->du;
    if (dd) *dd = s->dd;
  });
}

int qed_hv(qed_sector* s, const double* x, double* y) {
  return guard([&] { hv(*s, x, y); });
}

int qed_hv_variant(qed_sector* s, int variant, const double* x, double* y) {
  return guard([&] { hv(*s, x, y, variant); });
}

// Times `reps` launches of H.v (after `warmup`) with HIP events on the library stream;
// *ms = average milliseconds per launch.
int qed_hv_timed(qed_sector* s, int variant, const double* x, double* y, int warmup, int reps,
                 double* ms) {
  return guard([&] {
    for (int i = 0; i < warmup; ++i) hv(*s, x, y, variant);
    hipEvent_t e0, e1;
    QED_HIP(hipEventCreate(&e0));
    QED_HIP(hipEventCreate(&e1));
    QED_HIP(hipEventRecord(e0, stream()));
    for (int i = 0; i < reps; ++i) hv(*s, x, y, variant);
    QED_HIP(hipEventRecord(e1, stream()));
    QED_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QED_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / (reps > 0 ? reps : 1);
  });
}

int qed_hv_batch(int n, qed_sector* const* s, const double* const* x, double* const* y, int mode, int reps,
                 double* ms) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (n <= 0) return;
    std::vector<const qed_sector*> S(s, s + n);
    HvBatch B;
    hv_batch_prepare(B, S);
    if (mode == 1 && !B.tiled) fail("batched H.v: the tiled form needs rows that fit in LDS tiles");
    for (int j = 0; j < n; ++j) {
      B.host[j].x[0] = B.host[j].x[1] = x[j];
      B.host[j].y[0] = B.host[j].y[1] = y[j];
    }
    hv_batch_upload(B);
    std::vector<int> on(n, 1);
    DevBuf act = upload(on);
    HvBatchConfig cfg;
    cfg.block = mode == 1 ? 256 : 512;
    cfg.mode = mode == 1 ? 1 : 0;
    if (reps < 1) reps = 1;
    hipEvent_t e0, e1;
    QED_HIP(hipEventCreate(&e0));
    QED_HIP(hipEventCreate(&e1));
    QED_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r) hv_batch(B, cfg, 0, act.as<int>());
    QED_HIP(hipEventRecord(e1, stream()));
    QED_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QED_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = (double)t / reps;
  });
}

int qed_hv_host(qed_sector* s, const double* x, double* y) {
  return guard([&] {
    size_t bytes = s->D * sizeof(double);
    DevBuf dx(bytes), dy(bytes);
    QED_HIP(hipMemcpyAsync(dx.p, x, bytes, hipMemcpyHostToDevice, stream()));
    QED_HIP(hipMemcpyAsync(dy.p, y, bytes, hipMemcpyHostToDevice, stream()));
    hv(*s, dx.as<double>(), dy.as<double>());
    QED_HIP(hipMemcpyAsync(y, dy.p, bytes, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

int qed_diag(qed_sector* s, double* d) {
  return guard([&] {
    if (s->D == 0) return;
    if (!s->factorized()) {
      sector_diagonal(*s, d);
      return;
    }
    k_diag_interaction<<<(uint32_t)((s->D + 255) / 256), 256, 0, stream()>>>(view(*s), d, s->D);
    QED_HIP(hipGetLastError());
  });
}

int qed_start_vector(qed_sector* s, int seed, double* x_dev) {
  return guard([&] {
    std::vector<double> x(s->D);
    {
      std::lock_guard<std::mutex> lk(g_normal_mutex);
      const unsigned hc = std::thread::hardware_concurrency();
      normal_stream((unsigned)seed, s->D, x.data(), s->D < (1u << 16) ? 1 : (int)std::min(16u, std::max(1u, hc)));
    }
    double z = 0.0;
    for (double v : x) z += v * v;
    z = std::sqrt(z);
    if (z >= 1.0e-12)
      for (double& v : x) v *= 1.0 / z;
    QED_HIP(hipMemcpyAsync(x_dev, x.data(), s->D * 8, hipMemcpyHostToDevice, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

// n values of the start-vector stream for `seed` after a reset, computed with the
// standard library's distribution (which = 0) or the parallel restatement (which = 1,
// `threads` workers); the state is reset before and after.
int qed_util_normal_stream(int seed, uint64_t n, int which, int threads, double* out) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_normal_mutex);
    g_normal = NormalState();
    if (which == 0) {
      std::normal_distribution<double> nd(0.0, 1.0);
      std::default_random_engine gen((unsigned)seed);
      // two calls of n/2 values each, so that a carried-over deviate is exercised
      const uint64_t h = n / 2;
      for (uint64_t i = 0; i < h; ++i) out[i] = 0.0 + nd(gen);
      std::default_random_engine gen2((unsigned)seed + 1);
      for (uint64_t i = h; i < n; ++i) out[i] = 0.0 + nd(gen2);
    } else {
      const uint64_t h = n / 2;
      normal_stream((unsigned)seed, h, out, threads);
      normal_stream((unsigned)seed + 1, n - h, out + h, threads);
    }
    g_normal = NormalState();
  });
}

int qed_reset_start_vector_state(void) {
  return guard([] {
    std::lock_guard<std::mutex> lk(g_normal_mutex);
    g_normal = NormalState();
  });
}

int qed_apply_c(int L, int n_up, int n_dn, int a, int pm, double z, const double* x, double* y) {
  return guard([&] {
    if (a < 0 || a >= 2 * L) fail("orbital index out of range");
    if (pm != 1 && pm != -1) fail("pm must be +1 or -1");
    const auto& B = binom_host();
    bool down = a >= L;
    int site = down ? a - L : a;
    int tup = n_up + (down ? 0 : pm), tdn = n_dn + (down ? pm : 0);
    if (tup < 0 || tup > L || tdn < 0 || tdn > L || n_up < 0 || n_up > L || n_dn < 0 || n_dn > L)
      return;
    uint32_t tdu = B[L * 33 + tup], tdd = B[L * 33 + tdn], sdu = B[L * 33 + n_up];
    uint64_t DT = (uint64_t)tdu * tdd;
    if (DT == 0) return;
    DevBuf tu(tdu * 4), td(tdd * 4);
    k_unrank<<<(tdu + 255) / 256, 256, 0, stream()>>>(tu.as<uint32_t>(), tdu, L, tup, binom_dev());
    k_unrank<<<(tdd + 255) / 256, 256, 0, stream()>>>(td.as<uint32_t>(), tdd, L, tdn, binom_dev());
    k_apply_c<<<(uint32_t)((DT + 255) / 256), 256, 0, stream()>>>(
        tu.as<uint32_t>(), td.as<uint32_t>(), tdu, DT, L, n_up, n_dn, site, down ? 1 : 0, pm, z,
        x, sdu, y, binom_dev());
    QED_HIP(hipGetLastError());
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"
