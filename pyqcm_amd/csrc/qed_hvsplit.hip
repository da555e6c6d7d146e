// Split-halves H.v: y += H x in two streaming passes over the sector vector.
//
// The vector is the d_dn x d_up row-major matrix X[i_dn][i_up] of the reference's
// factorized order (src_ed/ED_basis.cpp:142-165). The products of
// HS_factorized_operator::multiply_add (src_ed/Operators/HS_factorized_operator.hpp:56-84)
// and of the density-density diagonal (src_ed/Operators/HS_interaction_operator.hpp) are
// regrouped by where their operands live:
//
//   pass 1 (row pass): y[r, :] += (diagonal + up-spin hops) x[r, :]. One workgroup holds
//     the row x[r, :] in LDS; the next row is loaded into registers while the current one
//     is used. Each up state is (a, b) with a the orbitals below the split and b those
//     above (qed_split.hpp); a wave carries up to 64 low strings a of one class (one per
//     lane, with its hops inside the low part held in registers) and runs through the
//     high parts b of that class, whose hops inside the high part and across the split
//     are uniform over the wave (scalar loads of a 256-byte record per b).
//   pass 2 (tile pass): y[r, c] += (down-spin hops) x[r', c] for tiles of one block of
//     rows (one high part b of the down spin: the rows reached by hops inside the low
//     part) by W columns. The tile is staged in LDS; the hops leaving the block read the
//     neighbour rows' W-column segments from global memory, which the column-panel order
//     of the workgroups keeps in the XCD's L2. Every hop list is uniform over a wave.
//
// Each pass reads x once and y once and writes y once from HBM (48 bytes per element for
// the pair); the re-reads go to LDS (row pass, in-block tile hops) and L2 (neighbour rows).
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "qed_sector.hpp"
#include "qed_split.hpp"

namespace qed {

using split::LRec2;
using split::SRec;
using split::Task;

namespace {

// read-only tables seen through the constant address space (scalar loads when uniform)
using CU32 = const __attribute__((address_space(4))) uint32_t;
using CF64 = const __attribute__((address_space(4))) double;
using CSRec = const __attribute__((address_space(4))) SRec;
using CTask = const __attribute__((address_space(4))) Task;

constexpr int kRowBS = 768;  // threads of a row-pass workgroup (three waves per SIMD)
constexpr int kRowWaves = kRowBS / 64;
constexpr int kDownBS = 256;

struct RowArgs {
  const SRec* rec;         // per unit and step, in sequence order
  const LRec2* lrec;       // unit u, lane l at u * 64 + l
  const uint32_t* rec0;    // first record of each unit
  const Task* tasks;       // tasks of the waves, concatenated
  const uint32_t* task0;   // kRowWaves + 1 offsets
  const uint32_t* bdn;     // down half basis
  const double* diag_dn;   // one-body diagonal of the down half states
  uint32_t du, dd;
  int k;                   // split orbital
  int nx;                  // low orbitals on bonds across the split
  double U;                // uniform Hubbard interaction on hub_sites
  uint32_t hub_sites;
};

// Row pass: z[r, :] = (diagonal + up-spin hops) x[r, :]. z is written, not read, so that
// no wave waits on a load of its output. KL: hop slots inside the low part held per lane
// (padding slots read another lane's address: LDS broadcasts); NXT, NPQ: low orbitals on
// bonds across the split and bonds per such orbital (cross slots qi * NPQ + m); PF: doubles
// of the next row held per thread (du <= PF * kRowBS). Every step of a wave reads its
// record (block offsets and amplitudes of one high part b) with scalar loads at fixed
// offsets; the gathers of a hop group are all issued before they are summed.
template <int KL, int NXT, int NPQ, int PF>
__global__ __launch_bounds__(kRowBS) void k_split_rows(RowArgs A, const double* __restrict__ x,
                                                       double* __restrict__ z) {
  extern __shared__ double xs[];
  const char* xsb = reinterpret_cast<const char*>(xs);
  const CSRec* __restrict__ rec = (const CSRec*)(A.rec);
  const CU32* __restrict__ rec0 = (const CU32*)(A.rec0);
  const CTask* __restrict__ tasks = (const CTask*)(A.tasks);
  const CU32* __restrict__ task0 = (const CU32*)(A.task0);
  const CU32* __restrict__ bdn = (const CU32*)(A.bdn);
  const CF64* __restrict__ diag_dn = (const CF64*)(A.diag_dn);
  const uint32_t du = A.du;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint32_t lomask = (1u << A.k) - 1u;
  const uint32_t task_a = task0[wave], task_b = task0[wave + 1];

  double pf[PF];
  uint32_t r = blockIdx.x;
  if (r < A.dd) {
    const double* xr = x + (uint64_t)r * du;
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const uint32_t c = threadIdx.x + q * kRowBS;
      pf[q] = c < du ? __builtin_nontemporal_load(xr + c) : 0.0;
    }
  }
  // lane record of the unit in registers (reloaded when the wave moves to another unit)
  uint32_t cur_unit = 0xffffffffu;
  uint32_t ra8 = 0, valid = 0, abits = 0;
  double e_lo = 0.0;
  uint32_t xra8[NXT], xneg[NXT];
  uint32_t xocc = 0;  // bit q: the low string holds cross orbital q
  uint32_t lra8[KL];
  double lamp[KL];

  for (; r < A.dd; r += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const uint32_t c = threadIdx.x + q * kRowBS;
      if (c < du) xs[c] = pf[q];
    }
    __syncthreads();
    const uint32_t rn = r + gridDim.x;
    if (rn < A.dd) {
      const double* xr = x + (uint64_t)rn * du;
#pragma unroll
      for (int q = 0; q < PF; ++q) {
        const uint32_t c = threadIdx.x + q * kRowBS;
        pf[q] = c < du ? __builtin_nontemporal_load(xr + c) : 0.0;
      }
    }
    const uint32_t sd = bdn[r];
    const double edn = diag_dn[r];
    const uint32_t sdh = sd & A.hub_sites;
    const uint32_t sdh_lo = sdh & lomask, sdh_hi = sdh >> A.k;
    double* __restrict__ zr = z + (uint64_t)r * du;

    for (uint32_t ti = task_a; ti < task_b; ++ti) {
      const uint32_t t_unit = tasks[ti].unit, t_t0 = tasks[ti].t0, t_t1 = tasks[ti].t1;
      if (t_unit != cur_unit) {
        cur_unit = t_unit;
        const LRec2* L = A.lrec + (uint64_t)t_unit * 64 + lane;
        ra8 = L->ra8;
        valid = L->valid;
        abits = L->abits;
        e_lo = L->e;
        const uint32_t xsa = L->xsa;
        xocc = xsa >> 16;
#pragma unroll
        for (int q = 0; q < NXT; ++q) {
          xra8[q] = L->xra8[q];
          xneg[q] = (xsa >> q & 1u) << 31;
        }
#pragma unroll
        for (int s = 0; s < KL; ++s) {
          lra8[s] = L->lra8[s];
          lamp[s] = L->lamp[s];
        }
      }
      const double lane_diag = e_lo + edn + A.U * (double)__popc(abits & sdh_lo);
      const uint32_t r0 = rec0[t_unit];
      for (uint32_t t = t_t0; t < t_t1; ++t) {
        const CSRec* R = rec + r0 + t;
        const uint32_t off8 = R->off8;
        const uint32_t col8 = off8 + ra8;
        const double own = *reinterpret_cast<const double*>(xsb + col8);
        const double bdiag = R->e + A.U * (double)__popc(R->bits & sdh_hi);
        double v[KL];
#pragma unroll
        for (int s = 0; s < KL; ++s) v[s] = *reinterpret_cast<const double*>(xsb + off8 + lra8[s]);
        double acc0 = (lane_diag + bdiag) * own, acc1 = 0.0;
#pragma unroll
        for (int s = 0; s < KL; s += 2) {
          acc0 += lamp[s] * v[s];
          if (s + 1 < KL) acc1 += lamp[s + 1] * v[s + 1];
        }
        // (the scheduler may not lift the next group's loads above this point: it keeps
        // the values in flight to one group)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int g = 0; g < split::kRecHigh; g += 6) {
          double h[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) h[q] = *reinterpret_cast<const double*>(xsb + R->hoff8[g + q] + ra8);
#pragma unroll
          for (int q = 0; q < 6; q += 2) {
            acc0 += R->hamp[g + q] * h[q];
            acc1 += R->hamp[g + q + 1] * h[q + 1];
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        const uint32_t xbj = R->xbj;
#pragma unroll
        for (int q = 0; q < NXT; ++q) {
          if (q >= A.nx) break;
          const uint32_t ai = xocc >> q & 1u;
          uint32_t xo[NPQ];
          double xc[NPQ], v[NPQ];
#pragma unroll
          for (int m = 0; m < NPQ; ++m) {
            xo[m] = R->xoff8[q * NPQ + m];
            xc[m] = R->xcoef[q * NPQ + m];
          }
          // the hop needs the cross orbital of a and that of b in opposite states; the
          // other lanes keep 0 (only the load sits under the lane mask)
#pragma unroll
          for (int m = 0; m < NPQ; ++m) {
            v[m] = 0.0;
            if (ai != (xbj >> (q * NPQ + m) & 1u)) v[m] = *reinterpret_cast<const double*>(xsb + xo[m] + xra8[q]);
          }
          double ax = 0.0;
#pragma unroll
          for (int m = 0; m < NPQ; ++m) ax += xc[m] * v[m];
          acc0 += __longlong_as_double(__double_as_longlong(ax) ^ ((uint64_t)xneg[q] << 32));
          __builtin_amdgcn_sched_barrier(0);
        }
        if (valid) __builtin_nontemporal_store(acc0 + acc1, reinterpret_cast<double*>(reinterpret_cast<char*>(zr) + col8));
      }
    }
  }
}

struct DownArgs {
  const uint32_t* ptr;     // per row: first entry (dd + 1)
  const uint32_t* nin;     // per row: entries inside the row's block (listed first)
  const uint32_t* off;     // entry: LDS byte offset of a tile row, or byte offset of an x row
  const double* amp;       // entry: signed amplitude
  const uint32_t* blk_off; // blocks (rows of one high part)
  const uint32_t* blk_len;
  uint32_t du, nblk, npanels;
};

constexpr int kDG = 8;  // hop values gathered before they are summed

// Tile pass: y[r, c] += z[r, c] + (down-spin hops) x. Workgroup g runs on XCD g % 8 and
// takes, on that XCD, the W-column panels congruent to it in turn (all blocks of a panel
// follow one another, so that the panel of x, dd * W * 8 bytes, stays in that XCD's L2).
// A wave sums one row of its block at a time over W lanes; the row's hop list (byte
// offsets and amplitudes) is read with scalar loads, eight entries per group.
template <int W>
__global__ __launch_bounds__(kDownBS) void k_split_down(DownArgs A, const double* __restrict__ x,
                                                        const double* __restrict__ z, double* __restrict__ y) {
  extern __shared__ double tile[];
  const CU32* __restrict__ ptr = (const CU32*)(A.ptr);
  const CU32* __restrict__ nin = (const CU32*)(A.nin);
  const CU32* __restrict__ eoff = (const CU32*)(A.off);
  const CF64* __restrict__ eamp = (const CF64*)(A.amp);
  const CU32* __restrict__ blk_off = (const CU32*)(A.blk_off);
  const CU32* __restrict__ blk_len = (const CU32*)(A.blk_len);
  const uint32_t g = blockIdx.x;
  const uint32_t xcd = g & 7u, j = g >> 3;
  const uint32_t panel = (j / A.nblk) * 8u + xcd;
  if (panel >= A.npanels) return;
  const uint32_t bi = j % A.nblk;
  const uint32_t du = A.du;
  const uint32_t c0 = panel * W;
  const uint32_t off = blk_off[bi], m = blk_len[bi];
  for (uint32_t e = threadIdx.x; e < m * W; e += kDownBS) {
    const uint32_t rr = e / W, c = e % W;
    tile[e] = (c0 + c < du) ? x[(uint64_t)(off + rr) * du + c0 + c] : 0.0;
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (lane >= (uint32_t)W || c0 + lane >= du) return;
  const uint32_t c = c0 + lane;
  const uint32_t lane8 = lane * 8u;
  const char* xc = reinterpret_cast<const char*>(x) + (uint64_t)c * 8u;
  const char* tb = reinterpret_cast<const char*>(tile) + lane8;
  for (uint32_t ra = wave; ra < m; ra += kDownBS / 64) {
    const uint32_t r = off + ra;
    const uint64_t rc = (uint64_t)r * du + c;
    const double yz = __builtin_nontemporal_load(y + rc) + __builtin_nontemporal_load(z + rc);
    const uint32_t p0 = ptr[r], p1 = ptr[r + 1], pin = p0 + nin[r];
    double acc0 = 0.0, acc1 = 0.0;
    // entry groups are read whole (the tables carry kDG padding entries at their end) and
    // every gather is issued: past the row's count it reads a harmless address (tile row 0,
    // or this row of x) and its amplitude is 0
    for (uint32_t p = p0; p < pin; p += kDG) {
      uint32_t e[kDG];
      double a[kDG], v[kDG];
      const CU32* ep = eoff + p;
      const CF64* ap = eamp + p;
#pragma unroll
      for (int q = 0; q < kDG; ++q) {
        e[q] = ep[q];
        a[q] = ap[q];
      }
#pragma unroll
      for (int q = 0; q < kDG; ++q) v[q] = *reinterpret_cast<const double*>(tb + (p + q < pin ? e[q] : 0u));
#pragma unroll
      for (int q = 0; q < kDG; q += 2) {
        acc0 += (p + q < pin ? a[q] : 0.0) * v[q];
        acc1 += (p + q + 1 < pin ? a[q + 1] : 0.0) * v[q + 1];
      }
    }
    const uint32_t self8 = r * 8u;
    for (uint32_t p = pin; p < p1; p += kDG) {
      uint32_t e[kDG];
      double a[kDG], v[kDG];
      const CU32* ep = eoff + p;
      const CF64* ap = eamp + p;
#pragma unroll
      for (int q = 0; q < kDG; ++q) {
        e[q] = ep[q];
        a[q] = ap[q];
      }
#pragma unroll
      for (int q = 0; q < kDG; ++q)
        v[q] = *reinterpret_cast<const double*>(xc + (uint64_t)(p + q < p1 ? e[q] : self8) * du);
#pragma unroll
      for (int q = 0; q < kDG; q += 2) {
        acc0 += (p + q < p1 ? a[q] : 0.0) * v[q];
        acc1 += (p + q + 1 < p1 ? a[q + 1] : 0.0) * v[q + 1];
      }
    }
    __builtin_nontemporal_store(yz + acc0 + acc1, y + rc);
  }
}

// Device tables of one sector's split plan.
struct SplitTables {
  bool ok = false;
  std::string why;
  int k = 0, kl = 0, nx = 0, npq = 0, pf = 0;
  DevBuf rec, lrec, rec0, tasks, task0;
  DevBuf dptr, dnin, didx, damp, blk_off, blk_len;
  DevBuf z;  // output of the row pass, read by the tile pass
  uint32_t nblk = 0, max_block = 0;
  static constexpr int W = 32;  // tile width (columns)
};

std::vector<split::Pair> merge_pairs(const std::vector<int>& pi, const std::vector<int>& pj,
                                     const std::vector<double>& pt) {
  std::map<std::pair<int, int>, double> m;
  std::vector<std::pair<int, int>> order;
  for (size_t q = 0; q < pi.size(); ++q) {
    auto key = std::make_pair(std::min(pi[q], pj[q]), std::max(pi[q], pj[q]));
    if (!m.count(key)) order.push_back(key);
    m[key] += pt[q];
  }
  std::vector<split::Pair> out;
  for (auto& k : order) out.push_back({k.first, k.second, m[k]});
  return out;
}

std::shared_ptr<SplitTables> build_split(const qed_sector& s) {
  auto T = std::make_shared<SplitTables>();
  auto no = [&](const std::string& m) {
    T->why = m;
    return T;
  };
  if (!s.hubbard_uniform) return no("interaction other than a uniform Hubbard term");
  if ((size_t)s.du * 8 > 160 * 1024 - 4096) return no("up row does not fit in LDS");
  if (s.du < 64 || s.dd < 64) return no("sector too small");
  const int L = s.L;
  if (L > 32 || L < 2) return no("orbital count");
  const auto pu = merge_pairs(s.pu_i, s.pu_j, s.pu_t);
  const auto pd = merge_pairs(s.pd_i, s.pd_j, s.pd_t);
  std::string why;
  const int k = split::choose_k(L, s.nup, pu, s.dsite_up, kRowWaves, &why);
  if (k < 0) return no(why.empty() ? "no split of the up orbitals suits the row pass" : why);
  const int kd = L / 2;
  const split::Spin Su = split::make_spin(L, s.nup, k, pu, s.dsite_up);
  const split::Spin Sd = split::make_spin(L, s.ndn, kd, pd, s.dsite_dn);
  if (Su.dim != s.du || Sd.dim != s.dd) return no("split dimensions disagree with the sector");
  if (!split::fits(Sd, &why)) return no(why);
  const split::RowPlan P = split::row_plan(Su, kRowWaves);
  T->k = k;
  T->kl = P.kl <= 8 ? 8 : P.kl <= 14 ? 14 : 16;
  T->nx = (int)Su.lx.size();
  T->npq = P.npq;
  T->pf = (int)((s.du + kRowBS - 1) / kRowBS);
  if (T->pf > 22) return no("up row too long");
  if (T->nx > 8 || P.npq > 4 || T->nx * std::max(1, P.npq) > split::kRecCross) return no("too many bonds across the split");
  const int npq_t = T->nx <= 4 ? 4 : T->nx <= 5 ? 3 : 2;  // cross slot layout of the kernel instances
  if (P.npq > npq_t) return no("too many bonds across the split per orbital");
  const split::XSlots X = split::x_slots(Su, npq_t);
  // tables of the row pass: per-step records in each unit's sequence order, lane records
  std::vector<SRec> recs;
  std::vector<uint32_t> rec0;
  std::vector<LRec2> lr(P.units.size() * 64);
  for (size_t u = 0; u < P.units.size(); ++u) {
    const split::Unit& U = P.units[u];
    rec0.push_back((uint32_t)recs.size());
    for (uint32_t bq : U.bseq) recs.push_back(split::s_record(Su, X, bq));
    std::vector<int> cnt(64, 0);
    for (size_t l = 0; l < 64; ++l) {
      if (l < U.a.size()) {
        lr[u * 64 + l] = split::l_record2(Su, U.pa, U.a[l]);
        cnt[l] = split::count_low(Su, U.a[l]);
      } else {
        LRec2 z{};
        lr[u * 64 + l] = z;
      }
    }
    // padding slots read the address of a lane that uses the slot (an LDS broadcast)
    for (int sl = 0; sl < split::kLaneLow; ++sl) {
      uint32_t ref = lr[u * 64].ra8;
      for (size_t l = 0; l < U.a.size(); ++l)
        if (cnt[l] > sl) {
          ref = lr[u * 64 + l].lra8[sl];
          break;
        }
      for (size_t l = 0; l < 64; ++l)
        if (cnt[l] <= sl) lr[u * 64 + l].lra8[sl] = ref;
    }
  }
  T->rec = upload(recs);
  T->rec0 = upload(rec0);
  T->lrec = upload(lr);
  std::vector<Task> tasks;
  std::vector<uint32_t> task0{0};
  for (const auto& wt : P.wave_tasks) {
    tasks.insert(tasks.end(), wt.begin(), wt.end());
    task0.push_back((uint32_t)tasks.size());
  }
  if (tasks.empty()) tasks.push_back({0, 0, 0, 0});
  T->tasks = upload(tasks);
  T->task0 = upload(task0);
  // tables of the tile pass
  const split::DownPlan D = split::down_plan(Sd);
  T->dptr = upload(D.ptr);
  T->dnin = upload(D.nin);
  // entries inside the block: LDS byte offset of the tile row (W doubles per row); the
  // others: the x row index (the kernel scales it by the row length in bytes)
  std::vector<uint32_t> eo(D.idx.size());
  for (uint32_t r = 0; r < Sd.dim; ++r)
    for (uint32_t p = D.ptr[r]; p < D.ptr[r + 1]; ++p)
      eo[p] = p < D.ptr[r] + D.nin[r] ? D.idx[p] * (uint32_t)T->W * 8u : D.idx[p] * 8u;
  eo.resize(eo.size() + kDG, 0u);  // whole-group reads past the last row
  T->didx = upload(eo);
  std::vector<double> ea(D.amp);
  ea.resize(ea.size() + kDG, 0.0);
  T->damp = upload(ea);
  T->blk_off = upload(D.blk_off);
  T->blk_len = upload(D.blk_len);
  T->nblk = (uint32_t)D.blk_off.size();
  T->max_block = D.max_block;
  if ((size_t)D.max_block * T->W * 8 > 64 * 1024) return no("down block too long for the tile");
  for (uint32_t r = 0; r < Sd.dim; ++r)
    if (D.ptr[r + 1] - D.ptr[r] > 64) return no("more than 64 down-spin hops from one state");
  QED_HIP(hipStreamSynchronize(stream()));
  T->ok = true;
  return T;
}

template <int KL, int NXT, int NPQ, int PF>
void launch_rows(const RowArgs& A, uint32_t grid, size_t lds, const double* x, double* z) {
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute((const void*)k_split_rows<KL, NXT, NPQ, PF>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  k_split_rows<KL, NXT, NPQ, PF><<<grid, kRowBS, lds, stream()>>>(A, x, z);
}

template <int KL, int NXT, int NPQ>
void launch_rows_pf(int pf, const RowArgs& A, uint32_t grid, size_t lds, const double* x, double* z) {
  if (pf <= 4) launch_rows<KL, NXT, NPQ, 4>(A, grid, lds, x, z);
  else if (pf <= 17) launch_rows<KL, NXT, NPQ, 17>(A, grid, lds, x, z);
  else launch_rows<KL, NXT, NPQ, 22>(A, grid, lds, x, z);
}

template <int KL>
void launch_rows_x(int nx, int pf, const RowArgs& A, uint32_t grid, size_t lds, const double* x, double* z) {
  if (nx <= 4) launch_rows_pf<KL, 4, 4>(pf, A, grid, lds, x, z);
  else if (nx <= 5) launch_rows_pf<KL, 5, 3>(pf, A, grid, lds, x, z);
  else launch_rows_pf<KL, 8, 2>(pf, A, grid, lds, x, z);
}

}  // namespace

// y += H x by the split-halves passes; returns false (nothing launched) when the sector's
// structure does not suit them.
bool hv_split(const qed_sector& s, const double* x, double* y, int passes) {
  if (!s.split) {
    s.split = build_split(s);
  }
  auto T = std::static_pointer_cast<SplitTables>(s.split);
  if (!T->ok) return false;
  if (T->z.bytes < s.D * 8) T->z.alloc(s.D * 8);  // the row pass's output (scratch)
  if (passes & 1) {
    RowArgs A;
    A.rec = T->rec.as<SRec>();
    A.lrec = T->lrec.as<LRec2>();
    A.rec0 = T->rec0.as<uint32_t>();
    A.tasks = T->tasks.as<Task>();
    A.task0 = T->task0.as<uint32_t>();
    A.bdn = s.bdn.as<uint32_t>();
    A.diag_dn = s.diag_dn.as<double>();
    A.du = s.du;
    A.dd = s.dd;
    A.k = T->k;
    A.nx = T->nx;
    A.U = s.hub_U;
    A.hub_sites = s.hub_sites;
    const uint32_t grid = std::min<uint32_t>(s.dd, (uint32_t)device_cu_count());
    const size_t lds = (size_t)s.du * 8;
    double* z = T->z.as<double>();
    if (T->kl <= 8) launch_rows_x<8>(T->nx, T->pf, A, grid, lds, x, z);
    else if (T->kl <= 14) launch_rows_x<14>(T->nx, T->pf, A, grid, lds, x, z);
    else launch_rows_x<16>(T->nx, T->pf, A, grid, lds, x, z);
    QED_HIP(hipGetLastError());
  }
  if (passes & 2) {
    DownArgs D;
    D.ptr = T->dptr.as<uint32_t>();
    D.nin = T->dnin.as<uint32_t>();
    D.off = T->didx.as<uint32_t>();
    D.amp = T->damp.as<double>();
    D.blk_off = T->blk_off.as<uint32_t>();
    D.blk_len = T->blk_len.as<uint32_t>();
    D.du = s.du;
    D.nblk = T->nblk;
    const uint32_t W = 32;
    D.npanels = (s.du + W - 1) / W;
    const uint64_t grid = (uint64_t)((D.npanels + 7) / 8) * 8ull * D.nblk;
    const size_t lds = (size_t)T->max_block * W * 8;
    const double* z = T->z.as<double>();
    k_split_down<32><<<(uint32_t)grid, kDownBS, lds, stream()>>>(D, x, z, y);
    QED_HIP(hipGetLastError());
  }
  return true;
}

std::string split_status(const qed_sector& s) {
  if (!s.split) s.split = build_split(s);
  auto T = std::static_pointer_cast<SplitTables>(s.split);
  return T->ok ? std::string() : T->why;
}

}  // namespace qed
