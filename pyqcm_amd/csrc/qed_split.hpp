// Host-side plan of the split-halves H.v.
//
// A half state s of L orbitals splits at orbital k into a low part a (orbitals < k) and a
// high part b (orbitals >= k): s = a | b << k. In ascending (colex) order the states with
// one high part b form a contiguous block of C(k, N - popcount(b)) states, ordered by the
// colex rank of a among the k-orbital strings with that particle number:
//   rank(s) = off(b) + rank_{N - |b|}(a).
// A hop c†_i c_j (both orbitals on one side of k) changes only a or only b, and keeps the
// class |b|; a hop across k changes both and moves one particle between the two parts.
// The plan below lists, per b, the hops inside the high part and, per a, the hops inside
// the low part, so that the kernels read them from small tables instead of per-state hop
// lists.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace qed {
namespace split {

constexpr int kMaxHigh = 16;   // hops inside the high part per b
constexpr int kMaxLow = 16;    // hops inside the low part per a
constexpr int kMaxXSites = 8;  // orbitals of either part that carry bonds across k
constexpr int kMaxPairs = 32;  // bonds across k

// One undirected bond of the one-body operator of a spin: the hop c†_i c_j + c†_j c_i
// with amplitude t (the sum of the elements on that pair).
struct Pair {
  int i, j;
  double t;
};

inline int popc(uint32_t x) { return __builtin_popcount(x); }

inline uint64_t binom(int n, int r) {
  if (r < 0 || r > n) return 0;
  uint64_t c = 1;
  for (int q = 1; q <= r; ++q) c = c * (uint64_t)(n - r + q) / (uint64_t)q;
  return c;
}

// Jordan-Wigner parity of the hop between orbitals i and j of state s (orbitals strictly
// between them), the convention of the hop tables.
inline int hop_parity(uint32_t s, int i, int j) {
  int lo = std::min(i, j), hi = std::max(i, j);
  uint32_t between = ((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u);
  return popc(s & between) & 1;
}

// The split structure of one spin's half space.
struct Spin {
  int L = 0, N = 0, k = 0;
  uint32_t dim = 0;
  // valid high parts, ascending
  std::vector<uint32_t> b_val, b_off;
  std::map<uint32_t, int> b_index;
  // classes by the particle number of the low part: ascending strings and ranks
  std::vector<std::vector<uint32_t>> a_of;       // a_of[pa]
  std::vector<std::map<uint32_t, uint32_t>> a_rank;
  std::vector<Pair> low, high, cross;
  std::vector<double> dsite;  // one-body diagonal per orbital
  std::vector<int> lx, hx;     // orbitals of the low / high part that carry bonds across k

  uint32_t rank(uint32_t s) const {
    uint32_t a = s & ((1u << k) - 1u), b = s >> k;
    auto it = b_index.find(b);
    if (it == b_index.end()) throw std::runtime_error("split: state outside the half space");
    return b_off[it->second] + a_rank[popc(a)].at(a);
  }
  double diag_low(uint32_t a) const {
    double e = 0.0;
    for (int i = 0; i < k; ++i)
      if (a >> i & 1u) e += dsite[i];
    return e;
  }
  double diag_high(uint32_t b) const {
    double e = 0.0;
    for (int i = k; i < L; ++i)
      if (b >> (i - k) & 1u) e += dsite[i];
    return e;
  }
};

// pairs: merged undirected bonds (i != j); dsite: diagonal per orbital (size L)
inline Spin make_spin(int L, int N, int k, const std::vector<Pair>& pairs, const std::vector<double>& dsite) {
  Spin S;
  S.L = L;
  S.N = N;
  S.k = k;
  S.dsite = dsite;
  S.dsite.resize(L, 0.0);
  S.a_of.assign(k + 1, {});
  S.a_rank.assign(k + 1, {});
  for (uint32_t a = 0; a < (1u << k); ++a) S.a_of[popc(a)].push_back(a);
  for (int p = 0; p <= k; ++p)
    for (uint32_t q = 0; q < S.a_of[p].size(); ++q) S.a_rank[p][S.a_of[p][q]] = q;
  uint32_t off = 0;
  for (uint32_t b = 0; b < (1u << (L - k)); ++b) {
    int pa = N - popc(b);
    if (pa < 0 || pa > k) continue;
    S.b_index[b] = (int)S.b_val.size();
    S.b_val.push_back(b);
    S.b_off.push_back(off);
    off += (uint32_t)S.a_of[pa].size();
  }
  S.dim = off;
  for (const Pair& p : pairs) {
    int i = std::min(p.i, p.j), j = std::max(p.i, p.j);
    Pair q{i, j, p.t};
    if (j < k) S.low.push_back(q);
    else if (i >= k) S.high.push_back(q);
    else S.cross.push_back(q);
  }
  for (const Pair& p : S.cross) {
    if (std::find(S.lx.begin(), S.lx.end(), p.i) == S.lx.end()) S.lx.push_back(p.i);
    if (std::find(S.hx.begin(), S.hx.end(), p.j) == S.hx.end()) S.hx.push_back(p.j);
  }
  std::sort(S.lx.begin(), S.lx.end());
  std::sort(S.hx.begin(), S.hx.end());
  return S;
}

// Whether the split kernels can take this structure.
inline bool fits(const Spin& S, std::string* why = nullptr) {
  auto no = [&](const char* m) {
    if (why) *why = m;
    return false;
  };
  if ((int)S.lx.size() > kMaxXSites || (int)S.hx.size() > kMaxXSites) return no("too many orbitals on bonds across the split");
  if ((int)S.cross.size() > kMaxPairs) return no("too many bonds across the split");
  for (uint32_t b : S.b_val) {
    int nh = 0;
    for (const Pair& p : S.high)
      if (((b >> (p.i - S.k)) ^ (b >> (p.j - S.k))) & 1u) ++nh;
    if (nh > kMaxHigh) return no("too many hops inside the high part");
  }
  for (int pa = 0; pa <= S.k; ++pa)
    for (uint32_t a : S.a_of[pa]) {
      int nl = 0;
      for (const Pair& p : S.low)
        if (((a >> p.i) ^ (a >> p.j)) & 1u) ++nl;
      if (nl > kMaxLow) return no("too many hops inside the low part");
    }
  return true;
}

// A unit of the row pass: up to 64 low strings of one class (one per lane) and the
// sequence of high parts b of that class they run through.
struct Unit {
  int pa = 0;
  std::vector<uint32_t> a;      // low strings (lanes)
  std::vector<uint32_t> bseq;   // indices into b_val
  int kl = 0;                   // hops inside the low part: maximum over the lanes
};

// A task of one wave: unit u, steps t0 .. t1-1 of its b sequence.
struct Task {
  uint32_t unit, t0, t1, pad;
};

struct RowPlan {
  int k = 0;
  std::vector<Unit> units;
  std::vector<std::vector<Task>> wave_tasks;  // per wave of the workgroup
  int kl = 0;                                 // maximum over the units
  int kh = 0;                                 // hops inside the high part: maximum over b
  int npq = 0;                                // bonds across the split per low orbital: maximum
  uint64_t steps = 0;                         // wave steps per row
  double cost = 0;                            // estimated LDS reads per row
};

inline int count_low(const Spin& S, uint32_t a) {
  int n = 0;
  for (const Pair& p : S.low)
    if (((a >> p.i) ^ (a >> p.j)) & 1u) ++n;
  return n;
}
inline int count_high(const Spin& S, uint32_t b) {
  int n = 0;
  for (const Pair& p : S.high)
    if (((b >> (p.i - S.k)) ^ (b >> (p.j - S.k))) & 1u) ++n;
  return n;
}

// Units of up to 64 lanes per class (a class of na strings is cut into ceil(na/64) equal
// chunks); each unit's b sequence is cut into pieces so that the largest piece is as
// small as the wave count allows; units too small to deserve a wave of their own share
// the last wave (which then reloads its lane records at each task).
inline RowPlan row_plan(const Spin& S, int nwaves) {
  RowPlan P;
  P.k = S.k;
  for (int pa = 0; pa <= S.k; ++pa) {
    const auto& A = S.a_of[pa];
    if (A.empty()) continue;
    std::vector<uint32_t> bs;
    for (size_t q = 0; q < S.b_val.size(); ++q)
      if (S.N - popc(S.b_val[q]) == pa) bs.push_back((uint32_t)q);
    if (bs.empty()) continue;
    const size_t nch = (A.size() + 63) / 64, per = (A.size() + nch - 1) / nch;
    for (size_t a0 = 0; a0 < A.size(); a0 += per) {
      Unit u;
      u.pa = pa;
      u.a.assign(A.begin() + a0, A.begin() + std::min(A.size(), a0 + per));
      u.bseq = bs;
      for (uint32_t a : u.a) u.kl = std::max(u.kl, count_low(S, a));
      P.kl = std::max(P.kl, u.kl);
      P.units.push_back(std::move(u));
    }
  }
  for (uint32_t b : S.b_val) P.kh = std::max(P.kh, count_high(S, b));
  std::map<int, int> per_lx;
  for (const Pair& p : S.cross) P.npq = std::max(P.npq, ++per_lx[p.i]);
  P.wave_tasks.assign(nwaves, {});
  if (P.units.empty()) return P;
  // pieces of at most `target` steps dealt to the least loaded wave, longest first (LPT);
  // a wave that takes pieces of several units reloads its lane records at each task, which
  // is counted as `switch_cost` steps. The target with the smallest maximum load wins.
  const double switch_cost = 1.5;
  uint64_t total = 0, longest = 0;
  for (const Unit& u : P.units) {
    total += u.bseq.size();
    longest = std::max<uint64_t>(longest, u.bseq.size());
  }
  double best = 1e300;
  std::vector<std::vector<Task>> best_tasks;
  for (uint64_t target = std::max<uint64_t>(1, (total + nwaves - 1) / nwaves); target <= longest; ++target) {
    struct Piece {
      uint32_t unit, t0, t1;
    };
    std::vector<Piece> pieces;
    for (uint32_t q = 0; q < P.units.size(); ++q) {
      const uint32_t n = (uint32_t)P.units[q].bseq.size();
      const uint32_t parts = (uint32_t)((n + target - 1) / target);
      for (uint32_t p = 0; p < parts; ++p) pieces.push_back({q, n * p / parts, n * (p + 1) / parts});
    }
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const Piece& x, const Piece& y) { return x.t1 - x.t0 > y.t1 - y.t0; });
    std::vector<std::vector<Task>> wt(nwaves);
    std::vector<double> load(nwaves, 0.0);
    for (const Piece& p : pieces) {
      int w = 0;
      double lw = 1e300;
      for (int q = 0; q < nwaves; ++q) {
        const double l = load[q] + (p.t1 - p.t0) + (wt[q].empty() ? 0.0 : switch_cost);
        if (l < lw) {
          lw = l;
          w = q;
        }
      }
      wt[w].push_back({p.unit, p.t0, p.t1, 0});
      load[w] = lw;
    }
    const double mx = *std::max_element(load.begin(), load.end());
    if (mx < best) {
      best = mx;
      best_tasks = wt;
    }
  }
  P.wave_tasks = best_tasks;
  uint64_t mx = 0;
  for (const auto& wt : P.wave_tasks) {
    uint64_t n = 0;
    for (const Task& t : wt) n += t.t1 - t.t0;
    mx = std::max(mx, n);
  }
  P.steps = mx;
  P.cost = (double)mx * (1 + P.kl + P.kh + (double)S.cross.size());
  return P;
}

// The per-step record of the row pass (one high part b of a unit's sequence), 384 bytes:
// everything a wave needs for the block of b, at fixed offsets.
constexpr int kRecHigh = 12;   // hops inside the high part (padded with zero-amplitude reads of b's own block)
constexpr int kRecCross = 16;  // bonds across the split, as (low orbital, partner) slots
struct SRec {
  uint32_t off8;   // byte offset of the block in the row
  uint32_t bits;   // b
  uint32_t xbj;    // bit s: occupation in b of the high orbital of cross slot s
  uint32_t pad0;
  double e;        // one-body diagonal of the high part
  uint32_t hoff8[kRecHigh];
  double hamp[kRecHigh];
  uint32_t xoff8[kRecCross];  // byte offset of the block of b ^ (high orbital of slot s)
  double xcoef[kRecCross];    // slot amplitude with the high part's Jordan-Wigner sign (0: unused)
  uint32_t pad1[6];
};
static_assert(sizeof(SRec) == 384, "SRec layout");

// The per-lane record of a unit (one low string a), 256 bytes.
constexpr int kLaneLow = 16;
constexpr int kLaneX = 8;
struct LRec2 {
  uint32_t ra8;    // byte offset of a in its block
  uint32_t valid;  // 1 for a lane that carries a string
  uint32_t abits;  // a
  uint32_t xsa;    // bit q: Jordan-Wigner sign of the low part for low cross orbital q;
                   // bit 16 + q: occupation of that orbital in a
  double e;        // one-body diagonal of the low part
  uint32_t xra8[kLaneX];    // byte offset of a ^ (low cross orbital q) in its block
  uint32_t lra8[kLaneLow];  // byte offsets of the targets of the hops inside the low part
  double lamp[kLaneLow];    // their signed amplitudes (0 for padding)
  uint32_t pad[2];
};
static_assert(sizeof(LRec2) == 256, "LRec2 layout");

// cross slots: slot s = q * npq + m is the m-th bond of low cross orbital lx[q]
struct XSlots {
  int npq = 0;
  std::vector<int> hx_of;    // per slot: index into hx (-1 for an unused slot)
  std::vector<double> t;     // per slot: amplitude (0 for an unused slot)
};
inline XSlots x_slots(const Spin& S, int npq) {
  XSlots X;
  X.npq = npq;
  X.hx_of.assign(S.lx.size() * npq, -1);
  X.t.assign(S.lx.size() * npq, 0.0);
  std::vector<int> fill(S.lx.size(), 0);
  for (const Pair& p : S.cross) {
    const int q = (int)(std::find(S.lx.begin(), S.lx.end(), p.i) - S.lx.begin());
    const int jx = (int)(std::find(S.hx.begin(), S.hx.end(), p.j) - S.hx.begin());
    const int s = q * npq + fill[q]++;
    X.hx_of[s] = jx;
    X.t[s] = p.t;
  }
  return X;
}

inline SRec s_record(const Spin& S, const XSlots& X, uint32_t bq) {
  const uint32_t b = S.b_val[bq];
  SRec r{};
  r.off8 = S.b_off[bq] * 8u;
  r.bits = b;
  r.e = S.diag_high(b);
  int n = 0;
  for (const Pair& p : S.high) {
    const int bi = p.i - S.k, bj = p.j - S.k;
    if (!(((b >> bi) ^ (b >> bj)) & 1u)) continue;
    const uint32_t b2 = b ^ (1u << bi) ^ (1u << bj);
    r.hoff8[n] = S.b_off[S.b_index.at(b2)] * 8u;
    r.hamp[n] = hop_parity(b << S.k, p.i, p.j) ? -p.t : p.t;
    ++n;
  }
  for (; n < kRecHigh; ++n) {
    r.hoff8[n] = r.off8;
    r.hamp[n] = 0.0;
  }
  for (size_t s = 0; s < X.hx_of.size(); ++s) {
    if (X.hx_of[s] < 0) continue;
    const int j = S.hx[X.hx_of[s]] - S.k;
    r.xbj |= (b >> j & 1u) << s;
    const int sb = popc(b & ((1u << j) - 1u)) & 1;
    r.xcoef[s] = sb ? -X.t[s] : X.t[s];
    auto it = S.b_index.find(b ^ (1u << j));
    r.xoff8[s] = it == S.b_index.end() ? 0u : S.b_off[it->second] * 8u;
  }
  return r;
}

inline LRec2 l_record2(const Spin& S, int pa, uint32_t a) {
  LRec2 r{};
  r.ra8 = S.a_rank[pa].at(a) * 8u;
  r.valid = 1;
  r.abits = a;
  r.e = S.diag_low(a);
  int n = 0;
  for (const Pair& p : S.low) {
    if (!(((a >> p.i) ^ (a >> p.j)) & 1u)) continue;
    r.lra8[n] = S.a_rank[pa].at(a ^ (1u << p.i) ^ (1u << p.j)) * 8u;
    r.lamp[n] = hop_parity(a, p.i, p.j) ? -p.t : p.t;
    ++n;
  }
  for (; n < kLaneLow; ++n) {
    r.lra8[n] = r.ra8;
    r.lamp[n] = 0.0;
  }
  for (size_t q = 0; q < S.lx.size(); ++q) {
    const int i = S.lx[q];
    r.xsa |= (uint32_t)(popc(a & ~((1u << (i + 1)) - 1u)) & 1) << q;
    r.xsa |= (a >> i & 1u) << (16 + q);
    const uint32_t a2 = a ^ (1u << i);
    r.xra8[q] = S.a_rank[popc(a2)].at(a2) * 8u;
  }
  return r;
}

// the split point with the smallest estimated row-pass cost among those the kernels take
inline int choose_k(int L, int N, const std::vector<Pair>& pairs, const std::vector<double>& dsite, int nwaves,
                    std::string* why = nullptr) {
  int best = -1;
  double best_cost = 0;
  for (int k = std::max(1, L / 2 - 2); k <= std::min(L - 1, L / 2 + 2); ++k) {
    const Spin S = make_spin(L, N, k, pairs, dsite);
    std::string w;
    if (!fits(S, &w)) {
      if (why) *why = w;
      continue;
    }
    const RowPlan P = row_plan(S, nwaves);
    if (P.kl > kLaneLow || P.kh > kRecHigh || (int)S.lx.size() > kLaneX ||
        (int)S.lx.size() * std::max(P.npq, 1) > kRecCross)
      continue;
    if (best < 0 || P.cost < best_cost) {
      best = k;
      best_cost = P.cost;
    }
  }
  return best;
}

// Down-spin hop lists per row for the tiled pass: per row r of block b, first the hops to
// rows of the same block (as ranks inside the block), then the others (as row indices).
struct DownPlan {
  std::vector<uint32_t> ptr;   // dim + 1
  std::vector<uint32_t> nin;   // per row: hops inside the block (listed first)
  std::vector<uint32_t> idx;   // rank inside the block, or row index
  std::vector<double> amp;
  std::vector<uint32_t> blk_off, blk_len;  // blocks (= high parts) in ascending order
  uint32_t max_block = 0;
};

inline DownPlan down_plan(const Spin& S) {
  DownPlan D;
  D.ptr.assign(S.dim + 1, 0);
  D.nin.assign(S.dim, 0);
  for (size_t q = 0; q < S.b_val.size(); ++q) {
    const uint32_t b = S.b_val[q];
    const int pa = S.N - popc(b);
    const auto& A = S.a_of[pa];
    D.blk_off.push_back(S.b_off[q]);
    D.blk_len.push_back((uint32_t)A.size());
    D.max_block = std::max<uint32_t>(D.max_block, (uint32_t)A.size());
    for (uint32_t a : A) {
      const uint32_t s = a | (b << S.k);
      const uint32_t r = S.b_off[q] + S.a_rank[pa].at(a);
      std::vector<std::pair<uint32_t, double>> in, out;
      for (const Pair& p : S.low) {
        if (!(((s >> p.i) ^ (s >> p.j)) & 1u)) continue;
        const uint32_t a2 = a ^ (1u << p.i) ^ (1u << p.j);
        in.push_back({S.a_rank[pa].at(a2), hop_parity(s, p.i, p.j) ? -p.t : p.t});
      }
      for (const Pair& p : S.high) {
        if (!(((s >> p.i) ^ (s >> p.j)) & 1u)) continue;
        out.push_back({S.rank(s ^ (1u << p.i) ^ (1u << p.j)), hop_parity(s, p.i, p.j) ? -p.t : p.t});
      }
      for (const Pair& p : S.cross) {
        if (!(((s >> p.i) ^ (s >> p.j)) & 1u)) continue;
        out.push_back({S.rank(s ^ (1u << p.i) ^ (1u << p.j)), hop_parity(s, p.i, p.j) ? -p.t : p.t});
      }
      // rows are visited in ascending order, so the lists are appended in row order
      D.nin[r] = (uint32_t)in.size();
      D.ptr[r + 1] = (uint32_t)(in.size() + out.size());
      for (const auto& e : in) {
        D.idx.push_back(e.first);
        D.amp.push_back(e.second);
      }
      for (const auto& e : out) {
        D.idx.push_back(e.first);
        D.amp.push_back(e.second);
      }
    }
  }
  for (uint32_t r = 0; r < S.dim; ++r) D.ptr[r + 1] += D.ptr[r];
  return D;
}

}  // namespace split
}  // namespace qed
