// Sectors that conserve the particle number or the spin but not both (the reference's
// mixings: spin flip, anomalous/Nambu, or both), trivial point group.
//
// The basis is the reference's ED_mixed_basis for a sector whose N or S is not conserved
// (src_ed/ED_basis.cpp:185-361, with the trivial group every state is its own
// representative): the union of the (N_up, N_down) blocks the sector allows, as 64-bit
// states (up | down << 32) in ascending order. The Hamiltonian terms follow
// HS_one_body_operator (src_ed/Operators/HS_one_body_operator.hpp:8-40; spin-flip
// elements are one-body elements between the two spin words), the density-density
// interaction, and HS_anomalous_operator (src_ed/Operators/HS_anomalous_operator.hpp:
// 8-40): for an element (r, c, v) of the operator (r < c, v as new_operator receives it)
// the pair (r, c) is annihilated from a state I by pair_annihilate (src_ed/binary_state.cpp:
// 36-43) and the element 2 v' pauli sits at (J, I) and (I, J), with v' = -v the value the
// operator stores for (c, r); with Jordan-Wigner operators that is 2 v (c_r c_c + h.c.).
//
// Device layout: the sorted states (8 B each); the product is a gather per basis state:
// each term applied to the state's bits, the target located by binary search.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/qed.h"
#include "qed_sector.hpp"

namespace qed {

namespace {

constexpr int kMaxOrbGeneral = 24;  // orbitals per spin

// one term of the product: kind 0 one-body hop or diagonal (c^+_r c_c, both directions
// listed), 1 density-density n_r n_c, 2 pair w (c_r c_c + h.c.); r, c as 64-bit masks
struct GTerm {
  uint64_t mr, mc;
  double w;
  int kind;
};

// Index of a state without a search, when every down word's allowed up words are those of
// one particle number or of one parity (the sectors of the spin-flip, anomalous and full
// mixings): states sort by down word, then up word, so
//   index(up, dn) = prefix[dn] + rank[up],
// with prefix[dn] the number of basis states of smaller down words and rank[up] the rank
// of up among the up words of its particle number (or of its parity); allowed[ndn] holds
// the allowed up particle numbers of down particle number ndn as bits.
struct LocateTables {
  const uint32_t* prefix;
  const uint32_t* rank;
  const uint64_t* allowed;
  int tabled;
};

struct GeneralTables {
  DevBuf states;  // uint64, ascending
  DevBuf terms;   // GTerm
  int nterms = 0;
  std::vector<uint64_t> host_states;
  DevBuf prefix, rank, allowed;
  bool tabled = false;
  LocateTables loc() const {
    return {prefix.as<uint32_t>(), rank.as<uint32_t>(), allowed.as<uint64_t>(), tabled ? 1 : 0};
  }
};

__host__ __device__ __forceinline__ uint64_t orbital_mask(int a, int L) {
  return a < L ? (1ull << a) : (1ull << (a - L + 32));
}

// index of s in the ascending array st[0..n), or n if absent
__device__ __forceinline__ uint32_t find_state(const uint64_t* __restrict__ st, uint32_t n, uint64_t s) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint64_t v = st[mid];
    if (v < s) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && st[lo] == s) ? lo : n;
}

// index of state s in the basis, or n when absent
__device__ __forceinline__ uint32_t locate(const LocateTables& T, const uint64_t* __restrict__ st, uint32_t n,
                                           uint64_t s) {
  if (!T.tabled) return find_state(st, n, s);
  const uint32_t up = (uint32_t)s, dn = (uint32_t)(s >> 32);
  if (!((T.allowed[__popc(dn)] >> __popc(up)) & 1ull)) return n;
  return T.prefix[dn] + T.rank[up];
}

// parity of the bits of s strictly between the single-bit masks a and b
__device__ __forceinline__ int between_parity(uint64_t s, uint64_t a, uint64_t b) {
  const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
  const uint64_t inside = (hi - 1ull) & ~((lo << 1) - 1ull);
  return __popcll(s & inside) & 1;
}

// y[I] += sum_J H[I, J] x[J]: every term applied to basis state I gives <J|T|I>, which is
// H[I, J] since the term list is Hermitian and real
__global__ __launch_bounds__(256) void k_hv_general(const uint64_t* __restrict__ st, uint32_t n, LocateTables loc,
                                                    const GTerm* __restrict__ terms, int nterms,
                                                    const double* __restrict__ x, double* __restrict__ y) {
  for (uint32_t I = blockIdx.x * blockDim.x + threadIdx.x; I < n; I += gridDim.x * blockDim.x) {
    const uint64_t s = st[I];
    double diag = 0.0, acc = 0.0;
    for (int k = 0; k < nterms; ++k) {
      const GTerm t = terms[k];
      if (t.kind == 1) {
        if ((s & t.mr) && (s & t.mc)) diag += t.w;
      } else if (t.kind == 0) {
        if (t.mr == t.mc) {
          if (s & t.mr) diag += t.w;
        } else if ((s & t.mc) && !(s & t.mr)) {
          const uint64_t s2 = s - t.mc + t.mr;
          const uint32_t J = locate(loc, st, n, s2);
          if (J == n) continue;
          const double v = t.w * x[J];
          acc += between_parity(s, t.mr, t.mc) ? -v : v;
        }
      } else {
        // <J| c_r c_c |I> = -(-1)^(bits between) when I holds both, and the same for the
        // conjugate when I holds neither (the bits between are those of I in both cases)
        const bool both = (s & t.mr) && (s & t.mc), none = !(s & t.mr) && !(s & t.mc);
        if (!both && !none) continue;
        const uint64_t s2 = both ? s - t.mr - t.mc : s + t.mr + t.mc;
        const uint32_t J = locate(loc, st, n, s2);
        if (J == n) continue;
        const double v = t.w * x[J];
        acc += between_parity(s, t.mr, t.mc) ? v : -v;
      }
    }
    y[I] += diag * x[I] + acc;
  }
}

// y[J] += z c_a x (pm = -1) or z c^+_a x (pm = +1) from the source basis to the target
// basis, gathered per target state: c_a |s> = (-1)^(bits of s below a) |s - a> and
// c^+_a |s> = (-1)^(bits of s below a) |s + a> (the 64-bit word's order, as Destroy,
// src_ed/ED_basis.cpp:60-77); the bits below a are the same in source and target
__global__ __launch_bounds__(256) void k_general_apply(const uint64_t* __restrict__ src, uint32_t ns,
                                                       LocateTables sloc, const uint64_t* __restrict__ tgt,
                                                       uint32_t nt, uint64_t m,
                                                       int pm, double z, const double* __restrict__ x,
                                                       double* __restrict__ y) {
  for (uint32_t J = blockIdx.x * blockDim.x + threadIdx.x; J < nt; J += gridDim.x * blockDim.x) {
    const uint64_t t = tgt[J];
    const bool has = (t & m) != 0;
    if (pm < 0 ? has : !has) continue;
    const uint64_t s = pm < 0 ? t + m : t - m;
    const uint32_t I = locate(sloc, src, ns, s);
    if (I == ns) continue;
    const double v = z * x[I];
    y[J] += (__popcll(t & (m - 1ull)) & 1) ? -v : v;
  }
}

// d[I] += H[I, I]: the one-body diagonal and density-density terms (no hop or pair term
// maps a state to itself)
__global__ __launch_bounds__(256) void k_diag_general(const uint64_t* __restrict__ st, uint32_t n,
                                                      const GTerm* __restrict__ terms, int nterms,
                                                      double* __restrict__ d) {
  for (uint32_t I = blockIdx.x * blockDim.x + threadIdx.x; I < n; I += gridDim.x * blockDim.x) {
    const uint64_t s = st[I];
    double acc = 0.0;
    for (int k = 0; k < nterms; ++k) {
      const GTerm t = terms[k];
      if (t.kind == 1) {
        if ((s & t.mr) && (s & t.mc)) acc += t.w;
      } else if (t.kind == 0 && t.mr == t.mc) {
        if (s & t.mr) acc += t.w;
      }
    }
    d[I] += acc;
  }
}

uint32_t grid_of(uint64_t n) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, (n + 255) / 256), 8192); }

}  // namespace

void hv_general(const qed_sector& s, const double* x, double* y, int variant) {
  auto T = std::static_pointer_cast<GeneralTables>(s.general);
  const uint32_t n = (uint32_t)s.D;
  if (n == 0) return;
  LocateTables loc = T->loc();
  if (variant == 0) loc.tabled = 0;  // the binary search (reference form of the lookup)
  k_hv_general<<<grid_of(n), 256, 0, stream()>>>(T->states.as<uint64_t>(), n, loc, T->terms.as<GTerm>(), T->nterms,
                                                 x, y);
  QED_HIP(hipGetLastError());
}

void diag_general(const qed_sector& s, double* d) {
  auto T = std::static_pointer_cast<GeneralTables>(s.general);
  const uint32_t n = (uint32_t)s.D;
  if (n == 0) return;
  k_diag_general<<<grid_of(n), 256, 0, stream()>>>(T->states.as<uint64_t>(), n, T->terms.as<GTerm>(), T->nterms, d);
  QED_HIP(hipGetLastError());
}

}  // namespace qed

using namespace qed;

extern "C" {

int qed_general_sector_create(int L, int npairs, const int* pairs, int n_ops, const int* types, const double* coefs,
                              const int* nels, const int* er, const int* ec, const double* ev, qed_sector** out) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    *out = nullptr;
    if (L <= 0 || L > kMaxOrbGeneral) fail("number of orbitals out of range");
    if (npairs < 0 || (npairs > 0 && !pairs)) fail("invalid particle-number blocks");
    // the basis: the blocks' states, ascending
    std::vector<uint32_t> half[kMaxOrbGeneral + 1];
    for (uint32_t b = 0; b < (1u << L); ++b) half[__builtin_popcount(b)].push_back(b);
    auto T = std::make_shared<GeneralTables>();
    uint64_t D = 0;
    std::vector<std::pair<int, int>> blocks;
    for (int k = 0; k < npairs; ++k) {
      const int nu = pairs[2 * k], nd = pairs[2 * k + 1];
      if (nu < 0 || nd < 0 || nu > L || nd > L) continue;
      blocks.push_back({nu, nd});
      D += (uint64_t)half[nu].size() * half[nd].size();
    }
    if (D >= (1ull << 31)) fail("sector too large for the mixed basis (2^31 states)");
    T->host_states.reserve(D);
    for (auto& bl : blocks)
      for (uint32_t d : half[bl.second])
        for (uint32_t u : half[bl.first]) T->host_states.push_back((uint64_t)u | ((uint64_t)d << 32));
    std::sort(T->host_states.begin(), T->host_states.end());
    T->host_states.erase(std::unique(T->host_states.begin(), T->host_states.end()), T->host_states.end());
    D = T->host_states.size();
    // lookup tables: every down particle number allows one up particle number, or every
    // up particle number of one parity
    std::vector<uint64_t> allowed(kMaxOrbGeneral + 1, 0);
    for (auto& bl : blocks) allowed[bl.second] |= 1ull << bl.first;
    bool single = true, parity = true;
    for (int nd = 0; nd <= L; ++nd) {
      const uint64_t a = allowed[nd];
      if (!a) continue;
      if (__builtin_popcountll(a) != 1) single = false;
      for (int nu = 0; nu <= L; ++nu) {
        const bool want = ((nu + nd) % 2) == ((__builtin_ctzll(a) + nd) % 2);
        if (((a >> nu) & 1ull) != (uint64_t)want) parity = false;
      }
    }
    if ((single || parity) && L <= 20) {
      const uint32_t nw = 1u << L;
      std::vector<uint32_t> rank(nw), prefix(nw);
      std::vector<uint32_t> cnt(L + 2, 0);
      for (uint32_t u = 0; u < nw; ++u) {
        const int pc = __builtin_popcount(u);
        rank[u] = single ? cnt[pc]++ : cnt[pc & 1]++;
      }
      std::vector<uint64_t> per(L + 1, 0);  // basis states per down word of each particle number
      for (int nd = 0; nd <= L; ++nd)
        for (int nu = 0; nu <= L; ++nu)
          if ((allowed[nd] >> nu) & 1ull) per[nd] += half[nu].size();
      uint64_t run = 0;
      for (uint32_t d = 0; d < nw; ++d) {
        prefix[d] = (uint32_t)run;
        run += per[__builtin_popcount(d)];
      }
      if (run != D) fail("mixed basis: inconsistent lookup tables");
      T->prefix.alloc((size_t)nw * 4);
      T->rank.alloc((size_t)nw * 4);
      T->allowed.alloc(allowed.size() * 8);
      QED_HIP(hipMemcpyAsync(T->prefix.p, prefix.data(), (size_t)nw * 4, hipMemcpyHostToDevice, stream()));
      QED_HIP(hipMemcpyAsync(T->rank.p, rank.data(), (size_t)nw * 4, hipMemcpyHostToDevice, stream()));
      QED_HIP(hipMemcpyAsync(T->allowed.p, allowed.data(), allowed.size() * 8, hipMemcpyHostToDevice, stream()));
      QED_HIP(hipStreamSynchronize(stream()));
      T->tabled = true;
    }
    // the terms
    std::vector<GTerm> terms;
    size_t off = 0;
    for (int o = 0; o < n_ops; ++o) {
      const double z = coefs[o];
      for (int e = 0; e < nels[o]; ++e, ++off) {
        const int r = er[off], c = ec[off];
        const double v = ev[off] * z;
        if (r < 0 || c < 0 || r >= 2 * L || c >= 2 * L) fail("operator element index out of range");
        if (v == 0.0) continue;
        const uint64_t mr = orbital_mask(r, L), mc = orbital_mask(c, L);
        if (types[o] == 0) {
          // one_body_operator lists (r, c, v) and (c, r, v) for r < c (one_body_operator.hpp:70-76)
          if (r == c) {
            terms.push_back({mr, mr, v, 0});
          } else if (r < c) {
            terms.push_back({mr, mc, v, 0});
            terms.push_back({mc, mr, v, 0});
          }
        } else if (types[o] == 1) {
          terms.push_back({mr, mc, v, 1});
        } else if (types[o] == 2) {
          if (r < c) terms.push_back({mr, mc, 2.0 * v, 2});
        } else {
          fail("unknown operator type");
        }
      }
    }
    T->nterms = (int)terms.size();
    T->states.alloc(std::max<uint64_t>(1, D) * 8);
    T->terms.alloc(std::max<size_t>(1, terms.size()) * sizeof(GTerm));
    if (D) QED_HIP(hipMemcpyAsync(T->states.p, T->host_states.data(), D * 8, hipMemcpyHostToDevice, stream()));
    if (!terms.empty())
      QED_HIP(hipMemcpyAsync(T->terms.p, terms.data(), terms.size() * sizeof(GTerm), hipMemcpyHostToDevice, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    auto s = std::make_unique<qed_sector>();
    s->L = L;
    s->nup = -1;
    s->ndn = -1;
    s->D = D;
    s->general = T;
    *out = s.release();
  });
}

int qed_general_basis(const qed_sector* s, uint64_t* states, uint64_t* n) {
  return guard([&] {
    if (!s->general) fail("not a mixed-number sector");
    auto T = std::static_pointer_cast<GeneralTables>(s->general);
    *n = T->host_states.size();
    if (states) std::copy(T->host_states.begin(), T->host_states.end(), states);
  });
}

int qed_general_apply(const qed_sector* src, const qed_sector* tgt, int orbital, int pm, double z, const double* x,
                      double* y) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (!src->general || !tgt->general) fail("qed_general_apply: both sectors must be mixed-number sectors");
    if (src->L != tgt->L) fail("qed_general_apply: the sectors have different orbitals");
    if (orbital < 0 || orbital >= 2 * src->L) fail("qed_general_apply: orbital out of range");
    if (pm != 1 && pm != -1) fail("qed_general_apply: pm must be 1 or -1");
    auto S = std::static_pointer_cast<GeneralTables>(src->general);
    auto T = std::static_pointer_cast<GeneralTables>(tgt->general);
    if (tgt->D && src->D)
      k_general_apply<<<grid_of(tgt->D), 256, 0, stream()>>>(S->states.as<uint64_t>(), (uint32_t)src->D, S->loc(),
                                                             T->states.as<uint64_t>(), (uint32_t)tgt->D,
                                                             orbital_mask(orbital, src->L), pm, z, x, y);
    QED_HIP(hipGetLastError());
    QED_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"
