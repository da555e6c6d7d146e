// ORACLE — test infrastructure only.
//
// CPU restatement of the reference ED solver's factorized path (pyqcm, src_ed/),
// written from the reference's documented behaviour, used as the checker for the
// HIP path and as the timed CPU baseline. Nothing in the product links this file:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Function-by-function provenance (reference file:line):
//   half_basis          src_ed/ED_basis.cpp:115-135    (ascending integers with popcount N)
//   half_index          src_ed/ED_basis.cpp:130-135    (lower_bound, returns dim if absent)
//   hop_sign            src_ed/binary_state.cpp:23-28, binary_state.hpp:42-47 (one_body, interphase)
//   HalfOp build        src_ed/Operators/HS_half_operator.hpp:17-67 (lower triangle, grouped by value)
//   Hermitian completion src_ed/Operators/one_body_operator.hpp:67-76
//   interaction elem    src_ed/Operators/HS_interaction_operator.hpp:32-55 (incl. the
//                       factorized-basis word order: up in the high word, mask() of an
//                       orbital < L in the low word, src_ed/ED_basis.cpp:157-160)
//   hv (y += H x)       src_ed/Hamiltonian/Hamiltonian_Factorized.hpp:50-58,
//                       src_ed/Operators/HS_factorized_operator.hpp:56-84,
//                       src_ed/Operators/HS_interaction_operator.hpp:62-70
//   start vector        src_util/vector_num.cpp:114-119 (default_random_engine(seed) per call,
//                       one process-wide normal_distribution, then normalize)
//   lanczos_gs          src_ed/Hamiltonian/Lanczos.hpp:41-143, Lanczos.cpp:25-52 (dstev)
//   lanczos_green       src_ed/Hamiltonian/Lanczos.hpp:156-183
//   create_destroy      src_ed/model.cpp:155-180, src_ed/ED_basis.cpp:79-110
//
// Parity pins for every function above live in tests/test_oracle.py (brute-force
// Jordan-Wigner Hamiltonians, closed forms, exhaustive rank checks, reference fixtures).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
// LAPACK / CBLAS (MKL runtime, the same LAPACK entry point the reference calls)
void dstev_(const char* jobz, const int* n, double* d, double* e, double* z, const int* ldz,
            double* work, int* info);
double cblas_ddot(const int n, const double* x, const int incx, const double* y, const int incy);
double cblas_dnrm2(const int n, const double* x, const int incx);
void cblas_daxpy(const int n, const double a, const double* x, const int incx, double* y,
                 const int incy);
void cblas_dscal(const int n, const double a, double* x, const int incx);
int MKL_Set_Threading_Layer(int code);
}

// BLAS-1 and dstev run sequentially, as in the reference's documented CPU runs
// (bench/ED/bench_ED2_laptop.txt: OPENBLAS_NUM_THREADS=1); 1 = MKL_THREADING_SEQUENTIAL.
static const int g_mkl_layer = MKL_Set_Threading_Layer(1);

namespace {

thread_local std::string g_err;

// One normal_distribution for the whole process, as the reference keeps one global
// instance (src_ed/Hamiltonian/Lanczos.cpp:7); its cached second deviate carries over
// between calls exactly as there.
std::normal_distribution<double> g_normal(0.0, 1.0);

int popcount32(uint32_t x) { return __builtin_popcount(x); }

std::vector<uint32_t> half_basis(int L, int N) {
  std::vector<uint32_t> b;
  if (N < 0 || N > L) return b;
  uint32_t top = (L >= 32) ? 0xffffffffu : ((1u << L) - 1u);
  for (uint64_t s = 0; s <= top; ++s)
    if (popcount32((uint32_t)s) == N) b.push_back((uint32_t)s);
  return b;
}

uint32_t half_index(const std::vector<uint32_t>& bin, uint32_t s) {
  auto it = std::lower_bound(bin.begin(), bin.end(), s);
  uint32_t I = (uint32_t)(it - bin.begin());
  if (I < bin.size() && bin[I] == s) return I;
  return (uint32_t)bin.size();
}

// c†_to c_from on a half state: 0 if Pauli-blocked, else +-1 with the parity of the
// occupied orbitals strictly between `from` and `to`; *out gets the new state.
int hop_sign(uint32_t s, int from, int to, uint32_t* out) {
  uint32_t mf = 1u << from, mt = 1u << to;
  if (from == to) {
    if (!(s & mf)) return 0;
    *out = s;
    return 1;
  }
  if ((s & mt) || !(s & mf)) return 0;
  uint32_t lo = from < to ? from : to, hi = from < to ? to : from;
  uint32_t between = ((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u);
  *out = s - mf + mt;
  return (popcount32(s & between) & 1) ? -1 : 1;
}

struct Elem {
  int r, c;
  double v;
};

// Per-spin factor of one one-body operator on one half basis.
struct HalfOp {
  std::vector<std::pair<double, std::vector<std::pair<uint32_t, uint32_t>>>> groups;
  std::vector<std::pair<uint32_t, double>> diag;
  uint32_t dim = 0;
};

HalfOp build_half(const std::vector<Elem>& el, const std::vector<uint32_t>& bin, int L,
                  bool spin_down) {
  HalfOp h;
  h.dim = (uint32_t)bin.size();
  int offset = spin_down ? L : 0;
  for (uint32_t I = 0; I < h.dim; ++I) {
    uint32_t s = bin[I];
    for (const Elem& x : el) {
      bool is_dn = x.r >= L;
      if (is_dn != spin_down) continue;
      int r = x.r - offset, c = x.c - offset;
      if (x.r != x.c) {
        uint32_t sp;
        int ph = hop_sign(s, c, r, &sp);
        if (ph == 0) continue;
        uint32_t J = half_index(bin, sp);
        if (J == h.dim) continue;
        double X = ph * x.v;
        if (std::fabs(X) < 1.0e-12) continue;
        if (J > I) continue;
        if (J == I) {
          h.diag.push_back({I, X});
          continue;
        }
        size_t g = 0;
        for (; g < h.groups.size(); ++g)
          if (h.groups[g].first == X) {
            h.groups[g].second.push_back({I, J});
            break;
          }
        if (g == h.groups.size()) h.groups.push_back({X, {{I, J}}});
      } else if (s & (1u << r)) {
        if (std::fabs(x.v) < 1.0e-12) continue;
        h.diag.push_back({I, x.v});
      }
    }
  }
  for (auto& g : h.groups) std::sort(g.second.begin(), g.second.end());
  std::stable_sort(h.diag.begin(), h.diag.end(),
                   [](const std::pair<uint32_t, double>& a, const std::pair<uint32_t, double>& b) {
                     return a.first < b.first;
                   });
  return h;
}

struct Op {
  int type;  // 0 one-body, 1 interaction
  double coef;
  HalfOp up, dn;                  // one-body
  std::vector<double> elem;       // interaction, one value per sector index
};

struct Sector {
  int L, nup, ndn;
  std::vector<uint32_t> bup, bdn;
  uint64_t dim;
  std::vector<Op> ops;
};

uint64_t mask64(int i, int L) {
  return (i >= L) ? (uint64_t(1) << (i - L + 32)) : (uint64_t(1) << i);
}

void hv_op(const Sector& S, const Op& op, const double* x, double* y) {
  const uint64_t du = S.bup.size(), dd = S.bdn.size();
  const double z = op.coef;
  if (op.type == 1) {
    for (uint64_t i = 0; i < S.dim; ++i) y[i] += x[i] * op.elem[i] * z;
    return;
  }
  // spin-down factor: row AXPYs of length dim_up
  for (const auto& w : op.dn.groups) {
    double z2 = w.first * z;
    for (const auto& e : w.second)
      for (uint64_t i = 0; i < du; ++i) y[e.first * du + i] += z2 * x[e.second * du + i];
    for (const auto& e : w.second)
      for (uint64_t i = 0; i < du; ++i) y[e.second * du + i] += z2 * x[e.first * du + i];
  }
  for (const auto& e : op.dn.diag)
    for (uint64_t i = 0; i < du; ++i) y[e.first * du + i] += z * e.second * x[e.first * du + i];
  // spin-up factor: gathers inside each row
  for (uint64_t i = 0; i < dd; ++i) {
    for (const auto& w : op.up.groups) {
      double z2 = w.first * z;
      for (const auto& e : w.second) y[e.first + du * i] += z2 * x[e.second + du * i];
      for (const auto& e : w.second) y[e.second + du * i] += z2 * x[e.first + du * i];
    }
    for (const auto& e : op.up.diag) y[e.first + du * i] += z * e.second * x[e.first + du * i];
  }
}

void hv_all(const Sector& S, const double* x, double* y) {
  for (const Op& op : S.ops) hv_op(S, op, x, y);
}

// Same arithmetic, parallel over down rows (OpenMP). Not the reference's loop order:
// this is the project's own parallel CPU variant, reported separately.
void hv_all_omp(const Sector& S, const double* x, double* y) {
  const uint64_t du = S.bup.size(), dd = S.bdn.size();
  // collect per-row down hops: for row r, list of (source row, amplitude*coef)
  std::vector<std::vector<std::pair<uint32_t, double>>> dn_rows(dd);
  std::vector<double> dn_diag(dd, 0.0), up_diag(du, 0.0);
  std::vector<std::vector<std::pair<uint32_t, double>>> up_cols(du);
  for (const Op& op : S.ops) {
    if (op.type != 0) continue;
    for (const auto& w : op.dn.groups)
      for (const auto& e : w.second) {
        dn_rows[e.first].push_back({e.second, w.first * op.coef});
        dn_rows[e.second].push_back({e.first, w.first * op.coef});
      }
    for (const auto& e : op.dn.diag) dn_diag[e.first] += e.second * op.coef;
    for (const auto& w : op.up.groups)
      for (const auto& e : w.second) {
        up_cols[e.first].push_back({e.second, w.first * op.coef});
        up_cols[e.second].push_back({e.first, w.first * op.coef});
      }
    for (const auto& e : op.up.diag) up_diag[e.first] += e.second * op.coef;
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t r = 0; r < (int64_t)dd; ++r) {
    double* yr = y + r * du;
    const double* xr = x + r * du;
    for (uint64_t i = 0; i < du; ++i) {
      double acc = (dn_diag[r] + up_diag[i]) * xr[i];
      for (const auto& h : up_cols[i]) acc += h.second * xr[h.first];
      yr[i] += acc;
    }
    for (const auto& h : dn_rows[r]) {
      const double* xs = x + (uint64_t)h.first * du;
      for (uint64_t i = 0; i < du; ++i) yr[i] += h.second * xs[i];
    }
    for (const Op& op : S.ops)
      if (op.type == 1)
        for (uint64_t i = 0; i < du; ++i) yr[i] += xr[i] * op.elem[r * du + i] * op.coef;
  }
}

void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, bool vec,
                    std::vector<double>& energy, std::vector<double>& evector) {
  int M = (int)a.size();
  energy.clear();
  evector.clear();
  if (M == 1) {
    energy.push_back(a[0]);
    if (vec) evector.push_back(1.0);
    return;
  }
  std::vector<double> d(a), e(b), work(2 * M), z((size_t)M * M);
  e.resize(M);
  char jobz = vec ? 'V' : 'N';
  int info = 0, n = M, ldz = M;
  dstev_(&jobz, &n, d.data(), e.data(), z.data(), &ldz, work.data(), &info);
  if (info != 0) throw std::runtime_error("dstev failed");
  energy.assign(d.begin(), d.end());
  if (vec) evector.assign(z.begin(), z.begin() + M);
}

}  // namespace

extern "C" {

const char* or_last_error() { return g_err.c_str(); }

int or_halfbasis(int L, int N, uint32_t* out) {
  auto b = half_basis(L, N);
  if (out) std::memcpy(out, b.data(), b.size() * sizeof(uint32_t));
  return (int)b.size();
}

// Builds one sector of a model. Operators are given as flattened element lists with
// 0-based spin-orbital indices (spin-down orbital = site + L), already converted from
// the 1-based user labels, i.e. what the reference's ED::new_operator receives.
void* or_sector_new(int L, int nup, int ndn, int n_ops, const int* types, const double* coefs,
                    const int* nels, const int* er, const int* ec, const double* ev) {
  try {
    Sector* S = new Sector;
    S->L = L;
    S->nup = nup;
    S->ndn = ndn;
    S->bup = half_basis(L, nup);
    S->bdn = half_basis(L, ndn);
    S->dim = (uint64_t)S->bup.size() * S->bdn.size();
    int off = 0;
    for (int k = 0; k < n_ops; ++k) {
      Op op;
      op.type = types[k];
      op.coef = coefs[k];
      std::vector<Elem> el;
      for (int i = 0; i < nels[k]; ++i) {
        Elem x{er[off + i], ec[off + i], ev[off + i]};
        if (op.type == 0) {
          if (x.r < x.c) {
            el.push_back(x);
            el.push_back({x.c, x.r, x.v});
          } else if (x.r == x.c) {
            el.push_back(x);
          }
        } else {
          if (x.r > x.c) std::swap(x.r, x.c);
          el.push_back(x);
        }
      }
      off += nels[k];
      if (op.type == 0) {
        op.up = build_half(el, S->bup, L, false);
        op.dn = build_half(el, S->bdn, L, true);
      } else {
        op.elem.assign(S->dim, 0.0);
        const uint64_t du = S->bup.size();
        for (uint64_t I = 0; I < S->dim; ++I) {
          // factorized basis word: up state in the high word, down state in the low word
          uint64_t b = (uint64_t)S->bdn[I / du] + ((uint64_t)S->bup[I % du] << 32);
          double val = 0.0;
          for (const Elem& x : el) {
            uint64_t m = mask64(x.r, L) + mask64(x.c, L);
            if ((b & m) == m) val += x.v;
          }
          op.elem[I] = val;
        }
      }
      S->ops.push_back(std::move(op));
    }
    return S;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void or_sector_free(void* h) { delete (Sector*)h; }
uint64_t or_sector_dim(void* h) { return ((Sector*)h)->dim; }

void or_hv(void* h, const double* x, double* y) { hv_all(*(Sector*)h, x, y); }
void or_hv_omp(void* h, const double* x, double* y, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  hv_all_omp(*(Sector*)h, x, y);
}

// Fills x with the reference's start vector: zeros + N(0,1) deviates from a fresh
// default_random_engine(seed), then normalized.
void or_start_vector(uint64_t n, int seed, double* x) {
  std::default_random_engine gen((unsigned)seed);
  for (uint64_t i = 0; i < n; ++i) x[i] = 0.0 + g_normal(gen);
  double z = cblas_dnrm2((int)n, x, 1);
  if (z >= 1.0e-12) cblas_dscal((int)n, 1.0 / z, x, 1);
}

void or_reset_normal() { g_normal.reset(); }

// Ground state by Lanczos: eigenvalue pass, then eigenvector replay.
// gs: in = start vector (normalized), out = ground state. Returns 0 or -1 on error.
// nthreads > 0: products by the OpenMP-over-rows form (the same terms grouped by output
// row, rows split over threads); want_vec = 0 skips the eigenvector replay.
static int lanczos_gs_impl(void* h, double accur, int max_iter, double* gs, double* E0, int* niter,
                           double* alpha_out, double* beta_out, int nthreads, int want_vec);

int or_lanczos_gs(void* h, double accur, int max_iter, double* gs, double* E0, int* niter,
                  double* alpha_out, double* beta_out) {
  return lanczos_gs_impl(h, accur, max_iter, gs, E0, niter, alpha_out, beta_out, 0, 1);
}

int or_lanczos_gs_threads(void* h, double accur, int max_iter, double* gs, double* E0, int* niter,
                          double* alpha_out, double* beta_out, int nthreads, int want_vec) {
  return lanczos_gs_impl(h, accur, max_iter, gs, E0, niter, alpha_out, beta_out, nthreads, want_vec);
}

// `steps` iterations of the ground-state recurrence of src_ed/Hamiltonian/Lanczos.hpp:58-90
// (rescale, product, alpha, axpy, beta) from a normalized psi, without the stopping test:
// a fixed amount of work for timing. nthreads 0 takes the serial product.
int or_lanczos_steps(void* h, const double* psi, int steps, int nthreads, double* alpha_out,
                     double* beta_out) {
  Sector& S = *(Sector*)h;
  const int n = (int)S.dim;
  std::vector<double> r(psi, psi + n), q(n, 0.0);
  double bprev = 0.0;
  for (int k = 0; k < steps; ++k) {
    if (k > 0) {
      cblas_dscal(n, 1.0 / bprev, r.data(), 1);
      cblas_dscal(n, -bprev, q.data(), 1);
    }
    if (nthreads > 0) {
      omp_set_num_threads(nthreads);
      hv_all_omp(S, r.data(), q.data());
    } else {
      hv_all(S, r.data(), q.data());
    }
    std::swap(r, q);
    const double a = cblas_ddot(n, q.data(), 1, r.data(), 1);
    cblas_daxpy(n, -a, q.data(), 1, r.data(), 1);
    bprev = cblas_dnrm2(n, r.data(), 1);
    if (alpha_out) alpha_out[k] = a;
    if (beta_out) beta_out[k] = bprev;
  }
  return 0;
}

}  // extern "C"

static int lanczos_gs_impl(void* h, double accur, int max_iter, double* gs, double* E0, int* niter,
                           double* alpha_out, double* beta_out, int nthreads, int want_vec) {
  Sector& S = *(Sector*)h;
  auto hv = [&](const double* x, double* y) {
    if (nthreads > 0) {
      omp_set_num_threads(nthreads);
      hv_all_omp(S, x, y);
    } else {
      hv_all(S, x, y);
    }
  };
  try {
    const int n = (int)S.dim;
    std::vector<double> r(gs, gs + n), q(n, 0.0), alpha, beta, energy, evector;
    double ritz = 2.0 * accur;
    while (ritz > accur) {
      if (!beta.empty()) {
        cblas_dscal(n, 1.0 / beta.back(), r.data(), 1);
        cblas_dscal(n, -beta.back(), q.data(), 1);
      }
      hv(r.data(), q.data());
      std::swap(r, q);
      double z = cblas_ddot(n, q.data(), 1, r.data(), 1);
      alpha.push_back(z);
      cblas_daxpy(n, -alpha.back(), q.data(), 1, r.data(), 1);
      beta.push_back(cblas_dnrm2(n, r.data(), 1));
      tridiag_lowest(alpha, beta, true, energy, evector);
      ritz = std::fabs(evector.back() * beta.back());
      if ((int)alpha.size() > max_iter) {
        g_err = "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations";
        return -1;
      }
    }
    *E0 = energy[0];
    *niter = (int)beta.size();
    for (size_t i = 0; i < alpha.size(); ++i) {
      if (alpha_out) alpha_out[i] = alpha[i];
      if (beta_out) beta_out[i] = beta[i];
    }
    if (!want_vec) return 0;
    // eigenvector replay
    std::vector<double> r2(gs, gs + n), q2(n, 0.0);
    tridiag_lowest(alpha, beta, true, energy, evector);
    cblas_dscal(n, evector[0], gs, 1);
    for (size_t j = 1; j < evector.size(); ++j) {
      hv(r2.data(), q2.data());
      cblas_daxpy(n, -alpha[j - 1], r2.data(), 1, q2.data(), 1);
      for (int i = 0; i < n; ++i) {
        double t = r2[i];
        r2[i] = q2[i] / beta[j - 1];
        q2[i] = -beta[j - 1] * t;
      }
      cblas_daxpy(n, evector[j], r2.data(), 1, gs, 1);
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

extern "C" {

// Lanczos tridiagonalization from a normalized psi (continued-fraction Green function).
// a_out/b_out must hold max_iter+1 entries. Returns the number of a coefficients.
int or_lanczos_green(void* h, const double* psi_in, double accur_cf, int max_iter, double* a_out,
                     double* b_out) {
  Sector& S = *(Sector*)h;
  const int n = (int)S.dim;
  std::vector<double> psi(psi_in, psi_in + n), q(n, 0.0), a, b;
  b.push_back(1.0);
  double prod = 1.0;
  while (prod > accur_cf && std::fabs(b.back()) > accur_cf && (int)a.size() < max_iter) {
    if (!a.empty()) {
      double bb = b.back();
      for (int i = 0; i < n; ++i) {
        double t = psi[i];
        psi[i] = q[i] / bb;
        q[i] = -bb * t;
      }
    }
    hv_all(S, psi.data(), q.data());
    double z = cblas_ddot(n, psi.data(), 1, q.data(), 1);
    a.push_back(z);
    cblas_daxpy(n, -a.back(), psi.data(), 1, q.data(), 1);
    double zz = cblas_dnrm2(n, q.data(), 1);
    b.push_back(zz);
    prod *= std::fabs(zz);
  }
  for (size_t i = 0; i < a.size(); ++i) a_out[i] = a[i];
  for (size_t i = 0; i < b.size(); ++i) b_out[i] = b[i];
  return (int)a.size();
}

// y += z * c_a x (pm = -1) or y += z * c†_a x (pm = +1) between factorized sectors.
// a is a spin-orbital index (down orbital = site + L). (nup, ndn) describe x's sector.
int or_create_destroy(int L, int nup, int ndn, int a, int pm, const double* x, double* y,
                      double z) {
  int tup = nup, tdn = ndn;
  bool down = a >= L;
  int site = down ? a - L : a;
  if (down) tdn += pm; else tup += pm;
  auto Bu = half_basis(L, nup), Bd = half_basis(L, ndn);
  auto Tu = half_basis(L, tup), Td = half_basis(L, tdn);
  if (Tu.empty() || Td.empty()) return 0;
  // Destroy() on a state of basis (su, sd) with dims (du): returns target label and parity
  auto destroy = [&](uint32_t label, const std::vector<uint32_t>& fu, const std::vector<uint32_t>& fd,
                     const std::vector<uint32_t>& gu, const std::vector<uint32_t>& gd,
                     uint32_t* out, int* par) -> bool {
    uint32_t du = (uint32_t)fu.size();
    uint32_t lu = label % du, ld = label / du;
    uint32_t su = fu[lu], sd = fd[ld];
    uint32_t m = 1u << site;
    if (down) {
      if (!(sd & m)) return false;
      *par = (popcount32(sd & (m - 1)) + popcount32(su)) & 1;
      *out = half_index(gd, sd ^ m) * (uint32_t)gu.size() + lu;
    } else {
      if (!(su & m)) return false;
      *par = popcount32(su & (m - 1)) & 1;
      *out = ld * (uint32_t)gu.size() + half_index(gu, su ^ m);
    }
    return true;
  };
  uint64_t dsrc = (uint64_t)Bu.size() * Bd.size(), dtgt = (uint64_t)Tu.size() * Td.size();
  if (pm == 1) {
    for (uint64_t I = 0; I < dtgt; ++I) {
      uint32_t J;
      int par;
      if (!destroy((uint32_t)I, Tu, Td, Bu, Bd, &J, &par)) continue;
      y[I] += z * (par ? -1.0 : 1.0) * x[J];
    }
  } else {
    for (uint64_t I = 0; I < dsrc; ++I) {
      uint32_t J;
      int par;
      if (!destroy((uint32_t)I, Bu, Bd, Tu, Td, &J, &par)) continue;
      y[J] += z * (par ? -1.0 : 1.0) * x[I];
    }
  }
  return 1;
}

// Symmetric tridiagonal eigen-decomposition through LAPACK dstev, exported so tests
// can pin the product's own tridiagonal solver against the routine the reference calls.
int or_dstev(int n, double* d, double* e, double* z) {
  std::vector<double> work(2 * n + 2);
  char jobz = z ? 'V' : 'N';
  int info = 0, ldz = n > 0 ? n : 1;
  dstev_(&jobz, &n, d, e, z, &ldz, work.data(), &info);
  return info;
}

}  // extern "C"
