// The ED solver interface (the reference's namespace ED, src_ed/qcm_ED.hpp:23-340) on top of
// the device sector API: model registry, instances, the sector loop and Boltzmann weights,
// Lehmann (band Lanczos / full diagonalization) and continued-fraction Green functions,
// self-energy, hybridization function and the Potthoff functional. Trivial symmetry group,
// normal or up/down mixing, real Hilbert space.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qed.h"
#include "qed_h5.hpp"
#include "qed_linalg.hpp"
#include "qed_sector.hpp"

namespace qed {
namespace ed {

using cd = std::complex<double>;

// ---- global parameters (src_util/global_parameter.cpp:12-91 defaults) ---------------
struct Globals {
  std::map<std::string, int> i{{"seed", 0},           {"max_iter_lanczos", 1000}, {"max_iter_BL", 600},
                               {"max_iter_CF", 400},  {"max_dim_full", 256},      {"n_states", 1},
                               {"kgrid_side", 32},    {"GF_lookup_depth", 64},   {"max_dim_print", 64}};
  std::map<std::string, double> d{{"accur_lanczos", 1e-12},     {"accur_band_lanczos", 1e-12},
                                  {"accur_deflation", 1e-7},    {"accur_Q_matrix", 1e-5},
                                  {"Qmatrix_wtol", 1e-5},       {"Qmatrix_vtol", 1e-10},
                                  {"temperature", 0.0},         {"minimum_weight", 0.01},
                                  {"accur_Davidson", 1e-5},     {"accur_continued_fraction", 0.01},
                                  {"band_lanczos_minimum_gap", 1e-5}};
  std::map<std::string, bool> b{{"verb_ED", false}, {"no_degenerate_BL", false}, {"continued_fraction", false},
                              {"block_Lanczos_QR", false}, {"combine_mcf", false}};
  std::map<std::string, char> c{{"GF_method", 'L'}, {"Ground_state_method", 'L'}, {"Hamiltonian_format", 'E'}};
};
Globals& G() {
  static Globals g;
  return g;
}
int gi(const char* n) { return G().i.at(n); }
double gd(const char* n) { return G().d.at(n); }
char gc(const char* n) { return G().c.at(n); }
bool gb(const char* n) { return G().b.at(n); }

// band Lanczos refuses a lowest pair closer than band_lanczos_minimum_gap when
// no_degenerate_BL is set (src_ed/Hamiltonian/Lanczos.hpp:235-236,368-370); -1: no guard
double degeneracy_gap() { return gb("no_degenerate_BL") ? gd("band_lanczos_minimum_gap") : -1.0; }
}  // namespace ed

bool verbose_ed() { return ed::gb("verb_ED"); }

namespace ed {

constexpr double SMALL_VALUE = 1.0e-12;
constexpr double MIDDLE_VALUE = 1.0e-8;
constexpr double MIN_GAP = 1.0e-6;

struct Op {
  std::string name, type;
  std::vector<int> r, c;
  std::vector<double> v;
  int target = 1;  // 1 cluster, 2 bath, 3 cluster-bath (src_ed/Operators/Hermitian_operator.hpp)
  bool interaction = false;
  bool anomalous = false;
  // the mixing the operator needs: 1 anomalous, 2 spin flip (one-body elements between the
  // spins, or same-spin pairing), as one_body_operator::check_spin_flip and
  // anomalous_operator::check_spin_flip set it (one_body_operator.hpp:198-206,
  // anomalous_operator.hpp:113-121)
  int mixing = 0;
};

// the reference's codes for a particle number or spin that is not conserved
// (src_ed/sector.hpp:24-25)
constexpr int kSecOdd = 9999, kSecEven = 10000;
bool sec_code(int x) { return x >= kSecOdd; }

struct Model {
  std::string name;
  int n_sites = 0, n_bath = 0, L = 0;
  std::vector<Op> ops;
  std::map<std::string, size_t> idx;
  int mixing = -1;
  bool closed = false;
  // point group (src_ed/symmetry_group.hpp): generators as given to the library (0-based
  // site images, then bath entries), the group's order, elements and characters, and the
  // symmetric orbitals of the sites (rows of sorb, n_sites x n_sites, and their irreps)
  std::vector<std::vector<int>> gens;
  bool bath_irrep = false;
  GroupTables group;
  std::vector<double> sorb;
  std::vector<int> sorb_irrep;
  bool symmetric() const { return group.g > 1; }
};

struct QMat {
  int L = 0;
  std::vector<double> e;  // M
  std::vector<double> Q;  // L x M column-major
  int M() const { return (int)e.size(); }
};

struct CFrac {
  std::vector<double> a, b;  // after the continued_fraction transform
  bool set = false;
};

// matrix continued fraction of one (spin, hole/particle) part: M floors of pa x pa blocks
// (energy-shifted A, B) and the pa x q weight W (sqrt(weight) folded in)
struct MCFrac {
  int pa = 0, q = 0;
  std::vector<double> A, B, W;
  int M() const { return pa ? (int)(A.size() / ((size_t)pa * pa)) : 0; }
};

// continued fractions of one block of p orbitals (an irrep of the symmetric orbitals, or
// the whole spinor): orbital k = sum_a W[k d + a] (component a of the Green function)
struct CFBlock {
  int p = 0;
  std::vector<double> W;     // p x d, row-major
  std::vector<CFrac> e, h;   // p x p (o1 >= o2 at o1 + o2 p), electron / hole parts
};

struct State {
  int nup = 0, ndn = 0, irrep = 0;
  int secN = 0, secS = 0;  // the sector (N, S; codes where not conserved)
  double energy = 0.0, weight = 1.0;
  DevBuf psi;
  QMat gf[2];
  std::vector<QMat> gfb[2];           // per irrep of the symmetric orbitals (their basis)
  std::vector<CFrac> cfe[2], cfh[2];  // n x n (o1 >= o2), electron / hole parts
  std::vector<CFBlock> cfb[2];        // symmetric orbitals or spinor: fractions per block
  MCFrac mcf[2][2];                   // [spin][0: holes, 1: particles]
  std::vector<MCFrac> mcfb[2][2];     // per irrep of the symmetric orbitals, W in the site basis
  MCFrac mcf_comb[2];                 // combined fraction (combine_mcf), evaluated alone
};

struct Instance {
  uint64_t label = 0;
  Model* model = nullptr;
  std::map<std::string, double> value;
  std::vector<std::pair<int, int>> sectors;  // (N, S); kSecOdd / kSecEven where not conserved
  std::vector<int> irreps;                   // R of each target sector
  int mixing = 0;
  bool correlated = false, gs_solved = false, gf_solved = false;
  bool cf = false;   // continued fractions ('F')
  bool mcf = false;  // matrix continued fractions ('M')
  bool q_to_mcf = false;  // Lehmann form converted to one matrix continued fraction
  bool from_file = false;  // states read from an instance file: Green functions only, no vectors
  std::vector<std::unique_ptr<State>> states;
  double GS_energy = 0.0, E0 = 0.0, SEF_bath = 0.0;
  // one-body matrices (index 0 up, 1 down), bath diagonalized
  std::vector<double> tc[2], tb[2], tcb[2];  // tc n x n, tb nb x nb, tcb n x nb (column-major)
  std::vector<double> eb[2], tcb_d[2];       // diagonal bath energies and rotated hybridization
  std::string gs_string;
  // recently evaluated G(z) per spin, most recent first (model_instance::look_up_table)
  std::deque<std::pair<std::complex<double>, std::vector<double>>> gf_cache[2];
};

std::map<std::string, std::unique_ptr<Model>>& models() {
  static std::map<std::string, std::unique_ptr<Model>> m;
  return m;
}
std::map<uint64_t, std::unique_ptr<Instance>>& instances() {
  static std::map<uint64_t, std::unique_ptr<Instance>> m;
  return m;
}

Instance& inst(uint64_t label) {
  auto it = instances().find(label);
  if (it == instances().end())
    fail("The cluster instance label " + std::to_string(label) + " is out of range.");
  return *it->second;
}

// spinor components per site: 2 with anomalous terms (Nambu), 2 with spin flip, 4 with both
int n_mixed(int mixing) { return ((mixing & 1) ? 2 : 1) * ((mixing & 2) ? 2 : 1); }
bool spinor_mixing(int mixing) { return (mixing & 3) != 0; }
// dimension of the Green function (model_instance_base.cpp:75-80)
int dim_gf(const Instance& I) { return I.model->n_sites * n_mixed(I.mixing); }

// a function that the spin-flip and anomalous mixings do not reach
void require_no_spinor(const Instance& I, const char* what) {
  if (spinor_mixing(I.mixing))
    fail(std::string(what) + " is not available with Hilbert space mixing " + std::to_string(I.mixing));
}

// the (N_up, N_down) blocks of a sector (ED_mixed_basis, src_ed/ED_basis.cpp:185-361)
std::vector<std::pair<int, int>> sector_blocks(int L, int N, int S) {
  std::vector<std::pair<int, int>> out;
  if (!sec_code(N) && !sec_code(S)) {
    out.push_back({(N + S) / 2, (N - S) / 2});
  } else if (sec_code(N) && !sec_code(S)) {
    for (int nu = 0; nu <= L; ++nu)
      if (nu - S >= 0 && nu - S <= L) out.push_back({nu, nu - S});
  } else if (!sec_code(N) && sec_code(S)) {
    for (int nu = 0; nu <= L; ++nu)
      if (N - nu >= 0 && N - nu <= L) out.push_back({nu, N - nu});
  } else {
    const int par = N == kSecOdd ? 1 : 0;
    for (int nu = 0; nu <= L; ++nu)
      for (int nd = 0; nd <= L; ++nd)
        if ((nu + nd) % 2 == par) out.push_back({nu, nd});
  }
  return out;
}

struct SectorDeleter {
  void operator()(qed_sector* s) const { qed_sector_destroy(s); }
};
using SectorPtr = std::unique_ptr<qed_sector, SectorDeleter>;

// flattened element lists of the active operators (value != 0), insertion order
void flatten(const Instance& I, std::vector<int>& types, std::vector<double>& coefs, std::vector<int>& nels,
             std::vector<int>& er, std::vector<int>& ec, std::vector<double>& ev, const std::string* only = nullptr) {
  for (const Op& op : I.model->ops) {
    double z;
    if (only) {
      if (op.name != *only) continue;
      z = 1.0;
    } else {
      auto it = I.value.find(op.name);
      if (it == I.value.end() || it->second == 0.0) continue;
      z = it->second;
    }
    types.push_back(op.interaction ? 1 : (op.anomalous ? 2 : 0));
    coefs.push_back(z);
    nels.push_back((int)op.v.size());
    er.insert(er.end(), op.r.begin(), op.r.end());
    ec.insert(ec.end(), op.c.begin(), op.c.end());
    ev.insert(ev.end(), op.v.begin(), op.v.end());
  }
}

// the sector (nup, ndn) of the instance's Hamiltonian (or of one operator): factorized for
// a trivial group, else the symmetry-adapted sector of irreducible representation irrep
SectorPtr make_sector(const Instance& I, int nup, int ndn, const std::string* only = nullptr, int irrep = 0) {
  std::vector<int> types, nels, er, ec;
  std::vector<double> coefs, ev;
  flatten(I, types, coefs, nels, er, ec, ev, only);
  qed_sector* s = nullptr;
  const Model& M = *I.model;
  int rc;
  if (M.symmetric()) {
    std::vector<int> gflat;
    for (const auto& gen : M.gens) gflat.insert(gflat.end(), gen.begin(), gen.end());
    rc = qed_mixed_sector_create(M.L, nup, ndn, (int)types.size(), types.data(), coefs.data(), nels.data(),
                                 er.data(), ec.data(), ev.data(), M.n_sites, (int)M.gens.size(), gflat.data(),
                                 M.bath_irrep ? 1 : 0, irrep, &s);
  } else {
    if (irrep != 0) fail("irreducible representation " + std::to_string(irrep) + " of the trivial group");
    rc = qed_sector_create(M.L, nup, ndn, (int)types.size(), types.data(), coefs.data(), nels.data(), er.data(),
                           ec.data(), ev.data(), &s);
  }
  if (rc != 0) fail(qed_last_error());
  return SectorPtr(s);
}

// a sector (N, S) of the spin-flip / anomalous mixings (codes where not conserved)
SectorPtr make_general_sector(const Instance& I, int N, int S, const std::string* only = nullptr) {
  std::vector<int> types, nels, er, ec;
  std::vector<double> coefs, ev;
  flatten(I, types, coefs, nels, er, ec, ev, only);
  const auto blocks = sector_blocks(I.model->L, N, S);
  std::vector<int> pairs;
  for (auto& b : blocks) {
    pairs.push_back(b.first);
    pairs.push_back(b.second);
  }
  qed_sector* s = nullptr;
  if (qed_general_sector_create(I.model->L, (int)blocks.size(), pairs.data(), (int)types.size(), types.data(),
                                coefs.data(), nels.data(), er.data(), ec.data(), ev.data(), &s) != 0)
    fail(qed_last_error());
  return SectorPtr(s);
}

// the sector of a state (factorized, symmetry-adapted or mixed-number)
SectorPtr state_sector(const Instance& I, const State& st, const std::string* only = nullptr) {
  if (spinor_mixing(I.mixing)) return make_general_sector(I, st.secN, st.secS, only);
  return make_sector(I, st.nup, st.ndn, only, st.irrep);
}

// a function of the trivial-group solver only
void require_trivial_group(const Instance& I, const char* what) {
  if (I.model->symmetric())
    fail(std::string(what) + " is not available for a model with a nontrivial point group (model " +
         I.model->name + ")");
}

// diagonalized bath of spin index sp: eb, and tcb rotated into the bath's eigenbasis
// (nc cluster rows, nbg bath orbitals)
void diagonalize_bath(Instance& I, int sp, int nc, int nbg) {
  I.eb[sp].assign(nbg, 0.0);
  I.tcb_d[sp] = I.tcb[sp];
  if (nbg == 0) return;
  std::vector<double> a(I.tb[sp]), w(nbg);
  bool diag = true;
  for (int i = 0; i < nbg && diag; ++i)
    for (int j = 0; j < nbg; ++j)
      if (i != j && a[i + (size_t)j * nbg] != 0.0) {
        diag = false;
        break;
      }
  if (diag) {
    for (int i = 0; i < nbg; ++i) I.eb[sp][i] = a[i + (size_t)i * nbg];
    return;
  }
  sym_eigen(nbg, a.data(), w.data(), true);
  I.eb[sp] = w;
  std::vector<double> t((size_t)nc * nbg, 0.0);
  for (int i = 0; i < nc; ++i)
    for (int m = 0; m < nbg; ++m) {
      double acc = 0.0;
      for (int k = 0; k < nbg; ++k) acc += I.tcb[sp][i + (size_t)k * nc] * a[k + (size_t)m * nbg];
      t[i + (size_t)m * nc] = acc;
    }
  I.tcb_d[sp] = t;
}

// One-body matrices of the spinor for the spin-flip, anomalous and full mixings
// (one_body_operator::set_hopping_matrix_templ, src_ed/Operators/one_body_operator.hpp:
// 326-411; anomalous_operator::set_hopping_matrix_templ, anomalous_operator.hpp:241-284):
// with M the operator's matrix over the 2L spin orbitals (one-body: (r, c, v), (c, r, v);
// anomalous: (r, c, v), (c, r, -v)) and e = [[M_uu, M_ud], [M_du, M_dd]] restricted to
// the operator's rows and columns (sites, bath or sites x bath):
//   spin flip: one-body e;
//   anomalous: one-body diag(M_uu, -M_dd^T) (hybrid: -M_dd), pairing 2 M_du below and
//     2 M_du^T above the diagonal blocks (hybrid: -2 M_ud above);
//   full: one-body diag(e, -e^T) (hybrid: -e), pairing 2 e below, 2 e^T above (hybrid: -2 e).
// Adds z times the spinor matrices of one operator to tc, tb and tcb (column-major,
// dimensions nm n_sites and nm n_bath) by the rules above.
void add_spinor_op(const Model& M, const Op& op, double z, int m, std::vector<double>& tc, std::vector<double>& tb,
                   std::vector<double>& tcb) {
  const int ns = M.n_sites, nb = M.n_bath, L = M.L, nm = n_mixed(m);
  const int dg = nm * ns, dbg = nm * nb;
  {
    std::vector<double> F((size_t)4 * L * L, 0.0);  // M over spin orbitals, 2L x 2L column-major
    auto at = [&](int a, int b) -> double& { return F[a + (size_t)b * 2 * L]; };
    for (size_t q = 0; q < op.v.size(); ++q) {
      const int r = op.r[q], c = op.c[q];
      const double v = op.v[q];
      if (op.anomalous) {
        if (r < c) {
          at(r, c) += v;
          at(c, r) -= v;
        }
      } else if (r < c) {
        at(r, c) += v;
        at(c, r) += v;
      } else if (r == c) {
        at(r, r) += v;
      }
    }
    // the operator's rows and columns (orbitals within a spin)
    int r0, nR, c0, nC;
    std::vector<double>* dst;
    int ld;  // leading dimension of the destination (its row count)
    if (op.target == 1) {
      r0 = 0, nR = ns, c0 = 0, nC = ns, dst = &tc, ld = dg;
    } else if (op.target == 2) {
      r0 = ns, nR = nb, c0 = ns, nC = nb, dst = &tb, ld = dbg;
    } else {
      r0 = 0, nR = ns, c0 = ns, nC = nb, dst = &tcb, ld = dg;
    }
    const bool hyb = op.target == 3;
    // e: the (2 nR) x (2 nC) spin block matrix of the operator
    auto e = [&](int a, int b) {
      const int ra = a < nR ? r0 + a : L + r0 + (a - nR);
      const int cb = b < nC ? c0 + b : L + c0 + (b - nC);
      return at(ra, cb);
    };
    auto put = [&](int a, int b, double val) { (*dst)[a + (size_t)b * ld] += z * val; };
    if (!op.anomalous) {
      if (m == 2) {
        for (int a = 0; a < 2 * nR; ++a)
          for (int b = 0; b < 2 * nC; ++b) put(a, b, e(a, b));
      } else if (m == 1) {
        for (int a = 0; a < nR; ++a)
          for (int b = 0; b < nC; ++b) {
            put(a, b, e(a, b));
            // lower-right: -M_dd^T (square blocks), -M_dd for the hybridization
            if (hyb) put(nR + a, nC + b, -e(nR + a, nC + b));
          }
        if (!hyb)
          for (int a = 0; a < nR; ++a)
            for (int b = 0; b < nC; ++b) put(nR + a, nC + b, -e(nR + b, nC + a));
      } else {
        for (int a = 0; a < 2 * nR; ++a)
          for (int b = 0; b < 2 * nC; ++b) {
            put(a, b, e(a, b));
            put(2 * nR + a, 2 * nC + b, hyb ? -e(a, b) : -e(b, a));
          }
      }
    } else {
      if (m == 1) {
        for (int a = 0; a < nR; ++a)
          for (int b = 0; b < nC; ++b) {
            put(nR + a, b, 2.0 * e(nR + a, b));                            // 2 M_du
            put(a, nC + b, hyb ? -2.0 * e(a, nC + b) : 2.0 * e(nR + b, a));  // 2 M_du^T or -2 M_ud
          }
      } else if (m == 3) {
        for (int a = 0; a < 2 * nR; ++a)
          for (int b = 0; b < 2 * nC; ++b) {
            put(2 * nR + a, b, 2.0 * e(a, b));
            put(a, 2 * nC + b, hyb ? -2.0 * e(a, b) : 2.0 * e(b, a));
          }
      } else {
        fail("anomalous operator " + op.name + " in an instance without anomalous mixing");
      }
    }
  }
}

void set_hopping_matrix_spinor(Instance& I) {
  Model& M = *I.model;
  const int ns = M.n_sites, nb = M.n_bath, m = I.mixing, nm = n_mixed(m);
  const int dg = nm * ns, dbg = nm * nb;
  I.SEF_bath = 0.0;
  for (int sp = 0; sp < 2; ++sp) {
    I.tc[sp].assign((size_t)dg * dg, 0.0);
    I.tb[sp].assign((size_t)dbg * dbg, 0.0);
    I.tcb[sp].assign((size_t)dg * dbg, 0.0);
  }
  for (const Op& op : M.ops) {
    if (op.interaction) continue;
    auto it = I.value.find(op.name);
    if (it == I.value.end() || it->second == 0.0) continue;
    add_spinor_op(M, op, it->second, m, I.tc[0], I.tb[0], I.tcb[0]);
  }
  diagonalize_bath(I, 0, dg, dbg);
  I.eb[1] = I.eb[0];
  I.tcb_d[1] = I.tcb_d[0];
  // the bath's part of the Potthoff functional: minus the negative ones among the n_bath
  // lowest energies of the spinor bath, no spin factor (model_instance::set_hopping_matrix,
  // src_ed/model_instance.hpp:469-473, which reads the first n_bath diagonal entries: the
  // lowest when the bath was diagonalized; a bath matrix that is already diagonal is
  // sorted here too, since its stored order lists the particle block first and the
  // Nambu functional then misses the hole block's negative energies)
  std::vector<double> e = I.eb[0];
  std::sort(e.begin(), e.end());
  for (int i = 0; i < nb; ++i)
    if (e[i] < 0) I.SEF_bath -= e[i];
}

// ---- one-body matrices (model_instance::set_hopping_matrix, model_instance.hpp:424-475) --
void set_hopping_matrix(Instance& I) {
  if (spinor_mixing(I.mixing)) {
    set_hopping_matrix_spinor(I);
    return;
  }
  Model& M = *I.model;
  const int ns = M.n_sites, nb = M.n_bath, L = M.L;
  I.SEF_bath = 0.0;
  const int nspin = (I.mixing & 4) ? 2 : 1;
  for (int sp = 0; sp < 2; ++sp) {
    I.tc[sp].assign((size_t)ns * ns, 0.0);
    I.tb[sp].assign((size_t)nb * nb, 0.0);
    I.tcb[sp].assign((size_t)ns * nb, 0.0);
  }
  for (const Op& op : M.ops) {
    if (op.interaction) continue;
    auto it = I.value.find(op.name);
    if (it == I.value.end() || it->second == 0.0) continue;
    // spin-flip and anomalous terms belong to the spinor matrices (mixing 1, 2, 3)
    if (op.anomalous || (op.mixing & 2))
      fail("operator " + op.name + " needs the spin-flip or anomalous mixing");
    const double z = it->second;
    for (size_t q = 0; q < op.v.size(); ++q) {
      int r = op.r[q], c = op.c[q];
      double v = op.v[q] * z;
      if (r > c) continue;
      if ((r >= L) != (c >= L)) fail("operator " + op.name + " couples the two spins");
      int sp = r >= L ? 1 : 0;
      int rr = r - sp * L, cc = c - sp * L;
      auto add = [&](int a, int b, double val) {
        if (a < ns && b < ns) I.tc[sp][a + (size_t)b * ns] += val;
        else if (a >= ns && b >= ns) I.tb[sp][(a - ns) + (size_t)(b - ns) * nb] += val;
        else if (a < ns && b >= ns) I.tcb[sp][a + (size_t)(b - ns) * ns] += val;
      };
      add(rr, cc, v);
      if (rr != cc) add(cc, rr, v);
    }
  }
  for (int sp = 0; sp < 2; ++sp) {
    I.eb[sp].assign(nb, 0.0);
    I.tcb_d[sp] = I.tcb[sp];
    if (nb == 0) continue;
    std::vector<double> a(I.tb[sp]), w(nb);
    bool diag = true;
    for (int i = 0; i < nb && diag; ++i)
      for (int j = 0; j < nb; ++j)
        if (i != j && a[i + (size_t)j * nb] != 0.0) {
          diag = false;
          break;
        }
    if (diag) {
      for (int i = 0; i < nb; ++i) I.eb[sp][i] = a[i + (size_t)i * nb];
    } else {
      sym_eigen(nb, a.data(), w.data(), true);
      I.eb[sp] = w;
      std::vector<double> t((size_t)ns * nb, 0.0);
      for (int i = 0; i < ns; ++i)
        for (int m = 0; m < nb; ++m) {
          double acc = 0.0;
          for (int k = 0; k < nb; ++k) acc += I.tcb[sp][i + (size_t)k * ns] * a[k + (size_t)m * nb];
          t[i + (size_t)m * ns] = acc;
        }
      I.tcb_d[sp] = t;
    }
    if (sp < nspin)
      for (int i = 0; i < nb; ++i)
        if (I.eb[sp][i] < 0) I.SEF_bath -= I.eb[sp][i];
  }
  if (I.mixing == 0) I.SEF_bath *= 2.0;
}

// ---- Q-matrix post-processing (src_ed/Q_matrix.hpp) ------------------------------------
QMat streamline(const QMat& q) {
  const double wtol = gd("Qmatrix_wtol"), vtol = gd("Qmatrix_vtol");
  if (vtol < 1e-12) return q;
  QMat out;
  out.L = q.L;
  const int L = q.L, M = q.M();
  for (int i = 0; i < M;) {
    int j = i + 1;
    while (j < M && std::fabs(q.e[j] - q.e[i]) <= wtol) ++j;
    const int n = j - i;
    if (n == 1) {
      double s = 0.0;
      for (int a = 0; a < L; ++a) s += q.Q[a + (size_t)i * L] * q.Q[a + (size_t)i * L];
      if (s >= vtol) {
        out.e.push_back(q.e[i]);
        out.Q.insert(out.Q.end(), q.Q.begin() + (size_t)i * L, q.Q.begin() + (size_t)(i + 1) * L);
      }
    } else {
      std::vector<double> QQ((size_t)L * L, 0.0), d(L);
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
          double s = 0.0;
          for (int k = i; k < j; ++k) s += q.Q[a + (size_t)k * L] * q.Q[b + (size_t)k * L];
          QQ[a + (size_t)b * L] = s;
        }
      sym_eigen(L, QQ.data(), d.data(), true);
      for (int k = 0; k < L; ++k)
        if (d[k] > vtol) {
          out.e.push_back(q.e[i]);
          double sq = std::sqrt(d[k]);
          for (int a = 0; a < L; ++a) out.Q.push_back(QQ[a + (size_t)k * L] * sq);
        }
    }
    i = j;
  }
  return out;
}

void check_norm(const QMat& q, double threshold) {
  const int L = q.L, M = q.M();
  double ztot = 0.0;
  for (int i = 0; i < L; ++i) {
    double z = 0.0;
    for (int k = 0; k < M; ++k) z += q.Q[i + (size_t)k * L] * q.Q[i + (size_t)k * L];
    ztot += (z - 1.0) * (z - 1.0);
    for (int j = 0; j < i; ++j) {
      double zz = 0.0;
      for (int k = 0; k < M; ++k) zz += q.Q[i + (size_t)k * L] * q.Q[j + (size_t)k * L];
      ztot += 2.0 * zz * zz;
    }
  }
  if (ztot > threshold) fail("Q-matrix does not satisfy unitary condition by " + std::to_string(ztot));
}

void append(QMat& a, const QMat& b) {
  if (b.M() == 0) return;
  if (a.M() == 0) {
    a = b;
    return;
  }
  a.e.insert(a.e.end(), b.e.begin(), b.e.end());
  a.Q.insert(a.Q.end(), b.Q.begin(), b.Q.end());
}

// Lehmann form of one sector from the start vectors phi (p x D on the device)
QMat sector_qmatrix(qed_sector& T, const double* phi, int p) {
  QMat q;
  q.L = p;
  const uint64_t D = T.D;
  if (D < (uint64_t)gi("max_dim_full")) {
    // full diagonalization (Hamiltonian_Dense::build_Q_matrix, Hamiltonian_Dense.hpp:157-176)
    const int n = (int)D;
    std::vector<double> H((size_t)n * n), e(n), ph((size_t)p * n);
    DevBuf ex(D * 8), hy(D * 8);
    for (int c = 0; c < n; ++c) {
      QED_HIP(hipMemsetAsync(ex.p, 0, D * 8, stream()));
      QED_HIP(hipMemsetAsync(hy.p, 0, D * 8, stream()));
      const double one = 1.0;
      QED_HIP(hipMemcpyAsync(ex.as<double>() + c, &one, 8, hipMemcpyHostToDevice, stream()));
      hv(T, ex.as<double>(), hy.as<double>());
      QED_HIP(hipMemcpyAsync(&H[(size_t)c * n], hy.p, D * 8, hipMemcpyDeviceToHost, stream()));
    }
    QED_HIP(hipMemcpyAsync(ph.data(), phi, (size_t)p * D * 8, hipMemcpyDeviceToHost, stream()));
    QED_HIP(hipStreamSynchronize(stream()));
    sym_eigen(n, H.data(), e.data(), true);
    q.e = e;
    q.Q.assign((size_t)p * n, 0.0);
    for (int m = 0; m < n; ++m)
      for (int i = 0; i < p; ++i) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += ph[(size_t)i * n + k] * H[k + (size_t)m * n];
        q.Q[i + (size_t)m * p] = s;
      }
    return q;
  }
  std::vector<double> e, Q;
  int M0 = (int)(14 * p * std::log(1.0 * D));
  int m = band_lanczos(T, phi, p, M0, gi("max_iter_BL"), gd("accur_deflation"), gd("accur_band_lanczos"), e, Q,
                       degeneracy_gap());
  q.e = e;
  q.Q = Q;
  (void)m;
  return q;
}

bool sector_valid(const Instance& I, int nup, int ndn) {
  const int L = I.model->L;
  return nup >= 0 && ndn >= 0 && nup <= L && ndn <= L;
}

// the character of irrep r at element i, real groups (every generator of order <= 2)
double character(const GroupTables& G, int r, int i) {
  const int ph = G.phi[(size_t)r * G.g + i];
  if (ph == 0) return 1.0;
  if (2 * ph == G.g) return -1.0;
  fail("complex characters are not supported");
  return 0.0;
}

// the irrep whose characters are the products of those of a and b (symmetry_group::tensor)
int irrep_tensor(const GroupTables& G, int a, int b) {
  for (int r = 0; r < G.g; ++r) {
    bool ok = true;
    for (int i = 0; i < G.g && ok; ++i)
      ok = G.phi[(size_t)r * G.g + i] == (G.phi[(size_t)a * G.g + i] + G.phi[(size_t)b * G.g + i]) % G.g;
    if (ok) return r;
  }
  fail("no product of the irreducible representations " + std::to_string(a) + " and " + std::to_string(b));
  return 0;
}

// symmetric orbitals of the sites (symmetry_group::build, src_ed/symmetry_group.cpp:293-327):
// for each irrep, each orbit of sites gives sum_i chi(r, e_i) |e_i(site)>, kept (normalized)
// unless it vanishes; n_sites of them in all
void build_symmetric_orbitals(Model& M) {
  const int n = M.n_sites, g = M.group.g;
  M.sorb.clear();
  M.sorb_irrep.clear();
  for (int r = 0; r < g; ++r) {
    std::vector<char> skip(n, 0);
    for (int site = 0; site < n; ++site) {
      if (skip[site]) continue;
      skip[site] = 1;
      std::vector<double> orb(n, 0.0);
      for (int i = 0; i < g; ++i) {
        const int img = M.group.elements[i][site];
        if (img < 0 || img >= n) fail("the point group maps a cluster site outside the cluster");
        skip[img] = 1;
        orb[img] += character(M.group, r, i);
      }
      double nrm = 0.0;
      for (double v : orb) nrm += v * v;
      nrm = std::sqrt(nrm);
      if (nrm < 1e-12) continue;
      for (double v : orb) M.sorb.push_back(v / nrm);
      M.sorb_irrep.push_back(r);
    }
  }
  if ((int)M.sorb_irrep.size() != n)
    fail("Problem with construction of symmetric orbitals: " + std::to_string(M.sorb_irrep.size()) +
         " found instead of " + std::to_string(n));
}

// Lehmann form of one state of a symmetry-adapted sector (model_instance::build_qmatrix,
// src_ed/model_instance.hpp:710-817): per irrep r_o of the symmetric orbitals, c_o|psi>
// and c^dagger_o|psi> lie in the sectors of irrep r (x) r_o with one particle less or
// more; a band Lanczos (or full diagonalization) there gives that irrep's block, which is
// streamlined per part and checked for unitarity in the symmetric basis, as the reference
// does, then turned to the site basis (Q_site = S_o^T Q_o) and concatenated with the others
// The site-basis Lehmann form of per-irrep blocks in the symmetric-orbital basis: each
// block's rows carried to the sites by the orbitals' coefficients, the blocks appended
QMat sym_blocks_to_site(const Model& M, const std::vector<QMat>& blocks) {
  const int n = M.n_sites;
  QMat tot;
  tot.L = n;
  for (int ro = 0; ro < (int)blocks.size(); ++ro) {
    std::vector<int> sel;
    for (int o = 0; o < n; ++o)
      if (M.sorb_irrep[o] == ro) sel.push_back(o);
    const QMat& blk = blocks[ro];
    const int p = (int)sel.size();
    if (p == 0 || blk.M() == 0) continue;
    if (blk.L != p) fail("Q matrix of irrep " + std::to_string(ro) + ": L = " + std::to_string(blk.L) + " for " +
                         std::to_string(p) + " symmetric orbitals");
    QMat site;
    site.L = n;
    site.e = blk.e;
    site.Q.assign((size_t)n * blk.M(), 0.0);
    for (int m = 0; m < blk.M(); ++m)
      for (int k = 0; k < p; ++k) {
        const double w = blk.Q[k + (size_t)m * p];
        const double* row = &M.sorb[(size_t)sel[k] * n];
        for (int a = 0; a < n; ++a) site.Q[a + (size_t)m * n] += row[a] * w;
      }
    append(tot, site);
  }
  tot.L = n;
  return tot;
}

void build_qmatrix_sym(Instance& I, State& st, int spin) {
  Model& M = *I.model;
  const int n = M.n_sites;
  SectorPtr B = make_sector(I, st.nup, st.ndn, nullptr, st.irrep);
  st.gfb[spin].assign(M.group.g, QMat());
  const double sw = std::sqrt(st.weight);
  for (int ro = 0; ro < M.group.g; ++ro) {
    std::vector<int> sel;
    for (int o = 0; o < n; ++o)
      if (M.sorb_irrep[o] == ro) sel.push_back(o);
    const int p = (int)sel.size();
    st.gfb[spin][ro].L = p;
    if (p == 0) continue;
    const int rt = irrep_tensor(M.group, st.irrep, ro);
    QMat parts[2];
    for (int s = 0; s < 2; ++s) {
      const int pm = 2 * s - 1;
      const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
      if (!sector_valid(I, tnup, tndn)) continue;
      SectorPtr T = make_sector(I, tnup, tndn, nullptr, rt);
      if (T->D == 0) continue;
      DevBuf phi((size_t)p * T->D * 8);
      QED_HIP(hipMemsetAsync(phi.p, 0, (size_t)p * T->D * 8, stream()));
      for (int k = 0; k < p; ++k)
        if (qed_mixed_apply_c(B.get(), T.get(), spin, pm, n, &M.sorb[(size_t)sel[k] * n], st.psi.as<double>(),
                              phi.as<double>() + (uint64_t)k * T->D) != 0)
          fail(qed_last_error());
      QMat q = sector_qmatrix(*T, phi.as<double>(), p);
      for (double& e : q.e) {
        e -= st.energy;
        if (pm == -1) e = -e;
      }
      parts[s] = streamline(q);
    }
    QMat blk;
    blk.L = p;
    append(blk, parts[0]);
    append(blk, parts[1]);
    blk.L = p;
    check_norm(blk, gd("accur_Q_matrix"));
    for (double& v : blk.Q) v *= sw;
    st.gfb[spin][ro] = blk;
  }
  st.gf[spin] = sym_blocks_to_site(M, st.gfb[spin]);
}

// Lehmann form of one state for the spin-flip, anomalous and full mixings
// (model_instance::build_qmatrix, src_ed/model_instance.hpp:710-817, trivial group): the
// spinor of build_symmetric_orbitals (src_ed/symmetry_group.cpp:616-645: per Nambu index,
// per spin, per site; the Nambu components are creation operators, model::create_or_destroy,
// src_ed/model.cpp:155-157), one band Lanczos per shifted sector (shift_sector,
// src_ed/symmetry_group.cpp:540-558: every component lands in the same one), one Q-matrix
// of dimension n_mixed x n_sites
std::vector<std::pair<int, bool>> spinor_components(const Instance& I) {
  const Model& M = *I.model;
  const int ns = M.n_sites, L = M.L, m = I.mixing;
  const int nambu_max = (m & 1) ? 2 : 1, spin_max = (m & 2) ? 2 : 1;
  std::vector<std::pair<int, bool>> spinor;
  for (int inb = 0; inb < nambu_max; ++inb)
    for (int isp = 0; isp < spin_max; ++isp)
      for (int i = 0; i < ns; ++i) {
        const int spin = inb ? (spin_max == 2 ? isp : (isp + 1) % 2) : isp;
        spinor.push_back({i + spin * L, inb != 0});
      }
  return spinor;
}

// the sector code or number after adding d particles (shift_sector,
// src_ed/symmetry_group.cpp:540-558): parity codes swap, numbers move by d
int shift_code(int x, int d) { return x == kSecOdd ? kSecEven : (x == kSecEven ? kSecOdd : x + d); }

void build_qmatrix_spinor(Instance& I, State& st) {
  const std::vector<std::pair<int, bool>> spinor = spinor_components(I);
  const int p = (int)spinor.size();
  SectorPtr B = state_sector(I, st);
  QMat parts[2];
  for (int s = 0; s < 2; ++s) {
    const int pm = 2 * s - 1;
    SectorPtr T = make_general_sector(I, shift_code(st.secN, pm), shift_code(st.secS, pm));
    if (T->D == 0) continue;
    DevBuf phi((size_t)p * T->D * 8);
    QED_HIP(hipMemsetAsync(phi.p, 0, (size_t)p * T->D * 8, stream()));
    for (int k = 0; k < p; ++k)
      if (qed_general_apply(B.get(), T.get(), spinor[k].first, spinor[k].second ? -pm : pm, 1.0,
                            st.psi.as<double>(), phi.as<double>() + (uint64_t)k * T->D) != 0)
        fail(qed_last_error());
    QMat q = sector_qmatrix(*T, phi.as<double>(), p);
    for (double& e : q.e) {
      e -= st.energy;
      if (pm == -1) e = -e;
    }
    parts[s] = streamline(q);
  }
  QMat tot;
  tot.L = p;
  append(tot, parts[0]);
  append(tot, parts[1]);
  tot.L = p;
  check_norm(tot, gd("accur_Q_matrix"));
  const double sw = std::sqrt(st.weight);
  for (double& v : tot.Q) v *= sw;
  st.gf[0] = tot;
}

void build_qmatrix(Instance& I, State& st, int spin) {
  if (spinor_mixing(I.mixing)) {
    build_qmatrix_spinor(I, st);
    return;
  }
  if (I.model->symmetric()) {
    build_qmatrix_sym(I, st, spin);
    return;
  }
  Model& M = *I.model;
  const int p = M.n_sites, L = M.L;
  QMat parts[2];
  for (int s = 0; s < 2; ++s) {
    const int pm = 2 * s - 1;
    const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
    if (!sector_valid(I, tnup, tndn)) continue;
    SectorPtr T = make_sector(I, tnup, tndn);
    if (T->D == 0) continue;
    DevBuf phi((size_t)p * T->D * 8);
    QED_HIP(hipMemsetAsync(phi.p, 0, (size_t)p * T->D * 8, stream()));
    for (int i = 0; i < p; ++i)
      if (qed_apply_c(L, st.nup, st.ndn, i + spin * L, pm, 1.0, st.psi.as<double>(),
                      phi.as<double>() + (uint64_t)i * T->D) != 0)
        fail(qed_last_error());
    QMat q = sector_qmatrix(*T, phi.as<double>(), p);
    for (double& e : q.e) {
      e -= st.energy;
      if (pm == -1) e = -e;
    }
    parts[s] = streamline(q);
  }
  QMat tot;
  tot.L = p;
  append(tot, parts[0]);
  append(tot, parts[1]);
  tot.L = p;
  check_norm(tot, gd("accur_Q_matrix"));
  const double sw = std::sqrt(st.weight);
  for (double& v : tot.Q) v *= sw;
  st.gf[spin] = tot;
}

// The continued fractions of one block of p orbitals (model_instance<double>::build_cf,
// src_ed/model_instance.cpp:34-99): for pm = -1, +1 the target sector target(pm) (null if
// none), and for o1 >= o2 the start vector A_o1 |psi> (+ A_o2 |psi>) accumulated by
// apply(T, k, pm, y), normalized, tridiagonalized by Lanczos; b[0] = norm x weight.
template <class Target, class Apply>
void cf_block(State& st, int p, Target target, Apply apply, CFBlock& blk) {
  blk.p = p;
  blk.e.assign((size_t)p * p, CFrac());
  blk.h.assign((size_t)p * p, CFrac());
  for (int pm = -1; pm < 2; pm += 2) {
    SectorPtr T = target(pm);
    if (!T || T->D == 0) continue;
    DevBuf psi(T->D * 8);
    for (int o1 = 0; o1 < p; ++o1)
      for (int o2 = 0; o2 <= o1; ++o2) {
        QED_HIP(hipMemsetAsync(psi.p, 0, T->D * 8, stream()));
        apply(*T, o1, pm, psi.as<double>());
        if (o1 != o2) apply(*T, o2, pm, psi.as<double>());
        const double nrm = dot(psi.as<double>(), psi.as<double>(), T->D);
        if (nrm <= 1e-14) continue;
        scal(1.0 / std::sqrt(nrm), psi.as<double>(), T->D);
        std::vector<double> a, b;
        lanczos_green(*T, psi.as<double>(), gd("accur_continued_fraction"), gi("max_iter_CF"), a, b);
        CFrac f;
        for (size_t i = 0; i < b.size(); ++i) b[i] *= b[i];
        b[0] = nrm * st.weight;
        for (double& x : a) x = (pm == 1) ? x - st.energy : -x + st.energy;
        f.a = a;
        f.b = b;
        f.set = true;
        (pm == 1 ? blk.e : blk.h)[o1 + (size_t)o2 * p] = f;
      }
  }
}

// continued fractions of the symmetric orbitals, one block per irrep (the orbitals'
// coefficients are the rows of W), or of the spinor of the spin-flip, anomalous and full
// mixings (one block, W the identity), with the start vectors of the Lehmann forms
void build_cf_blocks(Instance& I, State& st, int spin) {
  Model& M = *I.model;
  st.cfb[spin].clear();
  if (spinor_mixing(I.mixing)) {
    const std::vector<std::pair<int, bool>> spinor = spinor_components(I);
    const int p = (int)spinor.size();
    SectorPtr B = state_sector(I, st);
    CFBlock blk;
    blk.W.assign((size_t)p * p, 0.0);
    for (int k = 0; k < p; ++k) blk.W[(size_t)k * p + k] = 1.0;
    cf_block(
        st, p, [&](int pm) { return make_general_sector(I, shift_code(st.secN, pm), shift_code(st.secS, pm)); },
        [&](qed_sector& T, int k, int pm, double* y) {
          if (qed_general_apply(B.get(), &T, spinor[k].first, spinor[k].second ? -pm : pm, 1.0,
                                st.psi.as<double>(), y) != 0)
            fail(qed_last_error());
        },
        blk);
    st.cfb[0].push_back(std::move(blk));
    return;
  }
  const int n = M.n_sites;
  SectorPtr B = make_sector(I, st.nup, st.ndn, nullptr, st.irrep);
  for (int ro = 0; ro < M.group.g; ++ro) {
    std::vector<int> sel;
    for (int o = 0; o < n; ++o)
      if (M.sorb_irrep[o] == ro) sel.push_back(o);
    const int p = (int)sel.size();
    if (p == 0) continue;
    const int rt = irrep_tensor(M.group, st.irrep, ro);
    CFBlock blk;
    for (int k = 0; k < p; ++k) blk.W.insert(blk.W.end(), &M.sorb[(size_t)sel[k] * n], &M.sorb[(size_t)sel[k] * n] + n);
    cf_block(
        st, p,
        [&](int pm) -> SectorPtr {
          const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
          if (!sector_valid(I, tnup, tndn)) return nullptr;
          return make_sector(I, tnup, tndn, nullptr, rt);
        },
        [&](qed_sector& T, int k, int pm, double* y) {
          if (qed_mixed_apply_c(B.get(), &T, spin, pm, n, &M.sorb[(size_t)sel[k] * n], st.psi.as<double>(), y) != 0)
            fail(qed_last_error());
        },
        blk);
    st.cfb[spin].push_back(std::move(blk));
  }
}

// continued fractions (model_instance<double>::build_cf, src_ed/model_instance.cpp:34-99)
void build_cf(Instance& I, State& st, int spin) {
  if (spinor_mixing(I.mixing) || I.model->symmetric()) {
    build_cf_blocks(I, st, spin);
    return;
  }
  Model& M = *I.model;
  const int n = M.n_sites, L = M.L;
  st.cfe[spin].assign((size_t)n * n, CFrac());
  st.cfh[spin].assign((size_t)n * n, CFrac());
  for (int pm = -1; pm < 2; pm += 2) {
    const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
    if (!sector_valid(I, tnup, tndn)) continue;
    SectorPtr T = make_sector(I, tnup, tndn);
    if (T->D == 0) continue;
    DevBuf psi(T->D * 8);
    for (int o1 = 0; o1 < n; ++o1)
      for (int o2 = 0; o2 <= o1; ++o2) {
        QED_HIP(hipMemsetAsync(psi.p, 0, T->D * 8, stream()));
        if (qed_apply_c(L, st.nup, st.ndn, o1 + spin * L, pm, 1.0, st.psi.as<double>(), psi.as<double>()) != 0)
          fail(qed_last_error());
        if (o1 != o2 &&
            qed_apply_c(L, st.nup, st.ndn, o2 + spin * L, pm, 1.0, st.psi.as<double>(), psi.as<double>()) != 0)
          fail(qed_last_error());
        double nrm = dot(psi.as<double>(), psi.as<double>(), T->D);
        if (nrm <= 1e-14) continue;
        scal(1.0 / std::sqrt(nrm), psi.as<double>(), T->D);
        std::vector<double> a, b;
        lanczos_green(*T, psi.as<double>(), gd("accur_continued_fraction"), gi("max_iter_CF"), a, b);
        CFrac f;
        for (size_t i = 0; i < b.size(); ++i) b[i] *= b[i];
        b[0] = nrm * st.weight;
        for (double& x : a) x = (pm == 1) ? x - st.energy : -x + st.energy;
        f.a = a;
        f.b = b;
        f.set = true;
        (pm == 1 ? st.cfe[spin] : st.cfh[spin])[o1 + (size_t)o2 * n] = f;
      }
  }
}

// The matrix continued fraction of one start block of p vectors ph (D amplitudes each,
// overwritten): its modified Gram-Schmidt factor W (the first pa columns that do not
// deflate), block Lanczos on the orthonormal block, diagonal blocks shifted by the
// state's energy; W (pa x p) carries sqrt(weight). f is left empty when nothing remains.
void mcf_of_block(qed_sector& T, double* ph, int p, const State& st, int pm, MCFrac& f) {
  f = MCFrac();
  const uint64_t D = T.D;
  const double acc = gd("accur_deflation");
  std::vector<double> W((size_t)p * p, 0.0);
  int pa = p;
  for (int l = 0; l < p; ++l) {
    double* xl = ph + (uint64_t)l * D;
    for (int k = 0; k < l; ++k) {
      const double* xk = ph + (uint64_t)k * D;
      const double z = dot(xk, xl, D);
      W[k + (size_t)p * l] = z;
      axpy(-z, xk, xl, D);
    }
    const double nrm = nrm2(xl, D);
    if (nrm < acc) {
      pa = l;
      break;
    }
    W[l + (size_t)p * l] = nrm;
    scal(1.0 / nrm, xl, D);
  }
  if (pa == 0) return;
  const double sw = std::sqrt(st.weight);
  std::vector<std::vector<double>> A, B;
  const int M0 = (int)(14 * pa * std::log(1.0 * D));
  block_lanczos(T, ph, pa, M0, gi("max_iter_BL"), acc, gd("accur_band_lanczos"), gb("block_Lanczos_QR"), A, B);
  if (A.empty()) return;
  f.pa = pa;
  f.q = p;
  f.W.assign((size_t)pa * p, 0.0);
  for (int l = 0; l < p; ++l)
    for (int k = 0; k < pa; ++k) f.W[k + (size_t)pa * l] = sw * W[k + (size_t)p * l];
  const size_t pp = (size_t)pa * pa;
  f.A.resize(A.size() * pp);
  f.B.resize(A.size() * pp);
  for (size_t j = 0; j < A.size(); ++j)
    for (size_t t = 0; t < pp; ++t) {
      const bool diag = (t % pa) == (t / pa);
      const double a = A[j][t];
      f.A[j * pp + t] = (pm == 1) ? a - (diag ? st.energy : 0.0) : (diag ? st.energy : 0.0) - a;
      f.B[j * pp + t] = B[j][t];
    }
}

// matrix continued fractions (model_instance::build_mcf, src_ed/model_instance.hpp:846-946):
// start block c_i^(+-)|psi> over the sites; for the spin-flip, anomalous and full mixings
// the spinor's components (the Lehmann form's start vectors), q = n_mixed x n_sites; for a
// model with generators one fraction per irrep of the symmetric orbitals, its weight
// W (pa x p) times the orbitals' coefficients (p x n_sites), so that every block's
// W^T F W lands in the site basis.
void build_mcf(Instance& I, State& st, int spin) {
  Model& M = *I.model;
  const int L = M.L;
  if (M.symmetric()) {
    if (gb("combine_mcf")) fail("combine_mcf is not available for a model with a nontrivial point group (model " + M.name + ")");
    const int n = M.n_sites;
    SectorPtr B = make_sector(I, st.nup, st.ndn, nullptr, st.irrep);
    for (int part = 0; part < 2; ++part) st.mcfb[spin][part].clear();
    for (int ro = 0; ro < M.group.g; ++ro) {
      std::vector<int> sel;
      for (int o = 0; o < n; ++o)
        if (M.sorb_irrep[o] == ro) sel.push_back(o);
      const int p = (int)sel.size();
      if (p == 0) continue;
      const int rt = irrep_tensor(M.group, st.irrep, ro);
      for (int part = 0; part < 2; ++part) {
        const int pm = 2 * part - 1;
        const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
        if (!sector_valid(I, tnup, tndn)) continue;
        SectorPtr T = make_sector(I, tnup, tndn, nullptr, rt);
        if (T->D == 0) continue;
        DevBuf phi((size_t)p * T->D * 8);
        double* ph = phi.as<double>();
        QED_HIP(hipMemsetAsync(ph, 0, (size_t)p * T->D * 8, stream()));
        for (int k = 0; k < p; ++k)
          if (qed_mixed_apply_c(B.get(), T.get(), spin, pm, n, &M.sorb[(size_t)sel[k] * n], st.psi.as<double>(),
                                ph + (uint64_t)k * T->D) != 0)
            fail(qed_last_error());
        MCFrac f;
        mcf_of_block(*T, ph, p, st, pm, f);
        if (f.M() == 0) continue;
        std::vector<double> Wn((size_t)f.pa * n, 0.0);
        for (int k = 0; k < f.pa; ++k)
          for (int l = 0; l < p; ++l) {
            const double w = f.W[k + (size_t)f.pa * l];
            for (int a = 0; a < n; ++a) Wn[k + (size_t)f.pa * a] += w * M.sorb[(size_t)sel[l] * n + a];
          }
        f.W = Wn;
        f.q = n;
        st.mcfb[spin][part].push_back(std::move(f));
      }
    }
    return;
  }
  const bool spinor = spinor_mixing(I.mixing);
  const std::vector<std::pair<int, bool>> comps = spinor ? spinor_components(I) : std::vector<std::pair<int, bool>>();
  const int p = spinor ? (int)comps.size() : M.n_sites;
  SectorPtr B = spinor ? state_sector(I, st) : nullptr;
  for (int part = 0; part < 2; ++part) {
    const int pm = 2 * part - 1;
    MCFrac& f = st.mcf[spin][part];
    f = MCFrac();
    SectorPtr T;
    if (spinor) {
      T = make_general_sector(I, shift_code(st.secN, pm), shift_code(st.secS, pm));
    } else {
      const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
      if (!sector_valid(I, tnup, tndn)) continue;
      T = make_sector(I, tnup, tndn);
    }
    if (T->D == 0) continue;
    const uint64_t D = T->D;
    DevBuf phi((size_t)p * D * 8);
    double* ph = phi.as<double>();
    QED_HIP(hipMemsetAsync(ph, 0, (size_t)p * D * 8, stream()));
    for (int i = 0; i < p; ++i) {
      const int rc = spinor ? qed_general_apply(B.get(), T.get(), comps[i].first, comps[i].second ? -pm : pm, 1.0,
                                                st.psi.as<double>(), ph + (uint64_t)i * D)
                            : qed_apply_c(L, st.nup, st.ndn, i + spin * L, pm, 1.0, st.psi.as<double>(),
                                          ph + (uint64_t)i * D);
      if (rc != 0) fail(qed_last_error());
    }
    mcf_of_block(*T, ph, p, st, pm, f);
  }
}

// the reference's sector string (operator<< of sector, src_ed/sector.hpp:100-106)
std::string sector_name(int N, int S, int R = 0) {
  std::string o;
  if (S == kSecOdd && N == kSecOdd) o += "O";
  o += "R" + std::to_string(R);
  if (!sec_code(N)) o += ":N" + std::to_string(N);
  if (!sec_code(S)) o += ":S" + std::to_string(S);
  return o;
}

// ---- ground state (model_instance::low_energy_states, model_instance.hpp:245-417) ------
// model_instance::one_body_solve for the spin-flip, anomalous and full mixings
// (src_ed/model_instance.hpp:1164-1256): the spinor one-body matrix of cluster and bath
// diagonalized, the Lehmann form of its cluster rows, the energy of the negative levels
// plus the Nambu correction (the diagonal one-body elements of the spin-up orbitals for
// the anomalous mixing, of all orbitals for the full one, operators not of the
// cluster-bath kind; one_body_operator.hpp:95-103), halved for the full mixing
void one_body_solve_spinor(Instance& I) {
  Model& M = *I.model;
  const int ns = M.n_sites, nb = M.n_bath, dg = dim_gf(I), dbg = dg / ns * nb, d = dg + dbg;
  std::vector<double> H((size_t)d * d, 0.0), e(d);
  for (int a = 0; a < dg; ++a)
    for (int b = 0; b < dg; ++b) H[a + (size_t)b * d] = I.tc[0][a + (size_t)b * dg];
  for (int a = 0; a < dbg; ++a)
    for (int b = 0; b < dbg; ++b) H[(dg + a) + (size_t)(dg + b) * d] = I.tb[0][a + (size_t)b * dbg];
  for (int a = 0; a < dg; ++a)
    for (int b = 0; b < dbg; ++b) {
      H[a + (size_t)(dg + b) * d] = I.tcb[0][a + (size_t)b * dg];
      H[(dg + b) + (size_t)a * d] = I.tcb[0][a + (size_t)b * dg];
    }
  sym_eigen(d, H.data(), e.data(), true);
  auto st = std::make_unique<State>();
  st->weight = 1.0;
  QMat q;
  q.L = dg;
  q.e = e;
  q.Q.assign((size_t)dg * d, 0.0);
  for (int m = 0; m < d; ++m)
    for (int a = 0; a < dg; ++a) q.Q[a + (size_t)m * dg] = H[a + (size_t)m * d];
  st->gf[0] = q;
  double E = 0.0;
  for (double x : e)
    if (x < 0) E += x;
  double corr = 0.0;
  if (I.mixing & 1)
    for (const Op& op : M.ops) {
      auto it = I.value.find(op.name);
      if (it == I.value.end() || op.interaction || op.anomalous || op.target == 3) continue;
      for (size_t k = 0; k < op.v.size(); ++k)
        if (op.r[k] == op.c[k] && (I.mixing == 3 || op.r[k] < M.L)) corr += op.v[k] * it->second;
    }
  E += corr;
  if (I.mixing == 3) E *= 0.5;
  I.GS_energy = st->energy = I.E0 = E;
  I.states.clear();
  I.states.push_back(std::move(st));
  I.gs_solved = I.gf_solved = true;
  I.gs_string = "uncorrelated";
}

void one_body_solve(Instance& I) {
  if (spinor_mixing(I.mixing)) {
    one_body_solve_spinor(I);
    return;
  }
  Model& M = *I.model;
  const int ns = M.n_sites, nb = M.n_bath, d = ns + nb;
  I.GS_energy = 0.0;
  auto st = std::make_unique<State>();
  st->weight = 1.0;
  const int nspin = (I.mixing & 4) ? 2 : 1;
  for (int sp = 0; sp < nspin; ++sp) {
    std::vector<double> H((size_t)d * d, 0.0), e(d);
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < ns; ++b) H[a + (size_t)b * d] = I.tc[sp][a + (size_t)b * ns];
    for (int a = 0; a < nb; ++a)
      for (int b = 0; b < nb; ++b) H[(ns + a) + (size_t)(ns + b) * d] = I.tb[sp][a + (size_t)b * nb];
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < nb; ++b) {
        H[a + (size_t)(ns + b) * d] = I.tcb[sp][a + (size_t)b * ns];
        H[(ns + b) + (size_t)a * d] = I.tcb[sp][a + (size_t)b * ns];
      }
    sym_eigen(d, H.data(), e.data(), true);
    QMat q;
    q.L = ns;
    q.e = e;
    q.Q.assign((size_t)ns * d, 0.0);
    for (int m = 0; m < d; ++m)
      for (int a = 0; a < ns; ++a) q.Q[a + (size_t)m * ns] = H[a + (size_t)m * d];
    st->gf[sp] = q;
    for (double x : st->gf[0].e)
      if (x < 0) I.GS_energy += x;
  }
  if (I.mixing == 0) I.GS_energy *= 2.0;
  st->energy = I.GS_energy;
  I.E0 = I.GS_energy;
  I.states.clear();
  I.states.push_back(std::move(st));
  I.gs_solved = I.gf_solved = true;
  I.gs_string = "uncorrelated";
}

void finish_states(Instance& I, std::vector<std::unique_ptr<State>>& states);

// keeps the states of one sector within max_gap of the lowest energy seen so far
void add_states(Instance& I, int nup, int ndn, const std::vector<double>& evals, std::vector<DevBuf>& vecs,
                double max_gap, std::vector<std::unique_ptr<State>>& states, int irrep = 0, int secN = 0,
                int secS = 0) {
  if (evals[0] < I.GS_energy) I.GS_energy = evals[0];
  for (size_t k = 0; k < evals.size(); ++k) {
    if (evals[k] - I.GS_energy > max_gap) continue;
    auto st = std::make_unique<State>();
    st->nup = nup;
    st->ndn = ndn;
    st->irrep = irrep;
    st->secN = secN;
    st->secS = secS;
    st->energy = evals[k];
    st->psi = std::move(vecs[k]);
    states.push_back(std::move(st));
  }
}

void low_energy_states(Instance& I) {
  if (I.gs_solved) return;
  if (!I.correlated) {
    one_body_solve(I);
    return;
  }
  I.gs_solved = true;
  const double temperature = gd("temperature");
  double max_gap = -std::log(gd("minimum_weight")) * temperature;
  if (max_gap < MIN_GAP) max_gap = MIN_GAP;
  I.GS_energy = 1.0e12;
  std::vector<std::unique_ptr<State>> states;
  for (size_t is = 0; is < I.sectors.size(); ++is) {
    const int N = I.sectors[is].first, S = I.sectors[is].second, R = I.irreps[is];
    const bool spinor = spinor_mixing(I.mixing);
    const int nup = spinor ? -1 : (N + S) / 2, ndn = spinor ? -1 : (N - S) / 2;
    if (!spinor && !sector_valid(I, nup, ndn)) continue;
    SectorPtr H = spinor ? make_general_sector(I, N, S) : make_sector(I, nup, ndn, nullptr, R);
    if (H->D == 0) continue;
    int n_states = std::max(1, gi("n_states"));
    if ((uint64_t)n_states > H->D) n_states = (int)H->D;
    char method = gc("Ground_state_method");
    if (method == 'L' && n_states > 1) method = 'D';
    std::vector<double> evals;
    std::vector<DevBuf> vecs;
    try {
      if (method == 'D') {
        DevBuf all((size_t)n_states * H->D * 8);
        int nit = 0;
        davidson(*H, n_states, gd("accur_Davidson"), gi("max_iter_lanczos"), evals, all.as<double>(), &nit,
                 gb("verb_ED"));
        for (int k = 0; k < n_states; ++k) {
          DevBuf v(H->D * 8);
          QED_HIP(hipMemcpyAsync(v.p, all.as<double>() + (uint64_t)k * H->D, H->D * 8, hipMemcpyDeviceToDevice,
                                 stream()));
          vecs.push_back(std::move(v));
        }
      } else if (method == 'L') {
        DevBuf v(H->D * 8);
        if (qed_start_vector(H.get(), gi("seed"), v.as<double>()) != 0) fail(qed_last_error());
        double E = 0.0;
        int nit = 0;
        lanczos_gs(*H, gd("accur_lanczos"), gi("max_iter_lanczos"), v.as<double>(), &E, &nit, nullptr, nullptr);
        evals.push_back(E);
        vecs.push_back(std::move(v));
      } else {
        fail("Unknown method in Ground_state_method global parameter");
      }
    } catch (const std::exception& e) {
      fail(std::string(e.what()) + "\n  in cluster model '" + I.model->name + "', instance '" + I.model->name +
           "_" + std::to_string(I.label) + "', sector '" + sector_name(N, S, R) + "'");
    }
    add_states(I, nup, ndn, evals, vecs, max_gap, states, R, N, S);
  }
  finish_states(I, states);
}

// weights and selection of the low-energy states (compute_weights, model_instance.hpp:356-417)
void finish_states(Instance& I, std::vector<std::unique_ptr<State>>& states) {
  const double temperature = gd("temperature");
  double max_gap = -std::log(gd("minimum_weight")) * temperature;
  if (max_gap < MIN_GAP) max_gap = MIN_GAP;
  if (temperature < MIDDLE_VALUE) {
    double e0 = 1e12;
    for (auto& s : states) e0 = std::min(e0, s->energy);
    std::vector<std::unique_ptr<State>> kept;
    for (auto& s : states)
      if (s->energy - e0 <= SMALL_VALUE) kept.push_back(std::move(s));
    states.swap(kept);
    if (states.empty()) fail("instance " + I.model->name + "_" + std::to_string(I.label) + " has no low-energy state!");
    for (auto& s : states) s->weight = 1.0 / states.size();
  } else {
    std::vector<std::unique_ptr<State>> kept;
    for (auto& s : states)
      if (s->energy - I.GS_energy <= max_gap) kept.push_back(std::move(s));
    states.swap(kept);
    double Z = 0.0;
    for (auto& s : states) {
      s->weight = std::exp(-(s->energy - I.GS_energy) / temperature);
      Z += s->weight;
    }
    for (auto& s : states) s->weight /= Z;
  }
  const double minw = gd("minimum_weight");
  std::vector<std::unique_ptr<State>> kept;
  double tot = 0.0;
  for (auto& s : states)
    if (s->weight >= minw) {
      tot += s->weight;
      kept.push_back(std::move(s));
    }
  for (auto& s : kept) s->weight /= tot;
  I.states.swap(kept);
  bool updown = false;
  for (auto& s : I.states)
    if (s->nup != s->ndn) updown = true;
  if (I.mixing == 0 && updown) I.mixing = 4;
  I.GS_energy = 1.0e18;
  I.E0 = 0.0;
  std::ostringstream o;
  for (auto& s : I.states) {
    I.E0 += s->weight * s->energy;
    I.GS_energy = std::min(I.GS_energy, s->energy);
    if (o.tellp() > 0) o << "/";
    if (spinor_mixing(I.mixing)) o << sector_name(s->secN, s->secS, 0) << ":1";
    else o << sector_name(s->nup + s->ndn, s->nup - s->ndn, s->irrep) << ":1";
  }
  I.gs_string = o.str();
  set_hopping_matrix(I);
}

// Green-function format from GF_method (src_ed/model_instance_base.cpp:25-38):
// 'F' continued fractions, 'M' matrix continued fractions, otherwise the Lehmann
// Q-matrix; continued_fraction = true turns 'L' into 'F'
void set_gf_format(Instance& I) {
  char gf = gc("GF_method");
  if (gf == 'L' && gb("continued_fraction")) gf = 'F';
  I.cf = gf == 'F';
  I.mcf = gf == 'M';
  I.q_to_mcf = gf == 'L' && gb("combine_mcf");  // GF_format_Q_to_MCF, model_instance_base.cpp:29-33
}

// MGS factor W (p x p, upper triangle, column-major) of p host vectors of length n
std::vector<double> mgs_factor(std::vector<double> q, int p, size_t n) {
  std::vector<double> W((size_t)p * p, 0.0);
  for (int l = 0; l < p; ++l) {
    double* ql = q.data() + (size_t)l * n;
    for (int k = 0; k < l; ++k) {
      const double* qk = q.data() + (size_t)k * n;
      double z = 0.0;
      for (size_t i = 0; i < n; ++i) z += qk[i] * ql[i];
      W[k + (size_t)p * l] = z;
      for (size_t i = 0; i < n; ++i) ql[i] -= z * qk[i];
    }
    double nrm = 0.0;
    for (size_t i = 0; i < n; ++i) nrm += ql[i] * ql[i];
    nrm = std::sqrt(nrm);
    W[l + (size_t)p * l] = nrm;
    for (size_t i = 0; i < n; ++i) ql[i] /= nrm;
  }
  return W;
}

MCFrac mcf_from_blocks(int p, const std::vector<std::vector<double>>& A, const std::vector<std::vector<double>>& B,
                       const std::vector<double>& W) {
  MCFrac f;
  if (A.empty()) return f;
  f.pa = p;
  f.q = p;
  for (auto& a : A) f.A.insert(f.A.end(), a.begin(), a.end());
  for (auto& b : B) f.B.insert(f.B.end(), b.begin(), b.end());
  f.B.resize(f.A.size(), 0.0);
  f.W = W;
  return f;
}

// combine_mcf with GF_method 'M' (mcf_set::build_combined, src_ed/mcf_set.hpp:79-91;
// combine_via_lanczos, src_ed/matrix_continued_fraction.hpp:442-492): block Lanczos on the
// direct sum of the particle and hole block-tridiagonal matrices, start block W_e on the
// first particle floor and W_h on the first hole floor, M0 = Me + Mh
void build_combined(State& st, int spin) {
  const MCFrac& e = st.mcf[spin][1];
  const MCFrac& h = st.mcf[spin][0];
  st.mcf_comb[spin] = MCFrac();
  if (e.M() == 0 || h.M() == 0) return;
  if (e.pa != h.pa || e.pa != e.q || h.pa != h.q) return;  // a deflated start block: kept apart
  const int p = e.pa, Me = e.M(), Mh = h.M();
  const size_t pp = (size_t)p * p, n = (size_t)(Me + Mh) * p;
  std::vector<double> A(e.A), B(e.B);
  A.insert(A.end(), h.A.begin(), h.A.end());
  B.resize((size_t)Me * pp, 0.0);
  B.insert(B.end(), h.B.begin(), h.B.end());
  B.resize((size_t)(Me + Mh) * pp, 0.0);
  std::vector<double> phi((size_t)p * n, 0.0);
  for (int i = 0; i < p; ++i)
    for (int k = 0; k < p; ++k) {
      phi[(size_t)i * n + k] = e.W[k + (size_t)p * i];
      phi[(size_t)i * n + (size_t)Me * p + k] = h.W[k + (size_t)p * i];
    }
  const std::vector<double> W = mgs_factor(phi, p, n);
  std::vector<std::vector<double>> An, Bn;
  block_lanczos_blocktri(p, Me, Mh, A, B, phi, Me + Mh, gi("max_iter_BL"), gd("accur_deflation"),
                         gd("accur_band_lanczos"), gb("block_Lanczos_QR"), An, Bn);
  st.mcf_comb[spin] = mcf_from_blocks(p, An, Bn, W);
}

// combine_mcf with GF_method 'L' (build_mcf_from_qmatrix, src_ed/model_instance.hpp:
// 967-990; Q_matrix_to_mcf, src_ed/matrix_continued_fraction.hpp:275-314): block Lanczos on
// diag(e) of the Lehmann form from the rows of Q, M0 = M
void qmatrix_to_mcf(State& st, int spin) {
  const QMat& q = st.gf[spin];
  st.mcf_comb[spin] = MCFrac();
  const int L = q.L, M = q.M();
  if (M == 0 || L == 0) return;
  std::vector<double> phi((size_t)L * M);
  for (int a = 0; a < L; ++a)
    for (int k = 0; k < M; ++k) phi[(size_t)a * M + k] = q.Q[a + (size_t)L * k];
  const std::vector<double> W = mgs_factor(phi, L, M);
  std::vector<std::vector<double>> An, Bn;
  block_lanczos_diag(q.e, phi, L, M, gi("max_iter_BL"), gd("accur_deflation"), gd("accur_band_lanczos"),
                     gb("block_Lanczos_QR"), An, Bn);
  st.mcf_comb[spin] = mcf_from_blocks(L, An, Bn, W);
}

void green_function_solve(Instance& I) {
  if (I.gf_solved) return;
  low_energy_states(I);
  if (!I.correlated) return;
  set_gf_format(I);
  for (auto& s : I.states) {
    for (int spin = 0; spin < ((I.mixing & 4) ? 2 : 1); ++spin) {
      if (I.cf) {
        build_cf(I, *s, spin);
      } else if (I.mcf) {
        build_mcf(I, *s, spin);
        if (gb("combine_mcf")) build_combined(*s, spin);
      } else {
        build_qmatrix(I, *s, spin);
        if (I.q_to_mcf) qmatrix_to_mcf(*s, spin);
      }
    }
  }
  I.gf_solved = true;
}

// ---- batched solve of many instances -------------------------------------------------
// Same results as low_energy_states + green_function_solve applied to each instance in
// turn (start vectors are drawn in the same order), with the Krylov work of all
// instances batched: one lockstep Lanczos for every (instance, sector) ground state and
// one band Lanczos pass for every (instance, state, spin, particle/hole) Green function
// sector. Instances that need another method (Davidson, several states per sector,
// continued fractions) are solved one by one.
double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void batch_solve(std::vector<Instance*>& insts, bool graph, int chunk, BatchStats* stats) {
  double t0 = wall();
  struct GS {
    Instance* I;
    int nup, ndn;
    SectorPtr H;
    DevBuf v;
  };
  std::vector<GS> gs;
  std::vector<Instance*> batched;
  for (Instance* I : insts) {
    if (I->gs_solved) continue;
    // symmetry-adapted instances take the one-by-one path
    const bool plain = I->correlated && gc("Ground_state_method") == 'L' && std::max(1, gi("n_states")) == 1 &&
                       !I->model->symmetric() && !spinor_mixing(I->mixing);
    if (!plain) {
      low_energy_states(*I);
      continue;
    }
    I->gs_solved = true;
    I->GS_energy = 1.0e12;
    batched.push_back(I);
    for (auto& sec : I->sectors) {
      const int N = sec.first, S = sec.second;
      const int nup = (N + S) / 2, ndn = (N - S) / 2;
      if (!sector_valid(*I, nup, ndn)) continue;
      GS g{I, nup, ndn, make_sector(*I, nup, ndn), DevBuf()};
      if (g.H->D == 0) continue;
      g.v.alloc(g.H->D * 8);
      if (qed_start_vector(g.H.get(), gi("seed"), g.v.as<double>()) != 0) fail(qed_last_error());
      gs.push_back(std::move(g));
    }
  }
  QED_HIP(hipStreamSynchronize(stream()));
  if (stats) stats->t_setup = wall() - t0;
  t0 = wall();
  {
    std::vector<qed_sector*> S;
    std::vector<double*> psi;
    for (auto& g : gs) {
      S.push_back(g.H.get());
      psi.push_back(g.v.as<double>());
    }
    std::vector<double> E;
    std::vector<int> nit;
    lanczos_gs_batch(S, psi, gd("accur_lanczos"), gi("max_iter_lanczos"), graph, chunk, E, nit, stats);
    const double temperature = gd("temperature");
    double max_gap = -std::log(gd("minimum_weight")) * temperature;
    if (max_gap < MIN_GAP) max_gap = MIN_GAP;
    std::map<Instance*, std::vector<std::unique_ptr<State>>> states;
    std::set<Instance*> failed;
    for (size_t k = 0; k < gs.size(); ++k) {
      if (!(nit[k] > 0)) {
        failed.insert(gs[k].I);
        continue;
      }
      std::vector<double> ev{E[k]};
      std::vector<DevBuf> vv;
      vv.push_back(std::move(gs[k].v));
      add_states(*gs[k].I, gs[k].nup, gs[k].ndn, ev, vv, max_gap, states[gs[k].I]);
    }
    for (Instance* I : batched) {
      if (failed.count(I)) {
        // a sector's Lanczos did not converge within max_iter_lanczos: the instance is
        // reported with E0 = NaN and no states instead of failing the whole batch
        I->states.clear();
        I->GS_energy = I->E0 = std::nan("");
        I->gs_string = "failed";
        I->gf_solved = true;
        continue;
      }
      finish_states(*I, states[I]);
    }
  }
  gs.clear();
  QED_HIP(hipStreamSynchronize(stream()));
  if (stats) stats->t_gs = wall() - t0;
  t0 = wall();
  // Green functions
  struct GF {
    Instance* I;
    State* st;
    int spin, part;  // part 0: holes (pm = -1), 1: particles
    SectorPtr T;
    DevBuf phi;
  };
  std::vector<GF> gf;
  for (Instance* I : insts) {
    if (I->gf_solved || !I->correlated) continue;
    set_gf_format(*I);
    if (I->cf || I->mcf || I->q_to_mcf || I->model->symmetric() || spinor_mixing(I->mixing)) {
      green_function_solve(*I);
      continue;
    }
    Model& M = *I->model;
    const int p = M.n_sites, L = M.L;
    for (auto& stp : I->states) {
      State& st = *stp;
      for (int spin = 0; spin < ((I->mixing & 4) ? 2 : 1); ++spin)
        for (int part = 0; part < 2; ++part) {
          const int pm = 2 * part - 1;
          const int tnup = st.nup + (spin == 0 ? pm : 0), tndn = st.ndn + (spin == 1 ? pm : 0);
          GF g{I, &st, spin, part, SectorPtr(), DevBuf()};
          if (sector_valid(*I, tnup, tndn)) {
            g.T = make_sector(*I, tnup, tndn);
            if (g.T->D > 0) {
              g.phi.alloc((size_t)p * g.T->D * 8);
              QED_HIP(hipMemsetAsync(g.phi.p, 0, (size_t)p * g.T->D * 8, stream()));
              for (int i = 0; i < p; ++i)
                if (qed_apply_c(L, st.nup, st.ndn, i + spin * L, pm, 1.0, st.psi.as<double>(),
                                g.phi.as<double>() + (uint64_t)i * g.T->D) != 0)
                  fail(qed_last_error());
            }
          }
          gf.push_back(std::move(g));
        }
    }
  }
  std::vector<QMat> res(gf.size());
  std::vector<qed_sector*> S;
  std::vector<const double*> phi;
  std::vector<int> pv, which;
  for (size_t k = 0; k < gf.size(); ++k) {
    GF& g = gf[k];
    if (!g.T || g.T->D == 0) continue;
    const int p = g.I->model->n_sites;
    if (g.T->D < (uint64_t)gi("max_dim_full")) {
      res[k] = sector_qmatrix(*g.T, g.phi.as<double>(), p);
      continue;
    }
    S.push_back(g.T.get());
    phi.push_back(g.phi.as<double>());
    pv.push_back(p);
    which.push_back((int)k);
  }
  QED_HIP(hipStreamSynchronize(stream()));
  if (stats) stats->t_gf_setup = wall() - t0;
  t0 = wall();
  if (!S.empty()) {
    std::vector<std::vector<double>> ev, Q;
    band_lanczos_batch(S, phi, pv, gi("max_iter_BL"), gd("accur_deflation"), gd("accur_band_lanczos"), ev, Q,
                       stats, degeneracy_gap());
    for (size_t u = 0; u < which.size(); ++u) {
      QMat& q = res[which[u]];
      q.L = pv[u];
      q.e = std::move(ev[u]);
      q.Q = std::move(Q[u]);
    }
  }
  if (stats) stats->t_band = wall() - t0;
  t0 = wall();
  // assemble per (instance, state, spin) exactly as build_qmatrix does; an instance whose
  // Lehmann form fails the unitarity check is reported failed (E0 = NaN, no states) instead
  // of failing the whole batch (the one-instance path raises the reference's error)
  std::set<Instance*> bad;
  for (size_t k = 0; k < gf.size(); k += 2) {
    GF& h = gf[k];
    if (bad.count(h.I)) continue;
    QMat parts[2];
    for (int part = 0; part < 2; ++part) {
      GF& g = gf[k + part];
      if (!g.T || g.T->D == 0) continue;
      QMat q = std::move(res[k + part]);
      q.L = g.I->model->n_sites;
      for (double& e : q.e) {
        e -= g.st->energy;
        if (part == 0) e = -e;
      }
      parts[part] = streamline(q);
    }
    QMat tot;
    tot.L = h.I->model->n_sites;
    append(tot, parts[0]);
    append(tot, parts[1]);
    tot.L = h.I->model->n_sites;
    try {
      check_norm(tot, gd("accur_Q_matrix"));
    } catch (const Error&) {
      bad.insert(h.I);
      continue;
    }
    const double sw = std::sqrt(h.st->weight);
    for (double& v : tot.Q) v *= sw;
    h.st->gf[h.spin] = tot;
  }
  gf.clear();
  for (Instance* I : bad) {
    I->states.clear();
    I->GS_energy = I->E0 = std::nan("");
    I->gs_string = "failed";
    I->gf_solved = true;
  }
  for (Instance* I : insts)
    if (I->correlated && !I->cf) I->gf_solved = true;
  QED_HIP(hipStreamSynchronize(stream()));
  if (stats) {
    stats->t_assemble = wall() - t0;
    last_batch_stats() = *stats;
  }
}

int spin_index(const Instance& I, int spin_down) {
  if (spin_down && !(I.mixing & 4) && I.mixing != 0)
    fail("spin_down=True impossible with Hilbert space mixing " + std::to_string(I.mixing));
  return (spin_down && (I.mixing & 4)) ? 1 : 0;
}

// G(z) for nz frequencies, d x d complex column-major each
void green_function_eval(Instance& I, int nz, const double* z, int sp, double* out);

// G(z) for nz frequencies, d x d complex column-major each. Frequencies within 1e-8 of one
// evaluated before are served from a per-instance table of the last GF_lookup_depth
// evaluations, as the reference's look-up (model_instance::Green_function,
// src_ed/model_instance.hpp:566-582,624-632: the oldest matching entry wins)
void green_function(Instance& I, int nz, const double* z, int spin_down, double* out) {
  green_function_solve(I);
  const int sp = spin_index(I, spin_down);
  const int d = dim_gf(I);
  const size_t dd2 = (size_t)2 * d * d;
  auto& cache = I.gf_cache[sp];
  const size_t depth = (size_t)std::max(0, gi("GF_lookup_depth"));
  std::vector<int> miss;
  for (int iz = 0; iz < nz; ++iz) {
    const cd w(z[2 * iz], z[2 * iz + 1]);
    const std::vector<double>* hit = nullptr;
    for (auto& e : cache)
      if (std::abs(w - e.first) < MIDDLE_VALUE) hit = &e.second;
    if (hit) std::copy(hit->begin(), hit->end(), out + iz * dd2);
    else miss.push_back(iz);
  }
  if (miss.empty()) return;
  std::vector<double> zm(2 * miss.size()), gm(miss.size() * dd2);
  for (size_t m = 0; m < miss.size(); ++m) {
    zm[2 * m] = z[2 * miss[m]];
    zm[2 * m + 1] = z[2 * miss[m] + 1];
  }
  green_function_eval(I, (int)miss.size(), zm.data(), sp, gm.data());
  for (size_t m = 0; m < miss.size(); ++m) {
    std::copy(gm.begin() + m * dd2, gm.begin() + (m + 1) * dd2, out + miss[m] * dd2);
    if (depth == 0) continue;
    cache.emplace_front(cd(zm[2 * m], zm[2 * m + 1]),
                        std::vector<double>(gm.begin() + m * dd2, gm.begin() + (m + 1) * dd2));
    if (cache.size() > depth) cache.pop_back();
  }
}

// The p x p matrix of one block from per-pair values v(o1 >= o2): the diagonal entries and
// 0.5 (v(o1, o2) - v(o1, o1) - v(o2, o2)) off it (continued_fraction_set::Green_function),
// then added to g (d x d, column-major) as W^T G_block W.
void add_block(const CFBlock& blk, int d, const std::vector<cd>& v, cd* g) {
  const int p = blk.p;
  std::vector<cd> G((size_t)p * p, 0.0);
  for (int o = 0; o < p; ++o) G[o + (size_t)o * p] = v[o + (size_t)o * p];
  for (int o1 = 0; o1 < p; ++o1)
    for (int o2 = 0; o2 < o1; ++o2) {
      const cd t = 0.5 * (v[o1 + (size_t)o2 * p] - v[o1 + (size_t)o1 * p] - v[o2 + (size_t)o2 * p]);
      G[o1 + (size_t)o2 * p] = G[o2 + (size_t)o1 * p] = t;
    }
  for (int k = 0; k < p; ++k)
    for (int l = 0; l < p; ++l) {
      const cd gkl = G[k + (size_t)l * p];
      if (gkl == 0.0) continue;
      for (int b = 0; b < d; ++b) {
        const double wl = blk.W[(size_t)l * d + b];
        if (wl == 0.0) continue;
        for (int a = 0; a < d; ++a) g[a + (size_t)b * d] += blk.W[(size_t)k * d + a] * gkl * wl;
      }
    }
}

// G(z) of the block continued fractions of one state: all set fractions of all blocks in
// one device evaluation
void cf_blocks_eval(const std::vector<CFBlock>& blocks, int d, int nz, const double* z, double* out) {
  std::vector<int> off, dep, blk_of, pair_of;
  std::vector<double> aa, bb;
  for (size_t ib = 0; ib < blocks.size(); ++ib)
    for (int h = 0; h < 2; ++h) {
      const auto& cf = h ? blocks[ib].h : blocks[ib].e;
      for (size_t q = 0; q < cf.size(); ++q) {
        if (!cf[q].set) continue;
        off.push_back((int)aa.size());
        dep.push_back((int)cf[q].a.size());
        aa.insert(aa.end(), cf[q].a.begin(), cf[q].a.end());
        bb.insert(bb.end(), cf[q].b.begin(), cf[q].b.begin() + cf[q].a.size());
        blk_of.push_back((int)ib);
        pair_of.push_back((int)q);
      }
    }
  const int nf = (int)off.size();
  if (nf == 0) return;
  std::vector<double> vals((size_t)2 * nf * nz);
  if (qed_cf_eval(nf, off.data(), dep.data(), aa.data(), bb.data(), nz, z, vals.data()) != 0) fail(qed_last_error());
  for (int iz = 0; iz < nz; ++iz) {
    std::vector<std::vector<cd>> v(blocks.size());
    for (size_t ib = 0; ib < blocks.size(); ++ib) v[ib].assign((size_t)blocks[ib].p * blocks[ib].p, 0.0);
    for (int f = 0; f < nf; ++f)
      v[blk_of[f]][pair_of[f]] += cd(vals[2 * ((size_t)iz * nf + f)], vals[2 * ((size_t)iz * nf + f) + 1]);
    cd* g = reinterpret_cast<cd*>(out + (size_t)2 * iz * d * d);
    for (size_t ib = 0; ib < blocks.size(); ++ib) add_block(blocks[ib], d, v[ib], g);
  }
}

void green_function_eval(Instance& I, int nz, const double* z, int sp, double* out) {
  const int d = dim_gf(I);
  std::fill(out, out + (size_t)2 * nz * d * d, 0.0);
  for (auto& s : I.states) {
    if ((I.mcf || I.q_to_mcf) && I.correlated && s->mcf_comb[sp].M() > 0) {
      // one combined fraction (mcf_set::Green_function with combined[r])
      const MCFrac& f = s->mcf_comb[sp];
      mcf_eval(f.pa, f.q, f.M(), f.A.data(), f.B.data(), f.W.data(), nz, z, 0, out);
    } else if (I.mcf && I.correlated) {
      // particles added as they are, holes transposed (src_ed/mcf_set.hpp:111-131)
      for (int part = 1; part >= 0; --part) {
        const MCFrac& f = s->mcf[sp][part];
        if (f.M() > 0) mcf_eval(f.pa, f.q, f.M(), f.A.data(), f.B.data(), f.W.data(), nz, z, part == 0, out);
        for (const MCFrac& fb : s->mcfb[sp][part])
          mcf_eval(fb.pa, fb.q, fb.M(), fb.A.data(), fb.B.data(), fb.W.data(), nz, z, part == 0, out);
      }
    } else if (!I.cf || !I.correlated) {
      const QMat& q = s->gf[sp];
      if (q.M() == 0) continue;
      if (qed_qmatrix_gf(d, q.M(), q.e.data(), q.Q.data(), nz, z, out) != 0) fail(qed_last_error());
    } else if (!s->cfb[sp].empty()) {
      cf_blocks_eval(s->cfb[sp], d, nz, z, out);
    } else {
      // continued fractions: diagonal entries and the (o1 + o2) combinations
      std::vector<int> off, dep;
      std::vector<double> aa, bb;
      std::vector<int> which;  // (o1, o2, e/h)
      for (int h = 0; h < 2; ++h) {
        auto& cf = h ? s->cfh[sp] : s->cfe[sp];
        for (int o1 = 0; o1 < d; ++o1)
          for (int o2 = 0; o2 <= o1; ++o2) {
            const CFrac& f = cf[o1 + (size_t)o2 * d];
            if (!f.set) continue;
            off.push_back((int)aa.size());
            dep.push_back((int)f.a.size());
            aa.insert(aa.end(), f.a.begin(), f.a.end());
            bb.insert(bb.end(), f.b.begin(), f.b.begin() + f.a.size());
            which.push_back(o1 + d * o2);
          }
      }
      const int nf = (int)off.size();
      if (nf == 0) continue;
      std::vector<double> vals((size_t)2 * nf * nz);
      if (qed_cf_eval(nf, off.data(), dep.data(), aa.data(), bb.data(), nz, z, vals.data()) != 0)
        fail(qed_last_error());
      for (int iz = 0; iz < nz; ++iz) {
        std::vector<cd> comb((size_t)d * d, 0.0);
        for (int f = 0; f < nf; ++f) comb[which[f]] += cd(vals[2 * ((size_t)iz * nf + f)], vals[2 * ((size_t)iz * nf + f) + 1]);
        double* g = out + (size_t)2 * iz * d * d;
        for (int o1 = 0; o1 < d; ++o1) {
          g[2 * (o1 + (size_t)o1 * d)] += comb[o1 + (size_t)o1 * d].real();
          g[2 * (o1 + (size_t)o1 * d) + 1] += comb[o1 + (size_t)o1 * d].imag();
        }
        for (int o1 = 0; o1 < d; ++o1)
          for (int o2 = 0; o2 < o1; ++o2) {
            cd t = 0.5 * (comb[o1 + (size_t)o2 * d] - comb[o1 + (size_t)o1 * d] - comb[o2 + (size_t)o2 * d]);
            g[2 * (o1 + (size_t)o2 * d)] += t.real();
            g[2 * (o1 + (size_t)o2 * d) + 1] += t.imag();
            g[2 * (o2 + (size_t)o1 * d)] += t.real();
            g[2 * (o2 + (size_t)o1 * d) + 1] += t.imag();
          }
      }
    }
  }
}

void hybridization(Instance& I, int nz, const double* z, int spin_down, double* out) {
  low_energy_states(I);
  const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
  const int d = dim_gf(I), nb = I.model->n_bath * n_mixed(I.mixing);
  for (int iz = 0; iz < nz; ++iz) {
    cd w(z[2 * iz], z[2 * iz + 1]);
    double* g = out + (size_t)2 * iz * d * d;
    for (int a = 0; a < d; ++a)
      for (int b = 0; b < d; ++b) {
        cd s = 0.0;
        for (int i = 0; i < nb; ++i)
          s += I.tcb_d[sp][a + (size_t)i * d] * I.tcb_d[sp][b + (size_t)i * d] / (w - I.eb[sp][i]);
        g[2 * (a + (size_t)b * d)] = s.real();
        g[2 * (a + (size_t)b * d) + 1] = s.imag();
      }
  }
}

// ---- integrated Green function: the matrix of <c^dagger_j c_i> from the Green-function
// representation (model_instance::Green_function_average, src_ed/model_instance.hpp:
// 643-654; Q_matrix::integrated_Green_function, Q_matrix.hpp:161-171: poles below zero;
// continued_fraction_set::integrated_Green_function, continued_fraction_set.cpp:104-125:
// b_0 of the hole fractions; mcf_set::integrated_Green_function, mcf_set.hpp:146-158:
// (W_h^T W_h)^T). M(b, a) accumulates over the states, d x d column-major.
std::vector<cd> gf_average(Instance& I, int sp) {
  const int d = dim_gf(I);
  std::vector<cd> M((size_t)d * d, 0.0);
  auto at = [&](int r, int c) -> cd& { return M[r + (size_t)c * d]; };
  for (auto& s : I.states) {
    if (I.mcf && I.correlated) {
      std::vector<const MCFrac*> hs{&s->mcf[sp][0]};
      for (const MCFrac& fb : s->mcfb[sp][0]) hs.push_back(&fb);
      for (const MCFrac* hp : hs) {
        const MCFrac& h = *hp;
        if (h.M() == 0 || h.pa == 0) continue;
        for (int a = 0; a < h.q; ++a)
          for (int b = 0; b < h.q; ++b) {
            double w = 0.0;
            for (int k = 0; k < h.pa; ++k) w += h.W[k + (size_t)h.pa * a] * h.W[k + (size_t)h.pa * b];
            at(b, a) += w;  // (W^T W)^T
          }
      }
    } else if (I.cf && I.correlated && !s->cfb[sp].empty()) {
      // the hole fractions' weights b[0] per block, the same combination
      std::vector<cd> g((size_t)d * d, 0.0);
      for (const CFBlock& blk : s->cfb[sp]) {
        std::vector<cd> v((size_t)blk.p * blk.p, 0.0);
        for (size_t q = 0; q < v.size(); ++q)
          if (blk.h[q].set) v[q] = blk.h[q].b[0];
        add_block(blk, d, v, g.data());
      }
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) at(a, b) += g[a + (size_t)b * d];
    } else if (I.cf && I.correlated) {
      const auto& h = s->cfh[sp];
      if (h.size() != (size_t)d * d) continue;
      auto b0 = [&](int o1, int o2) { return h[o1 + (size_t)o2 * d].set ? h[o1 + (size_t)o2 * d].b[0] : 0.0; };
      for (int o = 0; o < d; ++o) at(o, o) += b0(o, o);
      for (int o1 = 0; o1 < d; ++o1)
        for (int o2 = 0; o2 < o1; ++o2) {
          const cd t = 0.5 * (b0(o1, o2) - at(o1, o1) - at(o2, o2));
          at(o1, o2) += t;
          at(o2, o1) += t;
        }
    } else {
      const QMat& q = s->gf[sp];
      for (int i = 0; i < q.M(); ++i) {
        if (q.e[i] >= 0.0) continue;
        for (int a = 0; a < q.L; ++a)
          for (int b = 0; b < q.L; ++b) at(b, a) += q.Q[a + (size_t)i * q.L] * q.Q[b + (size_t)i * q.L];
      }
    }
  }
  return M;
}

// Neville's polynomial interpolation of (xa, ya), evaluated at x (Numerical Recipes
// polint; the reference's polynomial_fit, src_ed/model_instance.cpp:184-222)
double neville(const std::vector<double>& xa, const std::vector<double>& ya, double x) {
  const size_t n = xa.size();
  std::vector<double> P(ya);
  for (size_t m = 1; m < n; ++m)
    for (size_t i = 0; i + m < n; ++i)
      P[i] = ((x - xa[i + m]) * P[i] + (xa[i] - x) * P[i + 1]) / (xa[i] - xa[i + m]);
  return P[0];
}

// ---- dynamic susceptibility (model_instance::susceptibility / susceptibility_poles,
// src_ed/model_instance.hpp:1024-1113) ----------------------------------------------------
// For each low-energy state: b = O|psi> with the operator's own Hamiltonian-space form
// (coefficient 1), the Lehmann form of b in the state's sector (band Lanczos, or full
// diagonalization below max_dim_full) and the poles shifted by the state's energy.
struct SusceptibilityPart {
  double weight;
  QMat q;
};

// O|psi> with the operator restricted to the state's own sector, as HS_operator.at(sec)
// (the symmetry-adapted and mixed-number sectors included: the part of O|psi> in the
// state's irrep and number sector)
std::vector<SusceptibilityPart> susceptibility_parts(Instance& I, const std::string& op) {
  if (!I.model->idx.count(op)) fail("Operator " + op + " is not defined in the model.");
  if (!I.value.count(op)) fail("operator " + op + " is not activated in instance " + std::to_string(I.label));
  low_energy_states(I);
  if (I.from_file) fail("the susceptibility needs the ground-state vectors, which an instance read from a file does not hold");
  std::vector<SusceptibilityPart> out;
  for (auto& st : I.states) {
    if (!st->psi.p) fail("the susceptibility needs the ground-state vectors of the instance");
    SectorPtr H = state_sector(I, *st);
    if (H->D == 0) continue;
    SectorPtr O = state_sector(I, *st, &op);
    DevBuf b(H->D * 8);
    QED_HIP(hipMemsetAsync(b.p, 0, H->D * 8, stream()));
    hv(*O, st->psi.as<double>(), b.as<double>());
    QMat q = sector_qmatrix(*H, b.as<double>(), 1);
    for (double& e : q.e) e -= st->energy;
    out.push_back({st->weight, std::move(q)});
  }
  return out;
}

// ---- instance files -------------------------------------------------------------------
// The reference's HDF5 layout (model_instance::write_hdf5 / read_hdf5,
// src_ed/model_instance.hpp:1339-1421; state::write_hdf5, src_ed/state.hpp:46-115;
// Q_matrix / Q_matrix_set, src_ed/Q_matrix.hpp:320-345, src_ed/Q_matrix_set.hpp:183-210;
// continued_fraction_set, src_ed/continued_fraction_set.cpp:136-176; mcf_set and its
// matrix continued fractions, src_ed/mcf_set.hpp:162-197, matrix_continued_fraction.hpp:
// 511-547; primitives src_util/hdf5_io.cpp), for the trivial symmetry group (one block).

const char* gf_format_name(const Instance& I) {
  if (I.correlated && I.mcf) return "mcf";
  if (I.correlated && I.cf) return "cf";
  return "bl";
}

void write_mcf(const h5::Node& g, const MCFrac& f) {
  const int M = f.M(), pa = f.pa;
  const size_t pp = (size_t)pa * pa;
  h5::write_attr(g, "p", pa);
  h5::write_attr(g, "floors", M);
  for (int j = 0; j < M; ++j) {
    h5::Node a = h5::create_group(g, "A_" + std::to_string(j));
    h5::write_mat(a, "data", pa, pa, f.A.data() + j * pp);
  }
  for (int j = 0; j < (int)(f.B.size() / std::max<size_t>(pp, 1)); ++j) {
    h5::Node b = h5::create_group(g, "B_" + std::to_string(j));
    h5::write_mat(b, "data", pa, pa, f.B.data() + j * pp);
  }
  h5::Node w = h5::create_group(g, "W");
  h5::write_mat_complex(w, "data", pa, pa ? f.q : 0, f.W.data(), nullptr);
}

MCFrac read_mcf(const h5::Node& g) {
  MCFrac f;
  f.pa = h5::read_attr_int(g, "p");
  const int M = h5::read_attr_int(g, "floors");
  int r, c;
  for (int j = 0; j < M; ++j) {
    auto a = h5::read_mat_real(h5::open_group(g, "A_" + std::to_string(j)), "data", r, c);
    if (r != f.pa || c != f.pa) fail("matrix continued fraction: block A_" + std::to_string(j) + " is not p x p");
    f.A.insert(f.A.end(), a.begin(), a.end());
  }
  for (int j = 0; j < M; ++j) {
    auto b = h5::read_mat_real(h5::open_group(g, "B_" + std::to_string(j)), "data", r, c);
    if (r != f.pa || c != f.pa) fail("matrix continued fraction: block B_" + std::to_string(j) + " is not p x p");
    f.B.insert(f.B.end(), b.begin(), b.end());
  }
  f.W = h5::read_mat_real(h5::open_group(g, "W"), "data", r, c);
  if (M > 0 && r != f.pa) fail("matrix continued fraction: W has " + std::to_string(r) + " rows for p = " +
                               std::to_string(f.pa));
  f.q = M > 0 ? c : 0;
  if (M == 0) f = MCFrac();
  return f;
}

void write_cf(const h5::Node& g, const std::vector<CFrac>& e, const std::vector<CFrac>& h, int d) {
  h5::write_attr(g, "is_complex", 0);
  for (int a = 0; a < d; ++a)
    for (int b = 0; b <= a; ++b) {
      h5::Node s = h5::create_group(g, "r0_a" + std::to_string(a) + "_b" + std::to_string(b));
      for (int k = 0; k < 2; ++k) {
        const std::vector<CFrac>& v = k ? h : e;
        static const CFrac none;
        const CFrac& f = v.size() == (size_t)d * d ? v[a + (size_t)b * d] : none;
        h5::Node x = h5::create_group(s, k ? "h" : "e");
        h5::write_vec(x, "a", f.set ? f.a : std::vector<double>());
        h5::write_vec(x, "b", f.set ? f.b : std::vector<double>());
      }
    }
}

void read_cf(const h5::Node& g, std::vector<CFrac>& e, std::vector<CFrac>& h, int d) {
  if (h5::read_attr_int(g, "is_complex") != 0) fail("complex continued fractions: this solver has a real Hilbert space");
  e.assign((size_t)d * d, CFrac());
  h.assign((size_t)d * d, CFrac());
  for (int a = 0; a < d; ++a)
    for (int b = 0; b <= a; ++b) {
      h5::Node s = h5::open_group(g, "r0_a" + std::to_string(a) + "_b" + std::to_string(b));
      for (int k = 0; k < 2; ++k) {
        h5::Node x = h5::open_group(s, k ? "h" : "e");
        CFrac f;
        f.a = h5::read_vec(x, "a");
        f.b = h5::read_vec(x, "b");
        if (f.b.size() < f.a.size()) fail("continued fraction r0_a" + std::to_string(a) + "_b" + std::to_string(b) +
                                          ": fewer b than a coefficients");
        f.set = !f.a.empty();
        (k ? h : e)[a + (size_t)b * d] = f;
      }
    }
}

void write_gf(const h5::Node& g, const Instance& I, const State& s, int spin, const std::string& fmt) {
  const int d = I.model->n_sites;
  if (I.model->symmetric()) {
    // one block per irrep, in the symmetric-orbital basis (Q_matrix_set::write_hdf5,
    // src_ed/Q_matrix_set.hpp:183-195)
    if (fmt != "bl") fail("the instance file of a model with generators holds its Lehmann form only");
    h5::write_attr(g, "nblocks", (int)s.gfb[spin].size());
    for (size_t r = 0; r < s.gfb[spin].size(); ++r) {
      const QMat& q = s.gfb[spin][r];
      h5::Node b = h5::create_group(g, "block_" + std::to_string(r));
      h5::write_attr(b, "L", q.L);
      h5::write_attr(b, "M", q.M());
      h5::write_vec(b, "e", q.e);
      h5::write_mat(b, "v", q.L, q.M(), q.Q.data());
    }
    return;
  }
  if (fmt == "bl") {
    const QMat& q = s.gf[spin];
    h5::write_attr(g, "nblocks", 1);
    h5::Node b = h5::create_group(g, "block_0");
    h5::write_attr(b, "L", q.L);
    h5::write_attr(b, "M", q.M());
    h5::write_vec(b, "e", q.e);
    h5::write_mat(b, "v", q.L, q.M(), q.Q.data());
  } else if (fmt == "cf") {
    write_cf(g, s.cfe[spin], s.cfh[spin], d);
  } else {
    h5::write_attr(g, "nblocks", 1);
    h5::Node b = h5::create_group(g, "block_0");
    h5::Node e = h5::create_group(b, "e");
    write_mcf(e, s.mcf[spin][1]);
    h5::Node h = h5::create_group(b, "h");
    write_mcf(h, s.mcf[spin][0]);
  }
}

void read_gf(const h5::Node& g, const Instance& I, State& s, int spin, const std::string& fmt, int mixing) {
  const int d = I.model->n_sites * n_mixed(mixing);
  if (I.model->symmetric()) {
    const int nbl = h5::read_attr_int(g, "nblocks");
    if (nbl != I.model->group.g)
      fail("the Green function has " + std::to_string(nbl) + " symmetry blocks; the model's group has " +
           std::to_string(I.model->group.g) + " irreps");
    s.gfb[spin].assign(nbl, QMat());
    for (int r = 0; r < nbl; ++r) {
      h5::Node b = h5::open_group(g, "block_" + std::to_string(r));
      QMat& q = s.gfb[spin][r];
      q.L = h5::read_attr_int(b, "L");
      const int M = h5::read_attr_int(b, "M");
      q.e = h5::read_vec(b, "e");
      int rr = 0, cc = 0;
      if (M > 0) q.Q = h5::read_mat_real(b, "v", rr, cc);
      if ((int)q.e.size() != M || (M > 0 && (rr != q.L || cc != M))) fail("Q matrix: e and v do not match L and M");
    }
    s.gf[spin] = sym_blocks_to_site(*I.model, s.gfb[spin]);
    return;
  }
  if (fmt == "cf") {
    read_cf(g, s.cfe[spin], s.cfh[spin], d);
    return;
  }
  if (h5::read_attr_int(g, "nblocks") != 1)
    fail("the Green function has " + std::to_string(h5::read_attr_int(g, "nblocks")) +
         " symmetry blocks; this solver has the trivial group (one block)");
  h5::Node b = h5::open_group(g, "block_0");
  if (fmt == "bl") {
    QMat q;
    q.L = h5::read_attr_int(b, "L");
    const int M = h5::read_attr_int(b, "M");
    q.e = h5::read_vec(b, "e");
    int r, c;
    q.Q = h5::read_mat_real(b, "v", r, c);
    if ((int)q.e.size() != M || r != q.L || c != M) fail("Q matrix: e and v do not match L and M");
    if (M > 0 && q.L != d) fail("Q matrix: L = " + std::to_string(q.L) + " for a cluster of " + std::to_string(d));
    s.gf[spin] = q;
  } else {
    s.mcf[spin][1] = read_mcf(h5::open_group(b, "e"));
    s.mcf[spin][0] = read_mcf(h5::open_group(b, "h"));
    for (auto& f : s.mcf[spin])
      if (f.M() > 0 && f.q != d) fail("matrix continued fraction: W has " + std::to_string(f.q) + " columns for a cluster of " +
                                      std::to_string(d));
  }
}

void write_instance_file(Instance& I, const std::string& path, const std::string& group) {
  green_function_solve(I);
  const std::string fmt = gf_format_name(I);
  if (spinor_mixing(I.mixing) && fmt == "cf")
    fail("the instance file of a spin-flip or anomalous instance takes its Lehmann form or matrix continued "
         "fraction, not continued fractions");
  try {
    h5::Node f = h5::open_or_create(path);
    if (h5::exists(f, group)) h5::unlink(f, group);
    h5::Node g = h5::create_group(f, group);
    h5::write_attr(g, "GS_energy", I.GS_energy);
    h5::write_attr(g, "GS_sectors", I.gs_string);
    h5::write_attr(g, "GF_format", fmt);
    h5::write_attr(g, "mixing", I.mixing);
    h5::write_attr(g, "complex_HS", 0);
    {
      h5::Node pg = h5::create_group(g, "params");
      for (auto& kv : I.value) h5::write_attr(pg, kv.first, kv.second);
    }
    h5::write_attr(g, "nstates", (int)I.states.size());
    for (size_t k = 0; k < I.states.size(); ++k) {
      const State& s = *I.states[k];
      h5::Node sg = h5::create_group(g, "state_" + std::to_string(k));
      h5::write_attr(sg, "sec", spinor_mixing(I.mixing) ? sector_name(s.secN, s.secS)
                                                         : sector_name(s.nup + s.ndn, s.nup - s.ndn, s.irrep));
      h5::write_attr(sg, "energy", s.energy);
      h5::write_attr(sg, "weight", s.weight);
      h5::write_attr(sg, "gf_format", fmt);
      write_gf(h5::create_group(sg, "gf"), I, s, 0, fmt);
      if (I.mixing & 4) write_gf(h5::create_group(sg, "gf_down"), I, s, 1, fmt);
    }
  } catch (const std::exception& e) {
    fail(std::string("HDF5 error while writing instance: ") + e.what());
  }
}

std::pair<int, int> sector_numbers(const std::string& s, int g = 1);
int sector_irrep(const std::string& s, int g);

void read_instance_file(Instance& I, const std::string& path, const std::string& group) {
  std::vector<std::unique_ptr<State>> states;
  double gse = 0.0;
  std::string fmt, gs;
  int mixing = 0;
  try {
    h5::Node f = h5::open_read(path);
    h5::Node g = h5::open_group(f, group);
    gse = h5::read_attr_double(g, "GS_energy");
    fmt = h5::read_attr_string(g, "GF_format");
    mixing = h5::read_attr_int(g, "mixing");
    gs = h5::attr_exists(g, "GS_sectors") ? h5::read_attr_string(g, "GS_sectors") : std::string();
    {
      h5::Node pg = h5::open_group(g, "params");
      for (auto& kv : I.value)
        if (h5::attr_exists(pg, kv.first)) {
          const double stored = h5::read_attr_double(pg, kv.first);
          if (std::abs(stored - kv.second) > MIDDLE_VALUE)
            std::printf("WARNING : The value of %s from the HDF5 solution (%g) differs from the expected value (%g).\n",
                        kv.first.c_str(), stored, kv.second);
          std::fflush(stdout);
        }
    }
    const int n = h5::read_attr_int(g, "nstates");
    for (int k = 0; k < n; ++k) {
      h5::Node sg = h5::open_group(g, "state_" + std::to_string(k));
      auto st = std::make_unique<State>();
      const auto ns = sector_numbers(h5::read_attr_string(sg, "sec"), I.model->group.g);
      if (spinor_mixing(mixing)) {
        st->nup = st->ndn = -1;
        st->secN = ns.first;
        st->secS = ns.second;
        if (fmt != "bl" && fmt != "mcf")
          fail("the instance file of a spin-flip or anomalous instance holds continued fractions");
      } else {
        st->nup = (ns.first + ns.second) / 2;
        st->ndn = (ns.first - ns.second) / 2;
        st->irrep = sector_irrep(h5::read_attr_string(sg, "sec"), I.model->group.g);
        if (I.model->symmetric() && fmt != "bl")
          fail("the instance file of a model with generators holds continued fractions");
      }
      st->energy = h5::read_attr_double(sg, "energy");
      st->weight = h5::read_attr_double(sg, "weight");
      std::string sf = h5::read_attr_string(sg, "gf_format");
      if (sf != "bl" && sf != "mcf") sf = "cf";
      if (h5::exists(sg, "gf")) read_gf(h5::open_group(sg, "gf"), I, *st, 0, sf, mixing);
      if (h5::exists(sg, "gf_down")) read_gf(h5::open_group(sg, "gf_down"), I, *st, 1, sf, mixing);
      states.push_back(std::move(st));
    }
  } catch (const std::exception& e) {
    fail(std::string("HDF5 error while reading instance: ") + e.what());
  }
  // the instance changes only once the whole group has been read
  I.states.swap(states);
  I.gf_cache[0].clear();
  I.gf_cache[1].clear();
  I.GS_energy = gse;
  I.mixing = mixing;
  I.cf = fmt == "cf" || (fmt != "bl" && fmt != "mcf");
  I.mcf = fmt == "mcf";
  I.q_to_mcf = false;  // a file holds the particle and hole parts, not a combined fraction
  I.correlated = true;
  I.gs_solved = I.gf_solved = true;
  I.from_file = true;
  I.E0 = 0.0;
  for (auto& s : I.states) I.E0 += s->weight * s->energy;
  I.gs_string = gs;
  set_hopping_matrix(I);  // one-body matrices and the bath term for the file's mixing
}

}  // namespace ed
}  // namespace qed

using namespace qed;
using namespace qed::ed;

namespace {
std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == c) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  out.push_back(cur);
  return out;
}

// "R0:N12:S0" -> (N, S)
// "R<r>:N<n>:S<s>" (sector::sector, src_ed/sector.cpp): irrep r of a group of order g,
// N particles, S = N_up - N_down
struct SectorLabel {
  int R, N, S;
};
SectorLabel parse_sector_label(const std::string& s, int g) {
  if (s.find('N') == std::string::npos || s.find('S') == std::string::npos || s.find('O') != std::string::npos) {
    // N or S not conserved (src_ed/sector.hpp:52-95): a missing label is "even", 'O'
    // makes both odd, and an odd value of the other makes it "odd"
    int R = 0, N = kSecEven, S = kSecEven;
    for (auto& tok : split(s, ':')) {
      if (tok.empty()) continue;
      std::string t = tok;
      if (t[0] == 'O') t = t.substr(1);
      if (t.empty()) continue;
      const char k = t[0];
      try {
        if (k == 'R') R = std::stoi(t.substr(1));
        else if (k == 'N') N = std::stoi(t.substr(1));
        else if (k == 'S') S = std::stoi(t.substr(1));
        else fail("sector string " + s + " does not conform to standard!");
      } catch (const std::invalid_argument&) {
        fail("sector string " + s + " does not conform to standard!");
      }
    }
    if (s.find('O') != std::string::npos) N = S = kSecOdd;
    if (S == kSecEven && N != kSecEven && (N % 2)) S = kSecOdd;
    if (N == kSecEven && S != kSecEven && (S % 2)) N = kSecOdd;
    if (R < 0 || R >= g) fail("sector " + s + " is not valid");
    if (!sec_code(N) && N < 0) fail("sector " + s + " is not valid");
    return {R, N, S};
  }
  int R = -1, N = -1000, S = -1000;
  for (auto& tok : split(s, ':')) {
    if (tok.empty()) continue;
    char k = tok[0];
    std::string v = tok.substr(1);
    try {
      if (k == 'R') R = std::stoi(v);
      else if (k == 'N') N = std::stoi(v);
      else if (k == 'S') S = std::stoi(v);
    } catch (...) {
      fail("sector " + s + " is not valid");
    }
  }
  if (R < 0 || R >= g || N == -1000 || S == -1000) fail("sector " + s + " is not valid");
  if ((N + S) % 2 != 0 || S > N || S < -N || N < 0) fail("sector " + s + " is not valid");
  return {R, N, S};
}

}  // namespace

std::pair<int, int> qed::ed::sector_numbers(const std::string& s, int g) {
  const SectorLabel l = parse_sector_label(s, g);
  return {l.N, l.S};
}
int qed::ed::sector_irrep(const std::string& s, int g) { return parse_sector_label(s, g).R; }

extern "C" {

int qed_ed_new_model(const char* name, int n_sites, int n_bath) {
  return qed_ed_new_model_generators(name, n_sites, n_bath, 0, nullptr, 0);
}

int qed_ed_new_model_generators(const char* name, int n_sites, int n_bath, int n_gen, const int* gens,
                                int bath_irrep) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (models().count(name)) fail(std::string("the name '") + name + "' has already been used for another cluster model");
    if (n_sites <= 0 || n_bath < 0 || n_sites + n_bath > 32) fail("invalid model size");
    if (n_gen < 0 || (n_gen > 0 && !gens)) fail("invalid generators");
    auto m = std::make_unique<Model>();
    m->name = name;
    m->n_sites = n_sites;
    m->n_bath = n_bath;
    m->L = n_sites + n_bath;
    for (int k = 0; k < n_gen; ++k) m->gens.emplace_back(gens + (size_t)k * m->L, gens + (size_t)(k + 1) * m->L);
    m->bath_irrep = n_gen > 0 && bath_irrep != 0;
    m->group = group_tables(m->L, n_sites, m->gens, m->bath_irrep);
    if (m->symmetric()) build_symmetric_orbitals(*m);
    models()[name] = std::move(m);
  });
}

int qed_ed_new_operator(const char* model, const char* name, const char* type, int nel, const int* r,
                        const int* c, const double* v) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    if (nel == 0) return;
    auto it = models().find(model);
    if (it == models().end()) fail(std::string("The model ") + model + " is not defined. Check spelling.");
    Model& M = *it->second;
    if (M.closed) fail("model " + M.name + " has already been instantiated and is closed for modifications");
    std::string t(type);
    Op op;
    op.name = name;
    op.type = t;
    if (t == "interaction") op.interaction = true;
    else if (t == "anomalous") op.anomalous = true;
    else if (t != "one-body") fail("type of operator " + op.name + " is not supported by this solver");
    bool touches_bath = false, touches_site = false, hybrid = false;
    for (int i = 0; i < nel; ++i) {
      int a = r[i], b = c[i];
      if (a < 0 || b < 0 || a >= 2 * M.L || b >= 2 * M.L) fail("operator element index out of range");
      if (op.interaction && a > b) std::swap(a, b);
      op.r.push_back(a);
      op.c.push_back(b);
      op.v.push_back(v[i]);
      bool ab = (a % M.L) >= M.n_sites, bb = (b % M.L) >= M.n_sites;
      touches_bath |= ab || bb;
      touches_site |= !ab || !bb;
      hybrid |= ab != bb;
    }
    op.target = hybrid ? 3 : (touches_bath && !touches_site ? 2 : 1);
    // one target per operator (set_target, src_ed/Operators/one_body_operator.hpp and
    // anomalous_operator.hpp:131-162): cluster only, bath only, or cluster-bath only
    if (!op.interaction) {
      bool mixed_target = false;
      for (int i = 0; i < nel; ++i) {
        const bool ab = (op.r[i] % M.L) >= M.n_sites, bb = (op.c[i] % M.L) >= M.n_sites;
        const int t = ab != bb ? 3 : (ab ? 2 : 1);
        mixed_target |= t != op.target;
      }
      if (mixed_target) fail("operator " + op.name + " has no unique target");
    }
    for (size_t q = 0; q < op.v.size(); ++q) {
      const bool cross = (op.r[q] < M.L) != (op.c[q] < M.L);
      if (op.anomalous && !cross) op.mixing |= 2;  // triplet pairing needs both spins
      if (!op.anomalous && !op.interaction && cross) op.mixing |= 2;
    }
    if (op.anomalous) {
      op.mixing |= 1;
      bool any = false;
      for (size_t q = 0; q < op.v.size(); ++q) any |= op.r[q] < op.c[q];
      if (!any) fail("building empty anomalous operator.");
    }
    if (M.idx.count(op.name)) {
      M.ops[M.idx[op.name]] = op;
    } else {
      M.idx[op.name] = M.ops.size();
      M.ops.push_back(op);
    }
  });
}

int qed_ed_new_model_instance(const char* model, int nparam, const char* const* names, const double* values,
                              const char* sectors, uint64_t label) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto it = models().find(model);
    if (it == models().end())
      fail(std::string("The model ") + model + " is not defined and so no model instance based on it is allowed.");
    Model& M = *it->second;
    M.closed = true;
    auto I = std::make_unique<Instance>();
    I->label = label;
    I->model = &M;
    for (int k = 0; k < nparam; ++k)
      if (M.idx.count(names[k])) I->value[names[k]] = values[k];
    if (I->value.empty()) fail("ED model instance (label = " + std::to_string(label) + ") for model " + M.name + " has no parameters!");
    // mixing: normal unless spin-asymmetric one-body terms or S != 0 sectors
    int mixing = 0;
    bool nonzero_anomalous = false;
    for (auto& kv : I->value) {
      const Op& op = M.ops[M.idx[kv.first]];
      if (op.interaction && kv.second != 0.0) I->correlated = true;
      // spin flip and anomalous terms set the mixing whatever their value
      // (model_instance_base.cpp:56-66)
      mixing |= op.mixing;
      if (op.anomalous) nonzero_anomalous = true;
      // spin-asymmetric one-body terms (check_spin_symmetry); spin-flip terms are
      // counted above
      if (op.interaction || op.anomalous || (op.mixing & 2) || kv.second == 0.0) continue;
      std::map<std::pair<int, int>, double> up, dn;
      for (size_t q = 0; q < op.v.size(); ++q) {
        if (op.r[q] < M.L) up[{op.r[q], op.c[q]}] += op.v[q];
        else dn[{op.r[q] - M.L, op.c[q] - M.L}] += op.v[q];
      }
      if (up != dn) mixing |= 4;
    }
    if (mixing > 5) mixing &= 3;
    if (mixing == 5) mixing = 1;
    if (n_mixed(mixing) == 4 && !nonzero_anomalous)
      fail("cluster anomalous operators cannot all be zero in the full Nambu case.");
    // a nontrivial group needs a format other than 'F' (model_instance_base.cpp:86-88);
    // every other format is served by the device's symmetry-adapted product
    if (M.symmetric() && gc("Hamiltonian_format") == 'F')
      fail("the Hamiltonian format 'factorized' is not compatible with a nontrivial point group symmetry");
    if (spinor_mixing(mixing) && gc("Hamiltonian_format") == 'F')
      fail("the Hamiltonian format 'factorized' is not compatible with Hilbert space mixing " +
           std::to_string(mixing));
    if (spinor_mixing(mixing) && M.symmetric())
      fail("spin-flip and anomalous terms together with a nontrivial point group are not supported by this solver");
    for (auto& s : split(sectors, '/')) {
      if (s.empty()) continue;
      const SectorLabel l = parse_sector_label(s, M.group.g);
      if (!sec_code(l.S) && l.S != 0 && mixing == 0) mixing = 4;
      // src_ed/model_instance_base.cpp:112-125
      if (sec_code(l.S) && !(mixing & 2))
        fail("sector string " + s + " defines a non conserved spin, but spin is conserved in the model (instance " +
             std::to_string(label) + ")");
      if (sec_code(l.N) && !(mixing & 1))
        fail("sector string " + s + " defines a non conserved particle number, but particle number is conserved "
             "in the model (instance " + std::to_string(label) + ")");
      if (!sec_code(l.S) && (mixing & 2))
        fail("sector string " + s + " defines a conserved spin, but spin is not conserved in the model (instance " +
             std::to_string(label) + ")");
      if (!sec_code(l.N) && (mixing & 1))
        fail("sector string " + s + " defines a conserved particle number, but particle number is not conserved "
             "in the model (instance " + std::to_string(label) + ")");
      I->sectors.push_back({l.N, l.S});
      I->irreps.push_back(l.R);
    }
    if (I->sectors.empty()) fail("no target sectors were specified for model " + M.name);
    if (M.mixing == -1) M.mixing = mixing;
    else if ((mixing | M.mixing) != M.mixing)
      fail("The new instance mixing state of model " + M.name + " (" + std::to_string(mixing) +
           ") is not compatible with the first instance mixing (" + std::to_string(M.mixing) + ")");
    else mixing = M.mixing;
    I->mixing = mixing;
    set_hopping_matrix(*I);
    instances()[label] = std::move(I);
  });
}

int qed_ed_erase_model_instance(uint64_t label) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { instances().erase(label); });
}

int qed_ed_ground_state_solve(uint64_t label, double* E0, char* buf, int buflen) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    low_energy_states(I);
    *E0 = I.GS_energy;
    if (buf && buflen > 0) {
      std::strncpy(buf, I.gs_string.c_str(), buflen - 1);
      buf[buflen - 1] = 0;
    }
  });
}

int qed_ed_green_function_solve(uint64_t label) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { green_function_solve(inst(label)); });
}

int qed_batch_last_timings(double* t, int n) {
  return guard([&] {
    const BatchStats& s = last_batch_stats();
    const double v[9] = {s.t_setup, s.t_gs, s.t_gf_setup, s.t_band, s.t_band_checks, s.t_assemble,
                         (double)s.hv_config.mode, (double)s.band_checks, (double)s.band_dense_checks};
    for (int i = 0; i < n && i < 9; ++i) t[i] = v[i];
  });
}

int qed_ed_batch_solve_instances(int n, const uint64_t* labels, int graph, int chunk, int* stats) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    std::vector<Instance*> v;
    for (int i = 0; i < n; ++i) v.push_back(&inst(labels[i]));
    BatchStats st;
    batch_solve(v, graph != 0, chunk > 0 ? chunk : 16, &st);
    if (stats) {
      stats[0] = st.lanczos_iterations;
      stats[1] = st.band_steps;
      stats[2] = st.hv_config.block;
      stats[3] = (int)st.hv_config.grid;
    }
  });
}

int qed_batch_solve(const char* model, int nparam, const char* const* names, int nb, const double* params,
                    const char* const* sectors, int graph, int cap, double* E0, int* Mout, double* e,
                    double* Q) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto it = models().find(model);
    if (it == models().end()) fail(std::string("The model ") + model + " is not defined");
    const int d = it->second->n_sites;
    // temporary instances under labels no user instance takes
    const uint64_t base = 0xB000000000000000ull;
    std::vector<uint64_t> labels(nb);
    for (int b = 0; b < nb; ++b) {
      labels[b] = base + (uint64_t)b;
      instances().erase(labels[b]);
      if (qed_ed_new_model_instance(model, nparam, names, params + (size_t)b * nparam, sectors[b], labels[b]) != 0)
        fail(qed_last_error());
    }
    struct Cleanup {
      std::vector<uint64_t>& l;
      ~Cleanup() {
        for (uint64_t x : l) instances().erase(x);
      }
    } cleanup{labels};
    std::vector<Instance*> v;
    for (uint64_t x : labels) v.push_back(&inst(x));
    BatchStats st;
    batch_solve(v, graph != 0, 16, &st);
    for (int b = 0; b < nb; ++b) {
      Instance& I = *v[b];
      E0[b] = I.GS_energy;
      int m = 0;
      for (auto& s : I.states) {
        const QMat& q = s->gf[0];
        for (int k = 0; k < q.M(); ++k) {
          if (m >= cap) fail("qed_batch_solve: more poles than the output capacity");
          e[(size_t)b * cap + m] = q.e[k];
          for (int i = 0; i < d; ++i) Q[((size_t)b * cap + m) * d + i] = q.Q[i + (size_t)k * d];
          ++m;
        }
      }
      Mout[b] = m;
    }
  });
}

int qed_ed_green_function(uint64_t label, int nz, const double* z, int spin_down, double* Gout) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { green_function(inst(label), nz, z, spin_down, Gout); });
}

int qed_ed_hybridization_function(uint64_t label, int nz, const double* z, int spin_down, double* Gam) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { hybridization(inst(label), nz, z, spin_down, Gam); });
}

int qed_ed_self_energy(uint64_t label, int nz, const double* z, int spin_down, double* S) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    const int d = dim_gf(I);
    std::vector<double> G((size_t)2 * nz * d * d), H((size_t)2 * nz * d * d);
    green_function(I, nz, z, spin_down, G.data());
    hybridization(I, nz, z, spin_down, H.data());
    const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
    for (int iz = 0; iz < nz; ++iz) {
      std::vector<cd> g((size_t)d * d);
      for (size_t q = 0; q < (size_t)d * d; ++q) g[q] = cd(G[2 * ((size_t)iz * d * d + q)], G[2 * ((size_t)iz * d * d + q) + 1]);
      if (!cinverse(d, g.data())) fail("cluster Green function is singular");
      cd w(z[2 * iz], z[2 * iz + 1]);
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
          size_t q = a + (size_t)b * d;
          cd s = -(g[q] + I.tc[sp][q] + (I.model->n_bath ? cd(H[2 * ((size_t)iz * d * d + q)], H[2 * ((size_t)iz * d * d + q) + 1]) : cd(0.0)));
          if (a == b) s += w;
          S[2 * ((size_t)iz * d * d + q)] = s.real();
          S[2 * ((size_t)iz * d * d + q) + 1] = s.imag();
        }
    }
  });
}

int qed_ed_hopping_matrix(uint64_t label, int spin_down, double* tc) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
    const int d = dim_gf(I);
    for (size_t q = 0; q < (size_t)d * d; ++q) {
      tc[2 * q] = I.tc[sp][q];
      tc[2 * q + 1] = 0.0;
    }
  });
}

int qed_ed_hopping_matrix_full(uint64_t label, int spin_down, int diag, double* H) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // model_instance::hopping_matrix_full (src_ed/model_instance.hpp:1491-1525): the
    // one-body matrix of cluster and bath, [[t_c, t_cb], [t_cb^T, t_b]], with the bath in
    // its eigenbasis (diag) or as the operators give it
    Instance& I = inst(label);
    const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
    const int dg = dim_gf(I), dbg = dg / I.model->n_sites * I.model->n_bath, d = dg + dbg;
    std::vector<double> F((size_t)d * d, 0.0);
    for (int a = 0; a < dg; ++a)
      for (int b = 0; b < dg; ++b) F[a + (size_t)b * d] = I.tc[sp][a + (size_t)b * dg];
    const std::vector<double>& tcb = diag ? I.tcb_d[sp] : I.tcb[sp];
    for (int a = 0; a < dg; ++a)
      for (int b = 0; b < dbg; ++b) {
        F[a + (size_t)(dg + b) * d] = tcb[a + (size_t)b * dg];
        F[(dg + b) + (size_t)a * d] = tcb[a + (size_t)b * dg];
      }
    for (int a = 0; a < dbg; ++a)
      for (int b = 0; b < dbg; ++b)
        F[(dg + a) + (size_t)(dg + b) * d] = diag ? (a == b ? I.eb[sp][a] : 0.0) : I.tb[sp][a + (size_t)b * dbg];
    for (size_t q = 0; q < F.size(); ++q) {
      H[2 * q] = F[q];
      H[2 * q + 1] = 0.0;
    }
  });
}

int qed_ed_bath(uint64_t label, int spin_down, double* eb, double* tcb) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
    std::copy(I.eb[sp].begin(), I.eb[sp].end(), eb);
    std::copy(I.tcb_d[sp].begin(), I.tcb_d[sp].end(), tcb);
  });
}

int qed_ed_qmatrix(uint64_t label, int spin_down, int* M, double* e, double* Q) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    green_function_solve(I);
    if ((I.cf || I.mcf || I.q_to_mcf) && I.correlated)
      fail("the Q-matrix is not available with the continued-fraction Green function");
    const int sp = spin_index(I, spin_down);
    const int d = dim_gf(I);
    int tot = 0;
    for (auto& s : I.states) tot += s->gf[sp].M();
    *M = tot;
    if (!e) return;
    int off = 0;
    for (auto& s : I.states) {
      const QMat& q = s->gf[sp];
      std::copy(q.e.begin(), q.e.end(), e + off);
      std::copy(q.Q.begin(), q.Q.end(), Q + (size_t)off * d);
      off += q.M();
    }
  });
}

int qed_ed_potthoff_functional(uint64_t label, double* omega) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    green_function_solve(I);
    *omega = I.SEF_bath + I.E0;
  });
}

int qed_ed_model_size(const char* model, int* n_sites, int* n_bath) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto it = models().find(model);
    if (it == models().end()) fail(std::string("model ") + model + " does not exist!");
    *n_sites = it->second->n_sites;
    *n_bath = it->second->n_bath;
  });
}

int qed_ed_mixing(uint64_t label, int* mixing) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { *mixing = inst(label).mixing; });
}

int qed_ed_dim_gf(uint64_t label, int* d) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] { *d = dim_gf(inst(label)); });
}

int qed_ed_cluster_averages(uint64_t label, int nop, const char* const* names, double* ave, double* var) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    low_energy_states(I);
    if (!I.correlated || I.from_file) {
      // from the Green function average (model_instance::cluster_averages,
      // src_ed/model_instance.hpp:152-163; one_body_operator::average_from_GF,
      // src_ed/Operators/one_body_operator.hpp:420-441): one-body operators on the cluster
      // only, no variance
      green_function_solve(I);
      if (spinor_mixing(I.mixing)) {
        // the trace of the integrated spinor G with the operator's spinor matrix E
        // (one_body_operator::average_from_GF, anomalous_operator::average_from_GF), plus
        // the Nambu correction of the diagonal one-body elements, halved for the full
        // mixing (model_instance::cluster_averages, src_ed/model_instance.hpp:144-163)
        const Model& Md = *I.model;
        const int dg = dim_gf(I), dbg = dg / Md.n_sites * Md.n_bath;
        const auto Mg = gf_average(I, 0);
        for (int k = 0; k < nop; ++k) {
          std::string nm(names[k]);
          if (!I.value.count(nm)) fail("operator " + nm + " is not activated in instance " + std::to_string(label));
          const Op& op = Md.ops[Md.idx.at(nm)];
          ave[k] = var[k] = std::nan("");
          if (op.interaction || op.target != 1) continue;
          std::vector<double> E((size_t)dg * dg, 0.0), Eb((size_t)dbg * dbg, 0.0), Ecb((size_t)dg * dbg, 0.0);
          add_spinor_op(Md, op, 1.0, I.mixing, E, Eb, Ecb);
          double a = 0.0;
          for (size_t q = 0; q < E.size(); ++q) a += Mg[q].real() * E[q];
          double corr = 0.0;
          if (!op.anomalous)
            for (size_t e = 0; e < op.v.size(); ++e)
              if (op.r[e] == op.c[e] && (I.mixing == 3 || op.r[e] < Md.L)) corr += op.v[e];
          if (I.mixing == 1) a += corr;
          if (I.mixing == 3) a = 0.5 * (a + corr);
          double norm = 1.0 / Md.n_sites;
          if (nm == "mu") norm *= -1.0;
          ave[k] = a * norm;
        }
        return;
      }
      const int d = I.model->n_sites, L = I.model->L;
      const int nsp = (I.mixing & 4) ? 2 : 1;
      std::vector<std::vector<cd>> M;
      for (int sp = 0; sp < nsp; ++sp) M.push_back(gf_average(I, sp));
      for (int k = 0; k < nop; ++k) {
        std::string nm(names[k]);
        if (!I.value.count(nm)) fail("operator " + nm + " is not activated in instance " + std::to_string(label));
        const Op& op = I.model->ops[I.model->idx.at(nm)];
        ave[k] = var[k] = std::nan("");
        if (op.interaction || op.target != 1) continue;
        double a = 0.0;
        for (int sp = 0; sp < nsp; ++sp)
          for (size_t e = 0; e < op.v.size(); ++e) {
            int r = op.r[e], c = op.c[e];
            if ((r >= L) != (sp == 1) || (c >= L) != (sp == 1)) continue;
            r -= sp * L;
            c -= sp * L;
            if (r >= d || c >= d) continue;
            a += op.v[e] * M[sp][r + (size_t)c * d].real();
            if (r != c) a += op.v[e] * M[sp][c + (size_t)r * d].real();
          }
        if (I.mixing == 0) a *= 2.0;
        double norm = 1.0 / d;
        if (nm == "mu") norm *= -1.0;
        ave[k] = a * norm;
      }
      return;
    }
    for (int k = 0; k < nop; ++k) {
      std::string nm(names[k]);
      if (!I.value.count(nm)) fail("operator " + nm + " is not activated in instance " + std::to_string(label));
      const Op& op = I.model->ops[I.model->idx.at(nm)];
      double a = 0.0, a2 = 0.0;
      for (auto& s : I.states) {
        SectorPtr H = state_sector(I, *s, &nm);
        DevBuf y(H->D * 8);
        QED_HIP(hipMemsetAsync(y.p, 0, H->D * 8, stream()));
        hv(*H, s->psi.as<double>(), y.as<double>());
        a += s->weight * dot(y.as<double>(), s->psi.as<double>(), H->D);
        a2 += s->weight * dot(y.as<double>(), y.as<double>(), H->D);
      }
      double v = a2 - a * a;
      if (op.target == 1) {
        double norm = 1.0 / I.model->n_sites;
        if (nm == "mu") norm *= -1.0;
        a *= norm;
        v *= norm * norm;
      }
      ave[k] = a;
      var[k] = v;
    }
  });
}

int qed_ed_green_function_average(uint64_t label, int spin_down, double* M) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    green_function_solve(I);
    const auto m = gf_average(I, spin_index(I, spin_down));
    for (size_t q = 0; q < m.size(); ++q) {
      M[2 * q] = m[q].real();
      M[2 * q + 1] = m[q].imag();
    }
  });
}

int qed_ed_green_function_density(uint64_t label, double* density) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& I = inst(label);
    green_function_solve(I);
    // model_instance::Green_function_density (src_ed/model_instance.hpp:662-675): the
    // spin-down particles of the spin-flip spinor, 1 - <c c+> of the Nambu components
    const int d = I.model->n_sites, dg = dim_gf(I);
    double n = 0.0;
    const auto m = gf_average(I, 0);
    for (int i = 0; i < d; ++i) n += m[i + (size_t)i * dg].real();
    if (I.mixing & 2)
      for (int i = d; i < 2 * d; ++i) n += m[i + (size_t)i * dg].real();
    else if (I.mixing & 1)
      for (int i = d; i < 2 * d; ++i) n += 1.0 - m[i + (size_t)i * dg].real();
    if (I.mixing & 4) {
      const auto md = gf_average(I, 1);
      for (int i = 0; i < d; ++i) n += md[i + (size_t)i * d].real();
    }
    n /= d;
    if (I.mixing == 0) n *= 2.0;
    *density = n;
  });
}

int qed_ed_fidelity(uint64_t label1, uint64_t label2, double* fidelity) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    Instance& A = inst(label1);
    Instance& B = inst(label2);
    low_energy_states(A);
    low_energy_states(B);
    // states of the same sector share a basis: the same (N_up, N_down) for the factorized
    // form; the same irrep or number sector of one model for the symmetry-adapted and
    // mixed-number bases
    const bool special = A.model->symmetric() || spinor_mixing(A.mixing) || B.model->symmetric() ||
                         spinor_mixing(B.mixing);
    if (special && (A.model != B.model || A.mixing != B.mixing))
      fail("the fidelity of symmetry-adapted or mixed-number states needs two instances of one model");
    double z = 0.0;
    for (auto& s1 : A.states)
      for (auto& s2 : B.states) {
        if (s1->nup != s2->nup || s1->ndn != s2->ndn || s1->irrep != s2->irrep || s1->secN != s2->secN ||
            s1->secS != s2->secS)
          continue;
        if (!s1->psi.p || !s2->psi.p) fail("the fidelity needs the ground-state vectors of both instances");
        if (A.model->L != B.model->L) fail("the two instances have different Hilbert spaces");
        auto binom = [](int n, int k) {
          uint64_t c = 1;
          for (int i = 1; i <= k; ++i) c = c * (uint64_t)(n - k + i) / (uint64_t)i;
          return c;
        };
        const uint64_t D = special ? state_sector(A, *s1)->D
                                   : binom(A.model->L, s1->nup) * binom(A.model->L, s1->ndn);
        const double zz = dot(s1->psi.as<double>(), s2->psi.as<double>(), D);
        z += zz * zz * s1->weight * s2->weight;
      }
    *fidelity = z;
  });
}

int qed_ed_tr_sigma_inf(uint64_t label, double* value) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // model_instance::tr_sigma_inf (src_ed/model_instance.hpp:1279-1311): with the Nambu
    // spinor, the hole components enter with the opposite sign
    Instance& I = inst(label);
    low_energy_states(I);
    const int d = dim_gf(I), d2 = (I.mixing & 1) ? d / 2 : d;
    std::vector<double> eps(3), S(3, 0.0), z(6), sig((size_t)2 * 3 * d * d);
    for (int i = 0; i < 3; ++i) {
      eps[i] = (i + 1) * 1.0e-5;
      z[2 * i] = 0.0;
      z[2 * i + 1] = 1.0 / eps[i];
    }
    for (int spin = 0; spin < ((I.mixing & 4) ? 2 : 1); ++spin) {
      if (qed_ed_self_energy(label, 3, z.data(), spin, sig.data()) != 0) fail(qed_last_error());
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < d; ++j) S[i] += (j < d2 ? 1.0 : -1.0) * sig[2 * ((size_t)i * d * d + j + (size_t)j * d)];
    }
    if (I.mixing == 0)
      for (double& v : S) v *= 2.0;
    *value = neville(eps, S, 0.0);
  });
}

int qed_ed_hybridization_lehmann(uint64_t label, int spin_down, double* E, double* T) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // model_instance::hybridization (src_ed/model_instance.hpp:1455-1466): the bath as
    // set_hopping_matrix leaves it, diagonalized in place (energies, and the cluster-bath
    // matrix in the bath's eigenbasis), of the spinor's dimensions for the spinor mixings
    Instance& I = inst(label);
    const int nb = I.model->n_bath;
    if (nb == 0) fail("System has no bath. Hybridization not defined.");
    low_energy_states(I);
    const int sp = (spin_down && (I.mixing & 4)) ? 1 : 0;
    const int dg = dim_gf(I), dbg = dg / I.model->n_sites * nb;
    for (int i = 0; i < dbg; ++i) E[i] = I.eb[sp][i];
    std::copy(I.tcb_d[sp].begin(), I.tcb_d[sp].begin() + (size_t)dg * dbg, T);
  });
}

int qed_ed_print_wavefunction(uint64_t label, char* buf, int buflen, int* needed) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // model_instance::print_wavefunction, src_ed/model_instance.hpp:1472-1485, and
    // state::write_wavefunction, src_ed/state.hpp:117-131: per state a header line, then,
    // when the sector dimension is at most max_dim_print, |psi_i|^2, psi_i and the basis
    // state, lowest orbital first; for the factorized format 'F' the down-spin then the
    // up-spin occupations (ED_factorized_basis::print_state, src_ed/ED_basis.cpp:167-170;
    // PrintBinaryDouble, src_ed/binary_state.cpp:104-112).
    // The operator formats (not 'F') print through the reference's mixed basis
    // (ED_mixed_basis::print_state, src_ed/ED_basis.cpp:391-399, trivial group): the state
    // b = up | down << 32 as PrintBinaryDouble writes it, the up-spin occupations first.
    Instance& I = inst(label);
    require_trivial_group(I, "print_wavefunction");
    low_energy_states(I);
    const int L = I.model->L;
    const bool factorized = gc("Hamiltonian_format") == 'F' && !spinor_mixing(I.mixing);
    std::ostringstream o;
    auto bits = [&](uint32_t v) {
      std::string t(L, '0');
      for (int k = 0; k < L; ++k) t[k] = (v >> k) & 1u ? '1' : '0';
      return t;
    };
    for (auto& st : I.states) {
      const bool sp = spinor_mixing(I.mixing);
      o << "state\n" << (sp ? sector_name(st->secN, st->secS) : sector_name(st->nup + st->ndn, st->nup - st->ndn))
        << '\t' << st->energy << '\t' << st->weight << std::endl;
      if (!st->psi.p) continue;
      SectorPtr H = state_sector(I, *st);
      if (H->D > (uint64_t)gi("max_dim_print")) continue;
      std::vector<double> psi(H->D);
      QED_HIP(hipMemcpy(psi.data(), st->psi.p, H->D * 8, hipMemcpyDeviceToHost));
      std::vector<uint64_t> basis(H->D);
      if (sp) {
        uint64_t n = 0;
        if (qed_general_basis(H.get(), basis.data(), &n) != 0) fail(qed_last_error());
      } else {
        std::vector<uint32_t> bu(H->du), bd(H->dd);
        QED_HIP(hipMemcpy(bu.data(), H->bup.p, H->du * 4, hipMemcpyDeviceToHost));
        QED_HIP(hipMemcpy(bd.data(), H->bdn.p, H->dd * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < H->D; ++i) basis[i] = (uint64_t)bu[i % H->du] | ((uint64_t)bd[i / H->du] << 32);
      }
      for (uint64_t i = 0; i < H->D; ++i) {
        const uint32_t up = (uint32_t)basis[i], dn = (uint32_t)(basis[i] >> 32);
        o << std::abs(psi[i]) * std::abs(psi[i]) << '\t' << psi[i] << '\t' << '|'
          << (factorized ? bits(dn) : bits(up)) << ',' << (factorized ? bits(up) : bits(dn)) << '>' << '\n';
      }
    }
    const std::string out = o.str();
    *needed = (int)out.size() + 1;
    if (buf && buflen > 0) {
      std::strncpy(buf, out.c_str(), buflen - 1);
      buf[buflen - 1] = 0;
    }
  });
}

int qed_ed_get_combined_mcf(uint64_t label, int spin_down, int* p, int* M, int* q, double* A, double* B,
                            double* W) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // ED::get_combined_mcf, src_ed/qcm_ED.cpp:455-540: the first state's combined fraction
    Instance& I = inst(label);
    if (!I.gf_solved) set_gf_format(I);
    if (!I.mcf && !I.q_to_mcf)
      fail("get_combined_mcf: the instance at label " + std::to_string(label) +
           " does not use the MCF format (GF_method must be 'M', or 'L' with combine_mcf=True).");
    if (!gb("combine_mcf"))
      fail("get_combined_mcf: the combined MCF is not built unless the global parameter 'combine_mcf' is set to True.");
    green_function_solve(I);
    if (I.states.empty()) fail("get_combined_mcf: no valid ground state found.");
    const MCFrac& f = I.states.front()->mcf_comb[spin_index(I, spin_down)];
    if (f.M() == 0) fail("get_combined_mcf: combined MCF is empty. Set combine_mcf=True before solving.");
    *p = f.pa;
    *M = f.M();
    *q = f.q;
    if (A) std::copy(f.A.begin(), f.A.end(), A);
    if (B) std::copy(f.B.begin(), f.B.end(), B);
    if (W) std::copy(f.W.begin(), f.W.end(), W);
  });
}

int qed_ed_density_matrix(uint64_t label, int nsites, const int* sites, int* dim, double* rho, uint64_t* states) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    // ED::density_matrix (src_ed/qcm_ED.cpp:205-216): empty for the factorized format
    // (density_matrix_factorized), otherwise model_instance::density_matrix_mixed
    // (src_ed/model_instance.hpp:1547-1625) on the first state: the basis states b
    // (up | down << 32, in basis order) split into b & mask_A and the rest, subsystem
    // states numbered by first appearance, rho(a, a') = sum over the rest of
    // psi(a, B) psi(a', B), and the subsystem states collapsed onto bits 0.. of each half
    Instance& I = inst(label);
    *dim = 0;
    if (gc("Hamiltonian_format") == 'F' && !spinor_mixing(I.mixing)) return;
    if (I.model->symmetric()) fail("computing the density matrix requires symmetries to be turned off!");
    low_energy_states(I);
    if (I.states.empty() || !I.states[0]->psi.p) fail("the density matrix needs the ground-state vector");
    const State& st = *I.states[0];
    SectorPtr H = state_sector(I, st);
    std::vector<uint64_t> basis(H->D);
    if (spinor_mixing(I.mixing)) {
      uint64_t n = 0;
      if (qed_general_basis(H.get(), basis.data(), &n) != 0) fail(qed_last_error());
    } else {
      std::vector<uint32_t> bu(H->du), bd(H->dd);
      QED_HIP(hipMemcpy(bu.data(), H->bup.p, H->du * 4, hipMemcpyDeviceToHost));
      QED_HIP(hipMemcpy(bd.data(), H->bdn.p, H->dd * 4, hipMemcpyDeviceToHost));
      for (uint64_t k = 0; k < H->D; ++k) basis[k] = (uint64_t)bu[k % H->du] | ((uint64_t)bd[k / H->du] << 32);
    }
    std::vector<double> psi(H->D);
    QED_HIP(hipMemcpy(psi.data(), st.psi.p, H->D * 8, hipMemcpyDeviceToHost));
    uint64_t mask = 0;
    for (int i = 0; i < nsites; ++i) {
      if (sites[i] < 0 || sites[i] >= I.model->L) fail("density_matrix: site out of range");
      mask |= 1ull << sites[i];
    }
    mask |= mask << 32;
    std::map<uint64_t, int> ia, ib;
    std::vector<uint64_t> binA;
    std::vector<std::vector<std::pair<int, uint64_t>>> listA;  // per B: (a, basis index)
    for (uint64_t k = 0; k < H->D; ++k) {
      const uint64_t A = basis[k] & mask, B = basis[k] ^ A;
      auto a = ia.find(A);
      if (a == ia.end()) {
        a = ia.emplace(A, (int)binA.size()).first;
        binA.push_back(A);
      }
      auto b = ib.find(B);
      if (b == ib.end()) {
        b = ib.emplace(B, (int)listA.size()).first;
        listA.emplace_back();
      }
      listA[b->second].push_back({a->second, k});
    }
    const int dA = (int)binA.size();
    *dim = dA;
    if (!rho) return;
    std::fill(rho, rho + (size_t)2 * dA * dA, 0.0);
    for (const auto& v : listA)
      for (const auto& x : v)
        for (const auto& y : v) rho[2 * (x.first + (size_t)y.first * dA)] += psi[x.second] * psi[y.second];
    for (int a = 0; a < dA; ++a) {
      uint64_t c = 0;
      for (int i = 0; i < nsites; ++i) {
        if (binA[a] & (1ull << sites[i])) c |= 1ull << i;
        if (binA[a] & (1ull << (sites[i] + 32))) c |= 1ull << (i + 32);
      }
      states[a] = c;
    }
  });
}

int qed_ed_susceptibility(uint64_t label, const char* op, int nz, const double* z, double* chi) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto parts = susceptibility_parts(inst(label), op);
    for (int j = 0; j < nz; ++j) {
      const cd w(z[2 * j], z[2 * j + 1]);
      cd c = 0.0;
      for (auto& p : parts) {
        cd g = 0.0;  // Q_matrix::Green_function for one start vector
        for (int i = 0; i < p.q.M(); ++i) g += p.q.Q[i] * p.q.Q[i] / (w - p.q.e[i]);
        c += p.weight * g;
      }
      chi[2 * j] = c.real();
      chi[2 * j + 1] = c.imag();
    }
  });
}

int qed_ed_susceptibility_poles(uint64_t label, const char* op, int cap, int* n, double* poles) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto parts = susceptibility_parts(inst(label), op);
    std::vector<std::pair<double, double>> chi, merged;
    for (auto& p : parts)
      for (int i = 0; i < p.q.M(); ++i) {
        if (p.q.e[i] < 1e-8) continue;
        const double r = std::abs(p.q.Q[i]) * std::abs(p.q.Q[i]) * p.weight;
        if (r > 1e-6) chi.push_back({p.q.e[i], r});
      }
    // consolidation: a pole within 1e-4 of a kept one adds its residue to it
    for (auto& c : chi) {
      size_t j = 0;
      for (; j < merged.size(); ++j)
        if (std::fabs(merged[j].first - c.first) < 1e-4) {
          merged[j].second += c.second;
          break;
        }
      if (j == merged.size()) merged.push_back(c);
    }
    *n = (int)merged.size();
    for (int k = 0; k < std::min(cap, *n); ++k) {
      poles[2 * k] = merged[k].first;
      poles[2 * k + 1] = merged[k].second;
    }
  });
}

int qed_ed_write_instance_hdf5(const char* filename, uint64_t label, const char* group) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto it = instances().find(label);
    if (it == instances().end()) fail("The label " + std::to_string(label) + " is out of range.");
    write_instance_file(*it->second, filename, group);
  });
}

int qed_ed_read_instance_hdf5(const char* filename, uint64_t label, const char* group) {
  std::lock_guard<std::recursive_mutex> lk(lib_mutex());
  return guard([&] {
    auto it = instances().find(label);
    if (it == instances().end()) fail("The label " + std::to_string(label) + " is out of range.");
    read_instance_file(*it->second, filename, group);
  });
}

int qed_ed_set_global_int(const char* name, int v) {
  return guard([&] { G().i[name] = v; });
}
int qed_ed_set_global_double(const char* name, double v) {
  return guard([&] { G().d[name] = v; });
}
int qed_ed_set_global_bool(const char* name, int v) {
  return guard([&] { G().b[name] = v != 0; });
}
int qed_ed_set_global_char(const char* name, char v) {
  return guard([&] { G().c[name] = v; });
}
int qed_ed_get_global_double(const char* name, double* v) {
  return guard([&] {
    auto it = G().d.find(name);
    if (it == G().d.end()) fail(std::string("unknown global parameter ") + name);
    *v = it->second;
  });
}

}  // extern "C"
