// One factorized sector (N_up, N_dn) of a cluster Hamiltonian resident on the device.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "qed_common.hpp"

struct qed_sector {
  int L = 0, nup = 0, ndn = 0;
  uint32_t du = 0, dd = 0;  // half-basis dimensions (up = fast index)
  uint64_t D = 0;           // du * dd

  // half bases (ascending bit strings) and hop tables per spin, CSR over half states:
  // ptr[i]..ptr[i+1] index packed entries (target rank | value index << 20)
  qed::DevBuf bup, bdn;
  qed::DevBuf up_ptr, up_ent, dn_ptr, dn_ent;
  qed::DevBuf up_vals, dn_vals;   // distinct signed amplitudes (double)
  qed::DevBuf diag_up, diag_dn;   // one-body diagonal per half state
  uint32_t up_nnz = 0, dn_nnz = 0, up_nvals = 0, dn_nvals = 0;
  uint32_t up_maxdeg = 0, dn_maxdeg = 0;

  // density-density interaction. Word layout of the reference's factorized basis:
  // w = dn_state | up_state << 32; element (r, c) active when both mask bits are set.
  std::vector<uint64_t> int_mask;
  std::vector<double> int_val;
  qed::DevBuf int_mask_d, int_val_d;
  bool hubbard_uniform = true;  // all elements (i, i+L) with one value -> U * popcount
  double hub_U = 0.0;
  uint32_t hub_sites = 0;
  std::vector<double> int_diag_host;  // unused unless requested

  // ELL copies of the hop tables, padded to multiples of 8 entries with zero-amplitude
  // self hops: up column-major (entry k of column c at k*du + c), down row-major
  qed::DevBuf up_ell, dn_ell;
  uint32_t up_w = 0, dn_w = 0;

  int hv_variant = -1;  // -1 = automatic
  // row visiting order of the tiled-order row kernel (built on its first use)
  mutable qed::DevBuf row_order;

  // the structure tables (bases, hop lists, ELL) are views into cached tables shared by
  // every sector with the same half-space structure; the amplitudes are this sector's own
  std::shared_ptr<void> keep_up, keep_dn;
  qed::DevBuf valbuf;  // owns up_vals, dn_vals, diag_up, diag_dn, int_val_d

  // one-body terms per spin for the split-halves H.v: hop elements (i, j, value) as read
  // from the operators, and the one-body diagonal per orbital
  std::vector<int> pu_i, pu_j, pd_i, pd_j;
  std::vector<double> pu_t, pd_t;
  std::vector<double> dsite_up, dsite_dn;
  mutable std::shared_ptr<void> split;  // plan and device tables, built on first use
  // symmetry-adapted (mixed) basis of one irreducible representation: the group tables,
  // the representatives and the factorized sector they live in (null: factorized sector)
  std::shared_ptr<void> mixed;
  // basis of a sector that conserves N or S but not both (spin-flip and anomalous
  // mixings): the sorted 64-bit states and the Hamiltonian's terms (null otherwise)
  std::shared_ptr<void> general;
  bool factorized() const { return !mixed && !general; }
};

namespace qed {

// Kernel-facing view of the sector (passed by value).
struct HvView {
  const uint32_t* bup;
  const uint32_t* bdn;
  const uint32_t* up_ptr;
  const uint32_t* up_ent;
  const uint32_t* dn_ptr;
  const uint32_t* dn_ent;
  const double* up_vals;
  const double* dn_vals;
  const double* diag_up;
  const double* diag_dn;
  const uint64_t* int_mask;
  const double* int_val;
  uint32_t du, dd;
  uint32_t up_nvals;
  int n_int;
  int hubbard_uniform;
  double hub_U;
  uint32_t hub_sites;
};

HvView view(const qed_sector& s);

// ELL hop entries: target rank | index of the signed amplitude << kEllTargetBits
constexpr uint32_t kEllTargetBits = 20;

// density-density energy of the factorized basis state (su, sd)
__device__ __forceinline__ double density_density(const HvView& v, uint32_t su, uint32_t sd) {
  if (v.hubbard_uniform) return v.hub_U * (double)__popc(su & sd & v.hub_sites);
  const uint64_t w = (uint64_t)sd | ((uint64_t)su << 32);
  double acc = 0.0;
  for (int k = 0; k < v.n_int; ++k) {
    const uint64_t m = v.int_mask[k];
    if ((w & m) == m) acc += v.int_val[k];
  }
  return acc;
}

// one sector of a batched H.v: y[parity] += H x[parity]
struct HvJob {
  HvView v;
  const uint32_t* up_ell;
  const uint32_t* dn_ell;
  uint32_t Wu, Wd;
  uint32_t row0;  // first tile (down row) of this job in the batch
  uint32_t a0, b0;  // first tiles of this job in the two passes of the tiled form
  const double* x[2];
  double* y[2];
};

struct HvBatch {
  std::vector<HvJob> host;  // fill x / y, then hv_batch_upload
  DevBuf dev;
  DevBuf tile_job;  // job index of every row tile (row kernel)
  HvJob* jobs = nullptr;
  int njobs = 0;
  uint32_t ntiles = 0;
  size_t lds = 0;
  std::vector<uint64_t> sig;  // shape signature (autotuner cache key)
  // two-pass tiled form (short rows): R rows per pass-A tile, W-column panels in pass B
  bool tiled = false;
  int R = 0, W = 0;
  uint32_t ntiles_a = 0, ntiles_b = 0;
  size_t lds_a = 0, lds_b = 0;
};

struct HvBatchConfig {
  int block = 256;
  uint32_t grid = 0;  // 0: one workgroup per tile; else a persistent grid of this size
  float ms = 0.0f;
  int mode = 0;  // 0: one workgroup per row; 1: two-pass tiled form
};

void hv_batch_prepare(HvBatch& B, const std::vector<const qed_sector*>& S);
void hv_batch_upload(HvBatch& B);
// times the candidate launch shapes on this batch (y[0] of every job receives the sums of
// the trial products: pass scratch vectors) and caches the fastest by shape signature
HvBatchConfig hv_batch_tune(const HvBatch& B, const int* active, bool force = false);
void hv_batch(const HvBatch& B, const HvBatchConfig& cfg, int parity, const int* active);

struct BatchStats {
  HvBatchConfig hv_config;
  int lanczos_iterations = 0;  // lockstep iterations of the ground-state pass
  int band_steps = 0;          // synchronized steps of the band Lanczos pass
  int band_checks = 0;         // eigenvalue tests of the band pass, of which dense
  int band_dense_checks = 0;
  // wall time (s) of the phases: sectors + start vectors, ground-state Lanczos,
  // Green-function start vectors, band Lanczos (of which: host eigenvalue tests), assembly
  double t_setup = 0, t_gs = 0, t_gf_setup = 0, t_band = 0, t_band_checks = 0, t_assemble = 0;
};
BatchStats& last_batch_stats();

// Lanczos ground states of many sectors in lockstep; psi[j]: start vector in, ground
// state out; the stopping test runs every `chunk` iterations (graph: replay a captured
// two-iteration hipGraph).
void lanczos_gs_batch(const std::vector<qed_sector*>& S, const std::vector<double*>& psi, double accur,
                      int max_iter, bool use_graph, int chunk, std::vector<double>& E0,
                      std::vector<int>& niter, BatchStats* stats);
// band Lanczos of many sectors (start vectors phi[j]: p[j] x D_j on the device)
void band_lanczos_batch(const std::vector<qed_sector*>& S, const std::vector<const double*>& phi,
                        const std::vector<int>& p, int max_iter_BL, double accur_deflation,
                        double accur_band_lanczos, std::vector<std::vector<double>>& evals,
                        std::vector<std::vector<double>>& Q, BatchStats* stats, double min_gap = -1.0);
void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, double* e0,
                    std::vector<double>& vec);

// y += H x on the library stream (device vectors). variant -1 chooses automatically.
void hv(const qed_sector& s, const double* x, double* y, int variant = -1);
// the split-halves passes (passes: 1 row pass, 2 tile pass, 3 both); false when the
// sector's structure does not suit them (nothing launched)
bool hv_split(const qed_sector& s, const double* x, double* y, int passes = 3);
std::string split_status(const qed_sector& s);
// one fused Lanczos product: y <- lz[0] (H x) + lz[1] y (device scalars) and, per row of
// the vector, the partial sums of <x, y_new> in lz_part (dd doubles); false when the
// sector does not take the fused kernel (nothing launched)
bool hv_lanczos(const qed_sector& s, const double* x, double* y, const double* lz, double* lz_part);
// y += H x on a symmetry-adapted sector (s.mixed set)
void hv_mixed(const qed_sector& s, const double* x, double* y);
// d += the diagonal of H on a symmetry-adapted or mixed-number sector (D doubles)
void sector_diagonal(const qed_sector& s, double* d);
void diag_general(const qed_sector& s, double* d);
// y += H x on a sector of the spin-flip / anomalous mixings (s.general set); variant 0
// locates the targets by binary search, otherwise by lookup tables where they exist
void hv_general(const qed_sector& s, const double* x, double* y, int variant = -1);
// The Abelian group of a list of generators (0-based site images, then bath entries:
// phase labels with bath_irrep, orbital images otherwise), on the host: order g, elements
// (g x N entries) and phi(a, b) in units of 2 pi / g (g x g); throws on invalid generators
// (src_ed/symmetry_group.cpp:36-100, 183-352)
struct GroupTables {
  int g = 1;
  std::vector<std::vector<int>> elements;
  std::vector<int> phi;
};
GroupTables group_tables(int N, int n_sites, const std::vector<std::vector<int>>& gens, bool bath_irrep);

// BLAS-1 on the library stream. dot/nrm2 return host scalars (synchronizing). `work`, where
// given, is device memory for the block partials (kMaxBlocks * 65 doubles); the library's
// own buffer otherwise.
double dot(const double* x, const double* y, uint64_t n, double* work = nullptr);
double nrm2(const double* x, uint64_t n);
void axpy(double a, const double* x, double* y, uint64_t n);
void scal(double a, double* x, uint64_t n);
// r *= ar; q *= aq (one pass)
void scal2(double ar, double* r, double aq, double* q, uint64_t n);
// y += a x and returns ||y|| (one pass, synchronizing)
double axpy_nrm2(double a, const double* x, double* y, uint64_t n, double* work = nullptr);
// multi-dot: out[k] = <x, ys[k]> for k < m (one pass over x)
void multi_dot(const double* x, const double* const* ys, int m, uint64_t n, double* out, double* work = nullptr);
// y -= sum_k c[k] xs[k] (one pass)
void multi_axpy(const double* c, const double* const* xs, int m, double* y, uint64_t n);
// ys[k] -= c[k] x for each k (one pass over x)
void multi_axpy_rev(const double* c, const double* x, double* const* ys, int m, uint64_t n);

// device memory of a Lanczos run supplied by the caller: the two recurrence vectors (D
// doubles each), the BLAS-1 block partials (kRedDoubles), the fused step's scalars (8) and
// row partials (kFusedPartDoubles; the fused step needs max(dd, 2048))
struct LanczosWork {
  double* r;
  double* q;
  double* red;
  double* sc;
  double* part;
};
constexpr size_t kRedDoubles = 2048 * 65 + 512;
constexpr size_t kFusedPartDoubles = 2048;
void lanczos_gs(qed_sector& s, double accur, int max_iter, double* psi, double* E0, int* niter,
                std::vector<double>* alpha, std::vector<double>* beta, const LanczosWork* ws = nullptr);
int band_lanczos(qed_sector& s, const double* phi, int p0, int M0, int max_iter_BL,
                 double accur_deflation, double accur_band_lanczos, std::vector<double>& evals,
                 std::vector<double>& Q, double min_gap = -1.0);
void lanczos_green(qed_sector& s, const double* psi, double accur_cf, int max_iter,
                   std::vector<double>& a, std::vector<double>& b);
// block Lanczos on p orthonormal device vectors q0 (p x D); floors A, B (p x p each,
// column-major): QR blocks if qr, polar (symmetric) blocks otherwise
int block_lanczos(qed_sector& s, const double* q0, int p, int M0, int max_iter_BL, double accur_deflation,
                  double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                  std::vector<std::vector<double>>& B);
// y += Op x on device vectors of the operator's dimension
using BlockOperator = std::function<void(const double*, double*)>;
int block_lanczos_op(uint64_t n, const BlockOperator& op, const double* q0, int p, int M0, int max_iter_BL,
                     double accur_deflation, double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                     std::vector<std::vector<double>>& B);
int block_lanczos_diag(const std::vector<double>& e, const std::vector<double>& phi, int p, int M0, int max_iter_BL,
                       double accur_deflation, double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                       std::vector<std::vector<double>>& B);
int block_lanczos_blocktri(int p, int Me, int Mh, const std::vector<double>& Ablk, const std::vector<double>& Bblk,
                           const std::vector<double>& phi, int M0, int max_iter_BL, double accur_deflation,
                           double accur_band_lanczos, bool qr, std::vector<std::vector<double>>& A,
                           std::vector<std::vector<double>>& B);
// G[z] += W^T F_0(z) W (transposed if transpose) for the matrix continued fraction with M
// floors of pa x pa blocks; W pa x q; host arrays, G q x q complex column-major per z
void mcf_eval(int pa, int q, int M, const double* A, const double* B, const double* W, int nz, const double* z,
              bool transpose, double* G);
// Davidson-Liu for the k lowest states; psi receives them (k x D); verb prints the
// progress every 10 iterations, as the reference's verb_ED
void davidson(qed_sector& s, int k, double accur, int max_iter, std::vector<double>& evals,
              double* psi, int* niter, bool verb = false);
// the global parameter verb_ED
bool verbose_ed();
std::recursive_mutex& lib_mutex();

}  // namespace qed
