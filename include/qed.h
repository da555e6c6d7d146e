/*
 * qed.h — C ABI of the MI355X exact-diagonalization impurity solver.
 *
 * Every entry point returns 0 on success and a nonzero status on failure; the message
 * of the last failure on the calling thread is returned by qed_last_error(). Pointers
 * named *_dev are device (HBM) addresses obtained from qed_alloc(); all others are host
 * addresses. Complex numbers are interleaved (re, im) doubles; matrices are column-major
 * as the reference's matrix<T> (src_util/matrix.hpp). Indices of spin-orbitals are
 * 0-based, spin-down orbital a = orbital + L, as ED::new_operator receives them after
 * the binding's 1-based -> 0-based conversion (src_python/qcm_ED_wrap.hpp:174-215).
 *
 * Each declaration names the reference interface it replaces (file:line in pyqcm).
 */
#ifndef QED_H
#define QED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qed_sector qed_sector;

/* ---- runtime ------------------------------------------------------------------------ */

/* Selects the HIP device and creates the library's stream. Replaces nothing: the
 * reference has no device. Called implicitly by every other entry point. */
int qed_init(int device);
/* Message of the last failure on this thread (src_util/parser.cpp:323-327 qcm_ED_throw). */
const char* qed_last_error(void);
int qed_sync(void);
int qed_alloc(void** ptr_dev, uint64_t bytes);
int qed_free(void* ptr_dev);
int qed_memcpy_h2d(void* dst_dev, const void* src, uint64_t bytes);
int qed_memcpy_d2h(void* dst, const void* src_dev, uint64_t bytes);
int qed_memcpy_d2d(void* dst_dev, const void* src_dev, uint64_t bytes);
int qed_memset(void* ptr_dev, int value, uint64_t bytes);

/* ---- basis (bit-exact with the reference) -------------------------------------------- */

/* Half basis of L orbitals with N particles, ascending = colex order; written on the
 * device by combinatorial-number-system unranking. out may be NULL to query the size.
 * Replaces ED_halfbasis::ED_halfbasis (src_ed/ED_basis.cpp:115-125). */
int qed_halfbasis(int L, int N, uint32_t* out, uint32_t* dim);
/* Ranks of n half states computed on the device (CNS); dim when the popcount is not N.
 * Replaces ED_halfbasis::index (src_ed/ED_basis.cpp:130-135). */
int qed_halfbasis_rank(int L, int N, const uint32_t* states, uint64_t n, uint32_t* ranks);

/* ---- one sector of a cluster Hamiltonian (the Hamiltonian<T> seam) ------------------- */

/* Builds H on the factorized sector (n_up, n_dn) of L orbitals per spin from the active
 * operators of a model: types[k] 0 = one-body, 1 = interaction (density-density),
 * coefs[k] the operator value, nels[k] element counts, and the concatenated element
 * lists (er, ec, ev) with er <= ec. Hermitian completion, sign conventions and the
 * lower-triangle storage follow the reference.
 * Replaces Hamiltonian_Factorized ctor + HS_ops_map
 * (src_ed/Hamiltonian/Hamiltonian_Factorized.hpp:31-87, src_ed/Operators/HS_half_operator.hpp:17-67,
 *  src_ed/Operators/HS_interaction_operator.hpp:221-243). */
int qed_sector_create(int L, int n_up, int n_dn, int n_ops, const int* types, const double* coefs,
                      const int* nels, const int* er, const int* ec, const double* ev,
                      qed_sector** out);
int qed_sector_destroy(qed_sector* s);

/* A sector of the symmetry-adapted (mixed) basis: the representatives of the orbits of an
 * Abelian point group in irreducible representation `irrep`, for fixed N_up, N_down. The
 * operators are given as for qed_sector_create; gens holds n_gen generators of L entries,
 * 0-based: site images for the first L' entries and, with bath_irrep, the phase labels of
 * the bath orbitals (L' = n_sites with bath_irrep, L otherwise). Generators must commute
 * and have order <= 2 (real characters). The sector works with every solver entry point
 * of a sector (qed_hv, qed_lanczos_gs, qed_band_lanczos, qed_davidson, qed_diag, ...)
 * except qed_apply_c. Replaces ED_mixed_basis + symmetry_group + HS_one_body_operator
 * (src_ed/ED_basis.cpp:185-361, src_ed/symmetry_group.cpp:18-454,
 * src_ed/Operators/HS_one_body_operator.hpp:8-40) for the unmixed sectors. */
int qed_mixed_sector_create(int L, int n_up, int n_dn, int n_ops, const int* types, const double* coefs,
                            const int* nels, const int* rows, const int* cols, const double* vals, int n_sites,
                            int n_gen, const int* gens, int bath_irrep, int irrep, qed_sector** out);
/* The basis of a symmetry-adapted sector: *n states (up | down << 32), ascending, written
 * to states unless it is NULL (the reference's ED_mixed_basis::binlist). */
int qed_mixed_basis(const qed_sector* s, uint64_t* states, uint64_t* n);
/* y += c_o x (pm = -1) or c^dagger_o x (pm = +1) between symmetry-adapted sectors, for the
 * symmetric orbital o = sum_a coef[a] c_{a, spin} over the n_sites sites (spin 0 = up):
 * src is x's sector, tgt the one of y (one particle less or more of that spin). Replaces
 * destruction_operator (src_ed/Operators/destruction_operator.cpp:14-33, with Destroy,
 * src_ed/ED_basis.cpp:60-77); the creation operator is its transpose, as the reference
 * applies it (src_ed/model.cpp:160-215). */
int qed_mixed_apply_c(const qed_sector* src, const qed_sector* tgt, int spin, int pm, int n_sites,
                      const double* coef, const double* x_dev, double* y_dev);
/* The group of a symmetry-adapted sector: its order g, its elements (g x L entries: site
 * images, bath phase labels) and phi (g x g, in units of 2 pi / g), as symmetry_group
 * builds them (src_ed/symmetry_group.cpp:183-352); elements and phi may be NULL. */
int qed_mixed_group(const qed_sector* s, int* g, int* elements, int* phi);
/* A sector of the reference's mixed basis for the spin-flip and anomalous mixings
 * (src_ed/ED_basis.cpp:185-361, trivial group): the union of the (N_up, N_down) blocks
 * pairs[2k], pairs[2k+1] (k < npairs), as 64-bit states ascending. Operator types 0
 * one-body (spin-flip elements allowed), 1 density-density interaction, 2 anomalous
 * (elements r < c as ED::new_operator receives them, HS_anomalous_operator,
 * src_ed/Operators/HS_anomalous_operator.hpp:8-40). qed_hv, qed_diag and the Krylov
 * entry points (qed_davidson included) take such a sector; its dim_up and dim_dn are 0. */
int qed_general_sector_create(int L, int npairs, const int* pairs, int n_ops, const int* types,
                              const double* coefs, const int* nels, const int* rows, const int* cols,
                              const double* vals, qed_sector** out);
/* The basis of such a sector (64-bit states up | down << 32, ascending), as
 * qed_mixed_basis. */
int qed_general_basis(const qed_sector* s, uint64_t* states, uint64_t* n);
/* y += z c_a x (pm = -1) or z c^dagger_a x (pm = +1) between two such sectors, spin
 * orbital a (down orbitals at L + i), with the Jordan-Wigner sign of the 64-bit word
 * (Destroy, src_ed/ED_basis.cpp:60-77; the creation operator is its transpose). */
int qed_general_apply(const qed_sector* src, const qed_sector* tgt, int orbital, int pm, double z,
                      const double* x_dev, double* y_dev);
int qed_sector_dims(const qed_sector* s, uint64_t* dim, uint32_t* dim_up, uint32_t* dim_dn);

/* y_dev += H x_dev. Replaces Hamiltonian_Factorized::mult_add
 * (src_ed/Hamiltonian/Hamiltonian_Factorized.hpp:50-58). */
int qed_hv(qed_sector* s, const double* x_dev, double* y_dev);
/* Same contract on host vectors (copies in and out), the std::vector form of mult_add. */
int qed_hv_host(qed_sector* s, const double* x, double* y);
/* d_dev += diagonal of H as the reference's factorized format reports it: interaction
 * terms only (HS_factorized_operator::diag is empty, src_ed/Operators/HS_factorized_operator.hpp:103).
 * Replaces Hamiltonian_Factorized::diag (src_ed/Hamiltonian/Hamiltonian_Factorized.hpp:65-68).
 * On a symmetry-adapted or mixed-number sector, the whole diagonal <I|H|I>, one-body terms
 * included, as the operator formats report it (Hamiltonian_CSR::diag,
 * src_ed/Hamiltonian/Hamiltonian_CSR.hpp:103-105; HS_nondiagonal_operator::diag,
 * src_ed/Operators/HS_nondiagonal_operator.hpp:176). */
int qed_diag(qed_sector* s, double* d_dev);

/* The reference's start vector: x = N(0,1) deviates of std::default_random_engine(seed)
 * through the process-wide normal_distribution, normalized; written to x_dev.
 * Replaces random() (src_util/vector_num.cpp:114-119). */
int qed_start_vector(qed_sector* s, int seed, double* x_dev);

/* Lanczos ground state: eigenvalue pass then eigenvector replay. psi_dev holds the
 * normalized start vector on entry and the ground state on exit. Fails with the
 * reference's message beyond max_iter iterations. alpha/beta (may be NULL) receive the
 * tridiagonal coefficients (capacity max_iter + 1).
 * Replaces Lanczos / LanczosEigenvalue / LanczosEigenvector (src_ed/Hamiltonian/Lanczos.hpp:12-143). */
int qed_lanczos_gs(qed_sector* s, double accur, int max_iter, double* psi_dev, double* E0,
                   int* niter, double* alpha, double* beta);

/* qed_lanczos_gs without device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (tight solver loops, harnesses that forbid hipMalloc
 * after start-up). Same arguments and results as qed_lanczos_gs; the two recurrence
 * vectors and the reduction partials live in a fixed device array of the library (one
 * per GPU, 34.7 MB, held until the process ends), so the sector dimension is limited to
 * 2^21 = 2,097,152: a larger sector fails with an error and nothing is launched. All calls
 * on one GPU share that array, which keeps the last call's contents; the library lock
 * serializes the calls of one process, and calls that overlap on one GPU otherwise race.
 * A sector set to a split-halves H.v variant takes the default row kernel here. */
int qed_lanczos_gs_noalloc(qed_sector* s, double accur, int max_iter, double* psi_dev, double* E0,
                           int* niter, double* alpha, double* beta);

/* Scalar Lanczos from a normalized psi_dev for the continued-fraction Green function;
 * a (capacity max_iter) and b (capacity max_iter + 1) receive the coefficients.
 * Replaces LanczosGreen (src_ed/Hamiltonian/Lanczos.hpp:156-183). */
int qed_lanczos_green(qed_sector* s, const double* psi_dev, double accur_cf, int max_iter,
                      double* a, double* b, int* n);

/* Band Lanczos on p start vectors phi_dev (p consecutive sector vectors). M0 <= 0 uses
 * 14 p ln(dim). On exit M poles: evals[M] and Q (p x M, column-major) = P U.
 * Capacity of evals/Q is cap columns (cap >= max_iter_BL + p is always enough).
 * Replaces Hamiltonian::build_Q_matrix + bandLanczos
 * (src_ed/Hamiltonian/Hamiltonian_base.hpp:229-254, src_ed/Hamiltonian/Lanczos.hpp:214-395). */
int qed_band_lanczos(qed_sector* s, const double* phi_dev, int p, int M0, int max_iter_BL,
                     double accur_deflation, double accur_band_lanczos, double* evals, double* Q,
                     int cap, int* M);

/* y_j += H_j x_j for n sectors (device vectors) in one batched launch; no counterpart in
 * the reference (its sectors are applied one at a time). mode 0: one workgroup per row;
 * mode 1: the two-pass tiled form for short rows (fails when the rows do not fit its LDS
 * tiles). The launch is repeated reps times (y accumulates each time); *ms, if not null,
 * receives the average launch time in milliseconds. */
int qed_hv_batch(int n, qed_sector* const* s, const double* const* x, double* const* y, int mode, int reps,
                 double* ms);

/* Block Lanczos on p orthonormal start vectors q0_dev (p consecutive sector vectors).
 * qr != 0: upper-triangular off-diagonal blocks from modified Gram-Schmidt; qr == 0 (the
 * reference's default): symmetric blocks from the polar decomposition of the residual
 * block. M0 <= 0 uses 14 p ln(dim), capped by max_iter_BL. On exit M floors: A and B hold
 * M blocks of p x p (column-major, block j at j p^2); capacity cap floors.
 * Replaces blockLanczos / blockLanczosSVD (src_ed/Hamiltonian/Lanczos.hpp:426-779). */
int qed_block_lanczos(qed_sector* s, const double* q0_dev, int p, int M0, int max_iter_BL,
                      double accur_deflation, double accur_band_lanczos, int qr, double* A, double* B,
                      int cap, int* M);

/* Matrix continued fraction G(z) = W^T F_0(z) W with F_M = 0 and
 * F_j = (z - A_j - B_j^T F_{j+1} B_j)^-1 (B_{M-1} unused), added to G (transposed when
 * transpose != 0) for nz frequencies z (complex interleaved). A, B: M blocks p x p; W: p x q
 * (column-major); G: nz blocks q x q complex column-major. p, q <= 24.
 * Replaces matrix_continued_fraction::evaluate (src_ed/matrix_continued_fraction.hpp:139-182). */
int qed_mcf_eval(int p, int q, int M, const double* A, const double* B, const double* W, int nz,
                 const double* z, int transpose, double* G);

/* Davidson-Liu for the k lowest states (device vectors psi_dev: k consecutive sector
 * vectors, out). Replaces Davidson (src_ed/Hamiltonian/Davidson.hpp:38-230). */
int qed_davidson(qed_sector* s, int k, double accur, int max_iter, double* evals, double* psi_dev,
                 int* niter);

/* y_dev += z * c_a x_dev (pm = -1) or z * c†_a x_dev (pm = +1); x lives in (n_up, n_dn)
 * of L orbitals per spin, y in the shifted sector. Replaces the factorized branch of
 * model::create_or_destroy (src_ed/model.cpp:155-180) and Destroy (src_ed/ED_basis.cpp:79-110). */
int qed_apply_c(int L, int n_up, int n_dn, int a, int pm, double z, const double* x_dev,
                double* y_dev);

/* ---- Green-function representations ---------------------------------------------- */

/* G(z_i) += sum_m Q_am conj(Q_bm) / (z_i - e_m) for nz frequencies; Q real (L x M col-major),
 * G (nz x L x L complex col-major, element (b,a) at b + a L as the reference writes it).
 * Replaces Q_matrix::Green_function + VDVH kernel (src_ed/Q_matrix.hpp:133-151,
 * src_util/VDVH_kernel.cpp:28-48). */
int qed_qmatrix_gf(int L, int M, const double* e, const double* Q, int nz, const double* z,
                   double* G);

/* Continued fractions: n fractions of depth[f] coefficients (a, b concatenated with
 * offsets off[f]); out[f + n*i] = b0/(z_i - a0 - b1/(z_i - a1 - ...)).
 * Replaces continued_fraction::evaluate (src_ed/continued_fraction.cpp:52-57). */
int qed_cf_eval(int n, const int* off, const int* depth, const double* a, const double* b, int nz,
                const double* z, double* out);

/* ---- lattice k-sums (CPT / VCA / CDMFT) ----------------------------------------- */

/* V(k) = V0 + sum_t Vt[t] exp(2 pi i k . nv[t]) (n x n complex each; nv[t] 3 doubles).
 * CPT: for every (w, k): Gcpt = Gc(w) (1 - V(k) Gc(w))^-1.
 *   mode 0: out = full Gcpt (nw x nk x n x n complex)
 *   mode 1: out = trace of Gcpt (nw x nk complex)
 *   mode 2: out = sum over k of Gcpt / nk (nw x n x n complex) (projected G, CDMFT host)
 * Replaces lattice_model_instance::set_V / set_Gcpt / projected_Green_function
 * (src_qcm/CPT.cpp:22-86, 283-345) and QCM::CPT_Green_function (src_qcm/QCM.cpp:328-344). */
int qed_cpt_batch(int n, int nw, const double* Gc, const double* V0, int nterm, const double* Vt,
                  const double* nv, int nk, const double* k, int mode, double* out);

/* Device time in ms of the k-sum kernels of the last qed_cpt_batch call (HIP events on the
 * library stream); no counterpart in the reference, for measurement. */
int qed_ksum_last_ms(double* ms);

/* VCA: out[w] = (1/nk) sum_k log|det(V(k) Gc(w) - 1)| per frequency, same V(k) model.
 * Replaces SEF_integrand_k summed by k_integral_grid
 * (src_qcm/lattice_model_instance.cpp:624-643, src_util/integrate.cpp:177-239). */
int qed_sef_batch(int n, int nw, const double* Gc, const double* V0, int nterm, const double* Vt,
                  const double* nv, int nk, const double* k, double* out);

/* CDMFT distance for B independent bath parameter sets (batched bath fit):
 *   d[b] = scale * sum_w weight[w] || Gamma_b(i w) + Ghost[w] ||_F^2,
 *   Gamma_b(z)_{ij} = sum_m tcb_b[i,m] tcb_b[j,m] / (z - eb_b[m]).
 * n = cluster sites, nb = bath orbitals. eb (B x nb), tcb (B x n x nb col-major per set).
 * Replaces lattice_model_instance::CDMFT_distance (src_qcm/CPT.cpp:519-555). */
int qed_cdmft_distance_batch(int n, int nb, int nw, const double* w, const double* weight,
                             const double* Ghost, int B, const double* eb, const double* tcb,
                             double scale, double* d);

/* Same distance with one host function per parameter set: Ghost holds nhost hosts
 * (nhost x nw x n x n complex, col-major per matrix), set b uses host host_of[b].
 * Identical arithmetic to qed_cdmft_distance_batch. */
int qed_cdmft_distance_multi(int n, int nb, int nw, const double* w, const double* weight,
                             int nhost, const double* Ghost, int B, const int* host_of,
                             const double* eb, const double* tcb, double scale, double* d);

/* Batched bath fit: B adaptive Nelder-Mead minimizations of the distance in one launch
 * (one wavefront per instance). Bath parameters x (nvar) map linearly onto a diagonal
 * bath: eb[m] = sum_p x[p] E[p nb + m], tcb[i + n m] = sum_p x[p] T[p n nb + i + n m].
 * Initial simplex x0 + initial_step e_i; stops when the simplex spread is <= xatol and
 * the value spread <= fatol (status 0) or after maxfev evaluations (status 1).
 * Replaces the Nelder-Mead minimization of CDMFT.CDMFT_step (pyqcm/cdmft.py:543-582 with
 * the options of pyqcm/cdmft.py:1614-1633), evaluating CDMFT_distance on the device. */
int qed_cdmft_fit_batch(int n, int nb, int nw, const double* w, const double* weight, int nhost,
                        const double* Ghost, int B, const int* host_of, int nvar, const double* E,
                        const double* T, const double* x0, double initial_step, double xatol,
                        double fatol, int maxfev, double scale, double* x, double* f, int* nfev,
                        int* status);

/* ---- the ED solver interface (namespace ED, src_ed/qcm_ED.hpp:23-340) -------------- */

/* ED::new_model (src_ed/qcm_ED.cpp:39-58) without generators: the trivial group. */
int qed_ed_new_model(const char* name, int n_sites, int n_bath);
/* ED::new_model (src_ed/qcm_ED.cpp:39-58) with n_gen generators of an Abelian point group
 * (gens: n_gen x (n_sites + n_bath), 0-based site images, then the bath entries: phase
 * labels when bath_irrep, else 0-based orbital images, as the reference's binding passes
 * them, src_python/qcm_ED_wrap.hpp:145-163). Generators of order > 2 (complex characters)
 * are rejected. Instances of such a model solve each "R<r>:N<n>:S<s>" sector in the
 * symmetry-adapted basis (qed_mixed_sector_create) and build the Green function from the
 * symmetric orbitals (src_ed/model_instance.hpp:710-817), reported in the site basis. */
int qed_ed_new_model_generators(const char* name, int n_sites, int n_bath, int n_gen, const int* gens,
                                int bath_irrep);
/* ED::new_operator (src_ed/qcm_ED.cpp:73-92): type "one-body", "interaction" or
 * "anomalous" (elements r < c, halved by the binding as the reference's); one-body elements
 * between the spins and anomalous operators set the instance's spin-flip / anomalous
 * mixing; an operator has one target (cluster, bath or cluster-bath); 0-based indices (the
 * binding converts 1-based labels). */
int qed_ed_new_operator(const char* model, const char* name, const char* type, int nel,
                        const int* r, const int* c, const double* v);
/* ED::new_model_instance (src_ed/qcm_ED.cpp:127-158); sectors "R0:N4:S0/R0:N6:S0". */
int qed_ed_new_model_instance(const char* model, int nparam, const char* const* names,
                              const double* values, const char* sectors, uint64_t label);
int qed_ed_erase_model_instance(uint64_t label);
/* ED::ground_state_solve (src_ed/qcm_ED.cpp:179-188); sector string into buf. */
int qed_ed_ground_state_solve(uint64_t label, double* E0, char* buf, int buflen);
/* ED::Green_function_solve (src_ed/qcm_ED.cpp, model_instance.hpp:517-555). */
int qed_ed_green_function_solve(uint64_t label);
/* ED::Green_function (model_instance.hpp:565-635) for nz frequencies at once:
 * G (nz x d x d complex, column-major), d = dim_GF. */
int qed_ed_green_function(uint64_t label, int nz, const double* z, int spin_down, double* G);
/* ED::self_energy (model_instance.hpp:683-700), batched over z. */
int qed_ed_self_energy(uint64_t label, int nz, const double* z, int spin_down, double* S);
/* ED::hybridization_function (model_instance.hpp:1004-1018), batched over z. */
int qed_ed_hybridization_function(uint64_t label, int nz, const double* z, int spin_down,
                                  double* Gam);
/* ED::hopping_matrix (cluster part tc, complex d x d). */
int qed_ed_hopping_matrix(uint64_t label, int spin_down, double* tc);
/* The one-body matrix of cluster and bath, (d + d_b) x (d + d_b) complex column-major
 * (d = dim_GF, d_b the bath's spinor dimension): [[t_c, t_cb], [t_cb^T, t_b]], the bath
 * diagonalized (diag != 0) or as the operators give it. Replaces ED::hopping_matrix_full
 * (src_ed/qcm_ED.cpp:278, src_ed/model_instance.hpp:1491-1525; binding hopping_matrix(...,
 * full=1), src_python/qcm_ED_wrap.hpp:94-106). */
int qed_ed_hopping_matrix_full(uint64_t label, int spin_down, int diag, double* H);
/* Bath matrices after diagonalization: eb (nb), tcb (d x nb column-major). */
int qed_ed_bath(uint64_t label, int spin_down, double* eb, double* tcb);
/* ED::qmatrix (model_instance.hpp:103): M poles, e (M) and Q (d x M column-major);
 * pass e = NULL to query M. */
int qed_ed_qmatrix(uint64_t label, int spin_down, int* M, double* e, double* Q);
/* ED::Potthoff_functional (src_ed/qcm_ED.cpp:330-334): E0 + SEF_bath. */
int qed_ed_potthoff_functional(uint64_t label, double* omega);
/* ED::model_size; dims of the Green function. */
int qed_ed_model_size(const char* model, int* n_sites, int* n_bath);
int qed_ed_dim_gf(uint64_t label, int* d);
/* ED::mixing (src_ed/qcm_ED.cpp:162-165): the Hilbert-space mixing of the instance (0
 * normal, 4 spin up and down apart). */
int qed_ed_mixing(uint64_t label, int* mixing);
/* ED::cluster_averages (model_instance.hpp:148-175): for each active operator, in the
 * order of names, average and variance. For an uncorrelated instance or one read from a
 * file they come from the Green function average: one-body operators on the cluster only,
 * variance NaN (NaN and NaN for the other operators). */
int qed_ed_cluster_averages(uint64_t label, int nop, const char* const* names, double* ave,
                            double* var);
/* ED::Green_function_average (src_ed/qcm_ED.cpp:240-248, model_instance.hpp:643-654): the
 * d x d matrix M(b, a) = <c^dagger_a c_b> integrated from the Green-function representation
 * of the solved instance (Lehmann: poles below zero; continued fractions: the hole b_0;
 * matrix continued fractions: (W_h^T W_h)^T), column-major complex, interleaved. */
int qed_ed_green_function_average(uint64_t label, int spin_down, double* M);
/* ED::Green_function_density (src_ed/qcm_ED.cpp:251-257, model_instance.hpp:661-675): the
 * trace of that matrix per cluster site (both spins). Solves the Green function first. */
int qed_ed_green_function_density(uint64_t label, double* density);
/* ED::fidelity (src_ed/qcm_ED.cpp:360-368, model_instance.hpp:1261-1274): sum over the
 * pairs of low-energy states in the same sector of |<psi1|psi2>|^2 w1 w2. */
int qed_ed_fidelity(uint64_t label1, uint64_t label2, double* fidelity);
/* ED::tr_sigma_inf (src_ed/qcm_ED.hpp:247-252, model_instance.hpp:1277-1312): trace of the
 * self-energy at z = i/eps, eps = 1e-5, 2e-5, 3e-5, extrapolated to eps = 0 by polynomial
 * interpolation. */
int qed_ed_tr_sigma_inf(uint64_t label, double* value);
/* ED::hybridization (src_ed/qcm_ED.hpp:287-294, model_instance.hpp:1455-1466): the diagonal
 * of the bath one-body matrix (n_bath values) and the cluster-bath matrix, n_sites x n_bath
 * column-major. Fails without bath orbitals. */
int qed_ed_hybridization_lehmann(uint64_t label, int spin_down, double* E, double* T);
/* ED::print_wavefunction (src_ed/qcm_ED.cpp:431-440): text of the low-energy states (a
 * header per state; the amplitudes and basis states when the sector dimension is at most
 * max_dim_print). *needed = the length of the text + 1; up to buflen - 1 characters are
 * copied into buf. */
int qed_ed_print_wavefunction(uint64_t label, char* buf, int buflen, int* needed);
/* ED::get_combined_mcf (src_ed/qcm_ED.hpp:312-340, qcm_ED.cpp:455-540): the combined matrix
 * continued fraction of the first low-energy state (GF_method 'M' or 'L' with combine_mcf
 * set): block size *p, *M floors, W of *p x *q; A and B receive M p x p blocks and W the
 * weight, column-major (pass NULL to query the sizes). */
int qed_ed_get_combined_mcf(uint64_t label, int spin_down, int* p, int* M, int* q, double* A, double* B,
                            double* W);
/* ED::density_matrix (src_ed/qcm_ED.cpp:205-216): for the factorized Hamiltonian format
 * (Hamiltonian_format 'F') the reference returns an empty matrix and basis
 * (density_matrix_factorized, src_ed/model_instance.hpp:1629-1633): *dim = 0; for the other
 * formats the reduced density matrix of the first state on the sites (0-based)
 * (density_matrix_mixed, src_ed/model_instance.hpp:1547-1625): *dim subsystem states, rho
 * (dim x dim complex, column-major) and the subsystem states collapsed onto bits 0.. of
 * each spin half (states, dim entries). Pass rho = NULL to query *dim. Trivial group. */
int qed_ed_density_matrix(uint64_t label, int nsites, const int* sites, int* dim, double* rho, uint64_t* states);
/* ED::susceptibility (src_ed/qcm_ED.cpp:307-315, model_instance.hpp:1024-1058): dynamic
 * susceptibility of operator `op` at nz complex frequencies z (interleaved re, im), the
 * Boltzmann-weighted sum over low-energy states of <psi|O (w - H + E)^-1 O|psi>, with the
 * Lehmann form of O|psi> from band Lanczos (full diagonalization below max_dim_full).
 * chi: nz complex values, interleaved. */
int qed_ed_susceptibility(uint64_t label, const char* op, int nz, const double* z, double* chi);
/* ED::susceptibility_poles (src_ed/qcm_ED.cpp:319-326, model_instance.hpp:1066-1113): the
 * poles (> 1e-8) and residues (> 1e-6) of the same sum, poles closer than 1e-4 merged, as
 * (pole, residue) pairs in poles[2k], poles[2k+1]; at most cap pairs are written and *n is
 * their total number. */
int qed_ed_susceptibility_poles(uint64_t label, const char* op, int cap, int* n, double* poles);
/* ED::write_instance_hdf5 (src_ed/qcm_ED.hpp:256-261, qcm_ED.cpp:373-389): solves the Green
 * function if needed and writes the instance to group `group` of the HDF5 file `filename`
 * (created if absent; an existing group of that name is replaced), in the reference's
 * layout (src_ed/model_instance.hpp:1339-1377, src_ed/state.hpp:46-69). libhdf5 is loaded
 * on first use (QED_HDF5_LIB overrides its path). */
int qed_ed_write_instance_hdf5(const char* filename, uint64_t label, const char* group);
/* ED::read_instance_hdf5 (src_ed/qcm_ED.hpp:263-266, qcm_ED.cpp:392-403): replaces the
 * states of instance `label` by those of the group (energies, weights, Green-function
 * representation), warns on stdout when a stored parameter differs from the instance's
 * by more than 1e-8, and marks the instance solved. The read instance holds no
 * ground-state vectors: its cluster averages come from the Green function average. */
int qed_ed_read_instance_hdf5(const char* filename, uint64_t label, const char* group);
/* Global parameters (src_util/global_parameter.cpp:12-91): int/double/bool/char maps. */
int qed_ed_set_global_int(const char* name, int v);
int qed_ed_set_global_double(const char* name, double v);
int qed_ed_set_global_bool(const char* name, int v);
int qed_ed_set_global_char(const char* name, char v);
int qed_ed_get_global_double(const char* name, double* v);

/* ---- batched CDMFT impurity solves ---------------------------------------------- */

/* Solves nb independent (cluster + bath) instances of one model topology in one call:
 * ground state by Lanczos and Green function by band Lanczos for each, all vectors in
 * one arena, the H.v of all instances issued as one persistent launch per Krylov step
 * (captured in a HIP graph when graph != 0). params: nb x nparam values in the order of
 * names; sectors: one sector string per instance. Outputs E0[nb] and the spin-up
 * Lehmann form of each instance, packed: Mout[nb] poles, e (nb x cap),
 * Q (nb x cap x d, pole-major: Q[(b cap + m) d + i]). New in this library: the batched
 * form of ED::ground_state_solve + ED::Green_function_solve (src_ed/qcm_ED.cpp). */
/* Solves n instances already created with qed_ed_new_model_instance together (ground
 * states and Green functions, same batched engine as qed_batch_solve); afterwards the
 * qed_ed_* queries on those labels use the stored results. chunk: Lanczos iterations
 * between stopping tests (0: 16). stats (4 ints, may be null): lockstep Lanczos
 * iterations, band Lanczos steps, and the autotuned H.v block size and grid. */
int qed_ed_batch_solve_instances(int n, const uint64_t* labels, int graph, int chunk, int* stats);

int qed_batch_solve(const char* model, int nparam, const char* const* names, int nb,
                    const double* params, const char* const* sectors, int graph, int cap,
                    double* E0, int* Mout, double* e, double* Q);

/* ---- learned warm start (MLP 3 -> 128 -> 128 -> 16) ------------------------------ */

/* Opaque MLP with parameters resident on the device. */
typedef struct qed_mlp qed_mlp;
int qed_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed, qed_mlp** out);
int qed_mlp_destroy(qed_mlp* m);
int qed_mlp_get_params(qed_mlp* m, float* params, int* count);
int qed_mlp_set_params(qed_mlp* m, const float* params, int count);
/* Forward pass on n samples (x: n x n_in row-major host), FP32; y: n x n_out. */
int qed_mlp_forward(qed_mlp* m, int n, const float* x, float* y);
/* Forward pass with bf16 weights and MFMA (inference). */
int qed_mlp_forward_bf16(qed_mlp* m, int n, const float* x, float* y);
/* Mean-squared-error loss and gradient (grad: same layout as params) on one batch. */
int qed_mlp_loss_grad(qed_mlp* m, int n, const float* x, const float* t, float* loss,
                      float* grad);
/* Adam training: epochs over (x, t) with minibatch size bs; loss history per epoch. */
int qed_mlp_train(qed_mlp* m, int n, const float* x, const float* t, int epochs, int bs, float lr,
                  uint64_t seed, float* loss_hist);

#ifdef __cplusplus
}
#endif

#endif /* QED_H */
