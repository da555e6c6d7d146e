"""Host-side logic of the lattice/CDMFT layer (no GPU): the Nelder-Mead coroutine against
scipy, Broyden, frequency grids, the tiling's Fourier decomposition, bath gauge fixing."""
import numpy as np
import pytest
import scipy.optimize

from pyqcm_amd import cdmft as C
from pyqcm_amd import lattice as LT


def _rosen_shifted(x):
    return scipy.optimize.rosen(x) + 0.1 * np.sum((x - 0.3) ** 2)


@pytest.mark.parametrize("n", [2, 5, 16])
def test_nelder_mead_coroutine_reproduces_scipy(n):
    """Same options as the reference's bath fit: adaptive, initial simplex x0 + 0.1 e_i."""
    rng = np.random.default_rng(n)
    x0 = rng.uniform(-1, 1, n)
    step, xatol, fatol = 0.1, 1e-6, 1e-8
    sim = np.repeat(x0[None], n + 1, 0)
    for i in range(n):
        sim[i + 1, i] += step
    calls = []

    def f(x):
        calls.append(np.copy(x))
        return _rosen_shifted(x)

    sol = scipy.optimize.minimize(f, x0, method="Nelder-Mead",
                                  options={"maxfev": 200000, "xatol": xatol, "fatol": fatol,
                                           "initial_simplex": sim, "adaptive": True, "disp": False})
    co = C.nelder_mead_coroutine(x0, step, xatol, fatol, 200000)
    mine = []

    def ev(reqs):
        out = []
        for _, pts in reqs:
            mine.extend(np.copy(p) for p in pts)
            out.append(np.array([_rosen_shifted(p) for p in pts]))
        return out

    (x, nfev, ok, fun), = C.run_coroutines([co], ev)
    assert ok and sol.success
    assert nfev == sol.nfev == len(calls) == len(mine)
    np.testing.assert_array_equal(np.array(mine), np.array(calls))
    np.testing.assert_array_equal(x, sol.x)
    assert fun == sol.fun


def test_run_coroutines_interleaves_independent_problems():
    rng = np.random.default_rng(3)
    x0s = [rng.uniform(-1, 1, 4) for _ in range(5)]
    singles = []
    for x0 in x0s:
        co = C.nelder_mead_coroutine(x0, 0.1, 1e-7, 1e-10, 100000)
        singles.append(C.run_coroutines([co], lambda r: [np.array([_rosen_shifted(p) for p in pts])
                                                         for _, pts in r])[0])
    cos = [C.nelder_mead_coroutine(x0, 0.1, 1e-7, 1e-10, 100000) for x0 in x0s]
    batch = C.run_coroutines(cos, lambda r: [np.array([_rosen_shifted(p) for p in pts])
                                             for _, pts in r])
    for a, b in zip(singles, batch):
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1]


def test_broyden_solves_nonlinear_system():
    A = np.array([[3.0, 0.5, 0.0], [0.2, 2.5, 0.3], [0.0, 0.4, 2.0]])
    b = np.array([1.0, -0.5, 0.25])

    def step(x):  # fixed-point map x -> x - 0.3 (A x + 0.1 x^3 - b)
        return x - 0.3 * (A @ x + 0.1 * x ** 3 - b)

    F = lambda x: x - step(x)
    x, it, _ = C.broyden(F, np.zeros(3), 0.0, xtol=1e-12, maxiter=50)
    np.testing.assert_allclose(A @ x + 0.1 * x ** 3, b, atol=1e-10)


def test_convergence_manager_depth():
    cm = C.parameters_convergence(1e-3, depth=2)
    xs = [np.array([1.0, 2.0]), np.array([1.1, 2.0]), np.array([1.1001, 2.0]),
          np.array([1.1002, 2.0]), np.array([1.1002, 2.0])]
    res = [cm.test(x) for x in xs]
    assert res == [False, False, False, True, True]


def test_legendre_grid_integrates_rational_functions():
    """sum_i p_i f(w_i) = (1/pi) int_0^inf f: 1/(1+w^2) -> 1/2, 1/(a^2+w^2) -> 1/(2a)."""
    w, p = LT.legendre_frequency_grid(1, 10, 10, 10, 10)
    assert abs(np.sum(p / (1 + w ** 2)) - 0.5) < 1e-6
    w, p = LT.legendre_frequency_grid(1, 10, 30, 30, 30)
    assert abs(np.sum(p / (1 + w ** 2)) - 0.5) < 1e-12
    assert abs(np.sum(p / (0.3 ** 2 + w ** 2)) - 1 / 0.6) < 1e-9
    w, p = LT.legendre_frequency_grid(1, 10, 5, 10, 5)
    assert len(w) == 20 and np.all(np.diff(w) > 0)
    g = C.frequency_grid()
    assert g.name == "legendre(1-10-5-10-5)" and np.all(g.cdmft_weight == 1)


def test_square_tiling_reproduces_dispersion():
    """Eigenvalues of t(K) for the 2x2 tiling = -2t(cos kx + cos ky) - 4t' cos kx cos ky
    at the 4 lattice momenta folded onto K."""
    T = LT.square_tiling(2, 2)
    vals = {"t": 1.0, "tp": -0.3, "mu": 0.0}
    rng = np.random.default_rng(0)
    for K in rng.uniform(0, 1, (5, 2)):
        ev = np.linalg.eigvalsh(T.tk(vals, [[K[0], K[1], 0]])[0])
        ks = []
        for a in range(2):
            for b in range(2):
                kx = np.pi * (K[0] + a)  # k = 2 pi (K + G)/2
                ky = np.pi * (K[1] + b)
                ks.append(-2 * (np.cos(kx) + np.cos(ky)) + 4 * 0.3 * np.cos(kx) * np.cos(ky))
        np.testing.assert_allclose(ev, np.sort(ks), atol=1e-12)


def test_chain_tiling_cluster_operator_elements():
    T = LT.chain_tiling(4)
    T.n_orb = 4
    el = T.cluster_operator_elements("t")
    assert (1, 2, -1.0) in el and (5, 6, -1.0) in el and len(el) == 6
    cdw = T.cluster_operator_elements("cdw")
    assert [v for (_, _, v) in cdw[:4]] == [1.0, -1.0, 1.0, -1.0]


def test_canonical_bath_is_gauge_invariant():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-3, 3, 8), rng.uniform(0.1, 1, 8)])
    y = x.copy()
    y[8:] *= np.array([1, -1, 1, 1, -1, 1, -1, 1])
    y[[0, 1]] = y[[1, 0]]
    y[[8, 9]] = y[[9, 8]]
    np.testing.assert_array_equal(C.canonical_bath(x), C.canonical_bath(y))


def test_canonical_map_is_the_signed_permutation_of_canonical_bath():
    rng = np.random.default_rng(4)
    for _ in range(20):
        x = np.concatenate([rng.uniform(-3, 3, 8), rng.uniform(-1, 1, 8)])
        Pm = C.canonical_map(x)
        np.testing.assert_array_equal(Pm @ x, C.canonical_bath(x))
        np.testing.assert_array_equal(Pm @ Pm.T, np.eye(16))


def test_broyden_matrix_start_is_the_inverse_jacobian():
    """For a linear residual F(x) = A x - b, starting from the exact inverse Jacobian A^-1
    the first step lands on the root and the next step is zero; a scalar start alpha gives
    (1 + alpha) times the unit matrix, as in the reference's broyden()."""
    A = np.array([[3.0, 0.5, 0.0], [0.2, 2.5, 0.3], [0.0, 0.4, 2.0]])
    b = np.array([1.0, -0.5, 0.25])
    F = lambda x: A @ x - b
    x, it, J = C.broyden(F, np.zeros(3), np.linalg.inv(A), xtol=1e-12, maxiter=5)
    assert it == 1
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-14)
    x2, it2, _ = C.broyden(F, np.zeros(3), 0.5, xtol=1e-12, maxiter=200)
    x3, it3, _ = C.broyden(F, np.zeros(3), 1.5 * np.eye(3), xtol=1e-12, maxiter=200)
    np.testing.assert_array_equal(x2, x3)
    assert it2 == it3


class _MapProblem:
    """A stand-in for a CDMFT problem whose 'impurity solve + bath fit' is the map g: the
    host is the current point, the fit returns g(host); the fit fails where x[0] > 5."""

    def __init__(self):
        self.calls = 0

    def solve_batch(self, values_list, x_list):
        self.calls += len(x_list)
        return [np.array(x, float) for x in x_list]

    def fit(self, hosts, idx, X, *args):
        hosts = np.asarray(hosts, float)
        # contraction to the fixed point (1, 1); beyond x[0] = 10 a map whose pull
        # vanishes, where Broyden's secant update breaks down
        far = hosts[:, :1] > 10.0
        Xn = np.where(far, hosts - 0.05 * np.tanh(hosts - 1.0), 1.0 + 0.5 * (hosts - 1.0))
        ok = (hosts[:, 0] <= 5.0) | (hosts[:, 0] > 10.0)
        return Xn, np.zeros(len(hosts)), np.ones(len(hosts), int), ok


def test_batched_cdmft_falls_back_to_the_second_start():
    """A problem whose first start fails (a failed fit; an undefined Broyden step or no
    convergence within the allowed iterations) starts again from its fallback point and
    converges there; without a fallback the same problems fail."""
    from pyqcm_amd.cdmft import BatchedCDMFT
    vals = [{}, {}, {}]
    x0 = [[7.0, 1.0], [40.0, -30.0], [0.5, 0.8]]  # fit fails; far away; close
    cold = [[1.2, 0.9]] * 3
    P = _MapProblem()
    plain = BatchedCDMFT(P, vals, x0, solver=P.solve_batch, fit="host")
    assert plain.failed[0] and plain.failed[1] and not plain.failed[2] and not plain.restarted.any()
    P = _MapProblem()
    B = BatchedCDMFT(P, vals, x0, solver=P.solve_batch, fit="host", fallback_list=cold, fallback_after=8)
    assert B.restarted[0] and B.restarted[1] and not B.restarted[2]
    assert (B.converged & ~B.failed).all()
    for i in range(3):
        np.testing.assert_allclose(B.x[i], [1.0, 1.0], atol=1e-4)
    assert B.niter[1] > 2  # both attempts counted


def _ring_bdg(N, t, mu, D, S, hx):
    """Real-space one-body matrix of an N-site ring in the spinor (c_up, c+_down) (Nambu,
    D and S singlet pairing on the bonds and sites, v (c_i up c_j down + h.c.) giving -v in
    both off-diagonal blocks) or (c_up, c_down) (spin flip, hx along x)."""
    if hx is None:
        H = np.zeros((2 * N, 2 * N))
        for i in range(N):
            j = (i + 1) % N
            H[i, j] = H[j, i] = -t
            H[N + i, N + j] = H[N + j, N + i] = t
            H[i, i], H[N + i, N + i] = -mu, mu
            H[i, N + i] = H[N + i, i] = -S
            for a, b in ((i, j), (j, i)):
                H[a, N + b] += -D
                H[N + b, a] += -D
        return H
    H = np.zeros((2 * N, 2 * N))
    for i in range(N):
        j = (i + 1) % N
        H[i, j] = H[j, i] = H[N + i, N + j] = H[N + j, N + i] = -t
        H[i, i] = H[N + i, N + i] = -mu
        H[i, N + i] = H[N + i, i] = hx
    return H


@pytest.mark.parametrize("form", ["nambu", "spin_flip"])
def test_tiling_spinor_terms_synthesize_the_ring(form):
    """The tiling's spinor Fourier terms of a 2-site cluster, summed over the two
    superlattice wavevectors of a 4-site ring with phases exp(2 pi i k (R - R')), give the
    ring's real-space spinor matrix built bond by bond: the inter-cluster pairing, hole
    and spin-flip blocks and their phases."""
    T = LT.chain_tiling(2)
    if form == "nambu":
        T.pairing("D", (1, 0, 0)).pairing("S", (0, 0, 0))
        vals, mixing = {"t": 1.0, "mu": 0.6, "D": 0.3, "S": 0.2}, 1
        ref = _ring_bdg(4, 1.0, 0.6, 0.3, 0.2, None)
    else:
        T.spin_flip("hx")
        vals, mixing = {"t": 1.0, "mu": 0.6, "hx": 0.4}, 2
        ref = _ring_bdg(4, 1.0, 0.6, 0.0, 0.0, 0.4)
    k = np.array([[0.0, 0, 0], [0.5, 0, 0]])
    hk = T.tk(vals, k, mixing=mixing)
    got = np.zeros((8, 8), complex)
    for s1 in range(4):
        for s2 in range(4):
            i, R, j, Rp = s1 % 2, s1 // 2, s2 % 2, s2 // 2
            ph = np.exp(2j * np.pi * k[:, 0] * (R - Rp))
            for h1 in (0, 1):
                for h2 in (0, 1):
                    got[s1 + 4 * h1, s2 + 4 * h2] = 0.5 * np.sum(ph * hk[:, i + 2 * h1, j + 2 * h2])
    np.testing.assert_allclose(got, ref, atol=1e-14)
