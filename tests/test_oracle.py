"""Pins the CPU oracle (oracle/) before it is trusted as the checker of the HIP path.

Every oracle function is checked against something independent of it: exhaustive
enumeration / closed forms (basis, rank), a Jordan-Wigner brute force (H.v, create and
destroy, Green functions), closed-form spectra (U = 0, atomic limit, two-site Hubbard),
the reference's own recorded outputs (tests/golden/reference_values.json) and an
independent re-implementation of the libstdc++ random stream the reference uses.
"""
import math
from itertools import combinations

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM


# ---------------------------------------------------------------- basis and rank

@pytest.mark.parametrize("L", [1, 2, 5, 8, 12])
def test_halfbasis_exhaustive(L):
    for N in range(L + 1):
        b = O.halfbasis(L, N)
        ref = sorted(sum(1 << p for p in c) for c in combinations(range(L), N))
        assert b.tolist() == ref
        # combinatorial number system closed form of the rank
        for i, s in enumerate(b.tolist()):
            bits = [p for p in range(L) if s >> p & 1]
            assert sum(math.comb(p, k + 1) for k, p in enumerate(bits)) == i


def test_halfbasis_sizes():
    assert len(O.halfbasis(16, 8)) == 12870
    assert len(O.halfbasis(12, 6)) == 924
    assert len(O.halfbasis(18, 9)) == 48620


# ---------------------------------------------------------------- H.v vs brute force

SMALL_MODELS = {
    "chain4": lambda: TM.chain(4),
    "chain5_tp": lambda: TM.chain(5, tp=-0.3, U=3.0, mu=1.1),
    "square2x3": lambda: TM.square(2, 3, tp=-0.3, U=8.0, mu=4.0),
    "plaq_4bath": lambda: TM.plaquette_bath(
        {"U": 6.0, "mu": 3.0, **{"eb%d" % x: 0.3 * x - 1.2 for x in range(1, 5)},
         **{"tb%d" % x: 0.2 + 0.1 * x for x in range(1, 5)}}, n_bath=4),
}


@pytest.mark.parametrize("name", sorted(SMALL_MODELS))
def test_hv_matches_jordan_wigner(name):
    m, vals = SMALL_MODELS[name]()
    L = m.L
    rng = np.random.default_rng(7)
    for (nup, ndn) in [(L // 2, L // 2), (L // 2 + 1, L // 2), (1, L - 1)]:
        S = O.Sector(m, vals, nup, ndn)
        H, _ = O.dense_hamiltonian(m, vals, nup, ndn)
        assert np.allclose(H, H.T)
        x = rng.standard_normal(S.dim)
        y0 = rng.standard_normal(S.dim)
        y = S.hv(x, y0)
        np.testing.assert_allclose(y, y0 + H @ x, rtol=0, atol=1e-12 * (1 + np.abs(H).max()) * S.dim)
        # parallel variant computes the same product
        np.testing.assert_allclose(S.hv_omp(x, y0, 2), y, rtol=0, atol=1e-12 * S.dim)


def test_hv_hermitian_12_sites():
    m, vals = TM.square(3, 4, tp=-0.3)
    S = O.Sector(m, vals, 6, 6)
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(S.dim), rng.standard_normal(S.dim)
    lhs, rhs = a @ S.hv(b), S.hv(a) @ b
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_hv_atomic_and_free_limits():
    # t = 0: diagonal closed form -mu N + U * double occupancy
    m, vals = TM.chain(6, t=0.0, U=5.0, mu=1.5)
    S = O.Sector(m, vals, 3, 3)
    d = S.hv(np.ones(S.dim))
    bu, bd = O.halfbasis(6, 3), O.halfbasis(6, 3)
    expect = np.array([-1.5 * 6 + 5.0 * bin(int(u) & int(dn)).count("1") for dn in bd for u in bu])
    np.testing.assert_allclose(d, expect, atol=1e-13)


# ---------------------------------------------------------------- ground state

def test_e0_reference_chain4(golden):
    """examples/1D_4/GS.tsv row 1: U=4, mu=2, t=1 (V, V0m at 1e-9: negligible)."""
    row = golden["chain4"]["rows"][0]
    m, vals = TM.chain(4, U=float(row["U"]), mu=float(row["mu"]), t=float(row["t"]))
    E0, gs, n, a, b = O.Sector(m, vals, 2, 2).lanczos_gs()
    assert abs(E0 - float(row["E0_1"])) < 1e-7


def test_e0_closed_forms():
    # U = 0: sum of the lowest one-particle levels (both spins) minus mu N
    m, vals = TM.square(2, 3, tp=-0.3, U=0.0, mu=0.7)
    nn, nnn = O.square_bonds(2, 3)
    T = np.zeros((6, 6))
    for (i, j) in nn:
        T[i, j] = T[j, i] = -1.0
    for (i, j) in nnn:
        T[i, j] = T[j, i] = 0.3
    eps = np.linalg.eigvalsh(T)
    for (nup, ndn) in [(3, 3), (2, 4), (4, 3)]:
        E0 = O.Sector(m, vals, nup, ndn).lanczos_gs()[0]
        expect = eps[:nup].sum() + eps[:ndn].sum() - 0.7 * (nup + ndn)
        assert abs(E0 - expect) < 1e-10
    # two-site Hubbard, N=2 Sz=0: U/2 - sqrt(U^2/4 + 4t^2) - 2 mu
    for U in (0.5, 4.0, 9.0):
        m, vals = TM.chain(2, U=U, mu=0.3)
        E0 = O.Sector(m, vals, 1, 1).lanczos_gs()[0]
        assert abs(E0 - (U / 2 - math.sqrt(U * U / 4 + 4) - 0.6)) < 1e-11


def test_e0_dense_and_residual():
    m, vals = TM.square(3, 3, tp=-0.3, U=8.0, mu=4.0)
    S = O.Sector(m, vals, 5, 4)
    E0, gs, n, a, b = S.lanczos_gs()
    H, _ = O.dense_hamiltonian(m, vals, 5, 4) if S.dim < 5000 else (None, None)
    assert np.linalg.norm(S.hv(gs) - E0 * gs) < 1e-7
    assert abs(np.linalg.norm(gs) - 1) < 1e-10
    if H is not None:
        assert abs(E0 - np.linalg.eigvalsh(H)[0]) < 1e-10 * abs(E0)


@pytest.mark.slow
def test_e0_reference_plaquette_8bath(golden):
    """tests/test_diagonalization/S_format.ref: GS_energy -75.0353, sector R0:N12:S0
    (computed there with the C2v group; the trivial-group ground state is the same)."""
    g = golden["plaquette_8bath"]
    m, vals = TM.plaquette_bath(g["params"])
    E0 = O.Sector(m, vals, 6, 6).lanczos_gs()[0]
    assert abs(E0 - g["E0"]) < 6e-5


def test_e0_reference_chain8(golden):
    """S_format_complex.ref: 8-site chain with imaginary hopping ti=1, U=1e-9. On an open
    chain the gauge c_j -> i^j c_j maps it onto real hopping t=1 with the same spectrum."""
    m, vals = TM.chain(8, t=1.0, U=1e-9, mu=0.0)
    E0 = O.Sector(m, vals, 4, 4).lanczos_gs()[0]
    assert abs(E0 - golden["chain8_complex"]["E0"]) < 1e-5
    assert abs(E0 + 4 * sum(math.cos(k * math.pi / 9) for k in range(1, 5))) < 1e-8


# ---------------------------------------------------------------- start vector

def _minstd(seed):
    x = seed % 2147483647
    if x == 0:
        x = 1
    while True:
        x = (16807 * x) % 2147483647
        yield x


def _libstdcxx_normals(seed, n, saved=None):
    """Independent re-implementation of std::default_random_engine (minstd_rand0),
    std::generate_canonical<double,53> and libstdc++'s Marsaglia-polar normal_distribution."""
    g = _minstd(seed)
    r = 2147483646.0

    def canon():
        s = float(next(g) - 1) * 1.0
        tmp = r
        s += float(next(g) - 1) * tmp
        tmp = tmp * r
        v = s / tmp
        return v if v < 1.0 else math.nextafter(1.0, 0.0)

    out = []
    while len(out) < n:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            x = 2.0 * canon() - 1.0
            y = 2.0 * canon() - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2 * math.log(r2) / r2)
        saved = x * mult
        out.append(y * mult)
    return np.array(out), saved


def test_start_vector_matches_libstdcxx_stream():
    O.lib().or_reset_normal()
    n1, n2 = 1001, 500
    v1 = np.zeros(n1)
    O.lib().or_start_vector(n1, 0, O._dp(v1))
    v2 = np.zeros(n2)
    O.lib().or_start_vector(n2, 5, O._dp(v2))
    r1, saved = _libstdcxx_normals(0, n1)
    r2, _ = _libstdcxx_normals(5, n2, saved=saved)   # odd n1 leaves one cached deviate
    np.testing.assert_allclose(v1, r1 / np.linalg.norm(r1), rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(v2, r2 / np.linalg.norm(r2), rtol=1e-14, atol=1e-16)


# ---------------------------------------------------------------- tridiagonal

def test_dstev_matches_numpy():
    rng = np.random.default_rng(0)
    d, e = rng.standard_normal(40), rng.standard_normal(39)
    w, Z = O.dstev(d, e)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(T), atol=1e-12)
    np.testing.assert_allclose(T @ Z[0], w[0] * Z[0], atol=1e-11)


# ---------------------------------------------------------------- create / destroy

@pytest.mark.parametrize("nup,ndn", [(2, 2), (1, 3), (3, 2)])
def test_create_destroy_matches_jordan_wigner(nup, ndn):
    L = 4
    c = O._fock_ops(L)
    rng = np.random.default_rng(11)
    src = O.sector_states(L, nup, ndn)
    x = rng.standard_normal(len(src))
    full = np.zeros(1 << (2 * L))
    full[src] = x
    for a in range(2 * L):
        for pm in (-1, 1):
            tup, tdn = (nup + pm, ndn) if a < L else (nup, ndn + pm)
            if not (0 <= tup <= L and 0 <= tdn <= L):
                continue
            y = O.create_destroy(L, nup, ndn, a, pm, x, z=0.5)
            op = c[a] if pm == -1 else c[a].T
            ref = 0.5 * (op @ full)[O.sector_states(L, tup, tdn)]
            np.testing.assert_allclose(y, ref, atol=1e-14)


# ---------------------------------------------------------------- Green functions

def _phi(m, vals, nup, ndn, gs, pm, orbitals):
    return [O.create_destroy(m.L, nup, ndn, a, pm, gs) for a in orbitals]


@pytest.mark.parametrize("dense", [True, False])
def test_qmatrix_matches_lehmann(dense):
    """Dense path (sectors below max_dim_full) on a 5-site chain; band Lanczos on an
    8-site chain whose sectors exceed the band-Lanczos iteration budget."""
    n = 5 if dense else 8
    m, vals = TM.chain(n, tp=-0.3, U=3.0, mu=1.1)
    nup = ndn = n // 2
    S = O.Sector(m, vals, nup, ndn)
    E0, gs, *_ = S.lanczos_gs()
    orbs = list(range(n))
    zs = 1j * np.array([0.1, 0.5, 1.3, 4.0]) + 0.05
    E0b, Gref = O.dense_gf(m, vals, nup, ndn, orbs, zs)
    assert abs(E0 - E0b) < 1e-11
    es, Qs = [], []
    for pm in (-1, 1):
        T = O.Sector(m, vals, nup + pm, ndn)
        phi = _phi(m, vals, nup, ndn, gs, pm, orbs)
        if dense:
            e, Q = O.dense_qmatrix(T.hv, T.dim, phi)
        else:
            e, Q = O.band_lanczos(T.hv, phi)
        e = e - E0
        if pm == -1:
            e = -e
        e, Q = O.streamline(e, Q)
        es.append(e)
        Qs.append(Q)
    e = np.concatenate(es)
    Q = np.hstack(Qs)
    assert O.check_norm(Q) < 1e-10
    for iz, z in enumerate(zs):
        np.testing.assert_allclose(O.qmatrix_gf(e, Q, z), Gref[iz], atol=1e-9)


def test_continued_fraction_matches_lehmann():
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    S = O.Sector(m, vals, 2, 2)
    E0, gs, *_ = S.lanczos_gs()
    zs = 1j * np.array([0.2, 1.0, 3.0])
    _, Gref = O.dense_gf(m, vals, 2, 2, [0, 1, 2, 3], zs)
    for o in range(4):
        G = np.zeros(len(zs), complex)
        for pm in (-1, 1):
            T = O.Sector(m, vals, 2 + pm, 2)
            psi = O.create_destroy(4, 2, 2, o, pm, gs)
            nrm = psi @ psi
            a, b = T.lanczos_green(psi / math.sqrt(nrm), accur_cf=1e-10, max_iter=400)
            ca, cb = O.cf_coefficients(a, b, E0, nrm, pm == 1)
            G += np.array([O.cf_evaluate(ca, cb, z) for z in zs])
        np.testing.assert_allclose(G, Gref[:, o, o], atol=1e-10)


# ---------------------------------------------------------------- lattice kernels

def test_cpt_identity_and_sef():
    rng = np.random.default_rng(5)
    n = 4
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Gc = np.linalg.inv(A + 3j * np.eye(n))
    V = rng.standard_normal((n, n))
    V = V + V.T
    Gcpt = O.cpt_green(Gc, V)
    # G_cpt^-1 = G_c^-1 - V
    np.testing.assert_allclose(np.linalg.inv(Gcpt), np.linalg.inv(Gc) - V, atol=1e-12)
    s = O.sef_integrand(Gc, V)
    assert abs(s - 2 * np.log(abs(np.linalg.det(np.eye(n) - V @ Gc)))) < 1e-12


# ---------------------------------------------------------------- warm-start MLP

def test_mlp_oracle_gradient_matches_finite_differences():
    """The oracle's backpropagation against central differences of its own loss."""
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(2)
    ni, nh, no = 3, 8, 5
    n_par = nh * ni + nh + nh * nh + nh + no * nh + no
    p = rng.normal(0, 0.5, n_par)
    x = rng.normal(size=(7, ni))
    t = rng.normal(size=(7, no))
    loss, g = MO.loss_grad(p, x, t, ni, nh, no)
    h = 1e-6
    fd = np.zeros(n_par)
    for i in range(n_par):
        e = np.zeros(n_par)
        e[i] = h
        fd[i] = (MO.loss_grad(p + e, x, t, ni, nh, no)[0] - MO.loss_grad(p - e, x, t, ni, nh, no)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    # the loss itself is the mean over samples and outputs
    y, _ = MO.forward(p, x, ni, nh, no)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-14


def test_mlp_oracle_adam_first_step_is_signed_lr():
    """Bias-corrected Adam: the first step moves every parameter by lr * g/(|g| + eps')."""
    from oracle import mlp_oracle as MO
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = MO.adam_step(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)


# ---------------------------------------------------------------- block Lanczos / MCF

def _mcf_gf(m, vals, nup, ndn, gs, E0, zs, qr):
    n = m.L if not hasattr(m, "n_sites") else m.n_sites
    G = np.zeros((len(zs), n, n), complex)
    for pm in (-1, 1):
        T = O.Sector(m, vals, nup + pm, ndn)
        phi = [O.create_destroy(m.L, nup, ndn, a, pm, gs) for a in range(n)]
        q, W = O.mgs_weight(phi)
        A, B = O.block_lanczos(T.hv, q, qr=qr)
        A = O.mcf_shift(A, E0, pm == 1)
        for iz, z in enumerate(zs):
            g = O.mcf_evaluate(A, B, W, z)
            G[iz] += g if pm == 1 else g.T
    return G


@pytest.mark.parametrize("qr", [False, True])
@pytest.mark.parametrize("n", [4, 8])
def test_mcf_matches_lehmann(n, qr):
    """Matrix continued fraction of the block Lanczos (QR and polar blocks) against the
    brute-force Lehmann sum: exhaustive on the 4-site chain, converged on the 8-site chain.
    QR blocks are not symmetric, so a transposed B in the recursion fails here."""
    m, vals = TM.chain(n, tp=-0.3, U=3.0, mu=1.1)
    nup = ndn = n // 2
    E0, gs, *_ = O.Sector(m, vals, nup, ndn).lanczos_gs()
    zs = 1j * np.array([0.1, 0.5, 1.3, 4.0]) + 0.05
    E0b, Gref = O.dense_gf(m, vals, nup, ndn, list(range(n)), zs)
    G = _mcf_gf(m, vals, nup, ndn, gs, E0, zs, qr)
    assert np.abs(G - Gref).max() < 1e-11


def test_mcf_deflated_block_truncates():
    """5-site chain: the hole block deflates at a singular value of ~2e-8 (below
    accur_deflation = 1e-7) and the recurrence stops there, as the reference's does; the
    Green function is then exact up to that truncation (~7e-9)."""
    m, vals = TM.chain(5, tp=-0.3, U=3.0, mu=1.1)
    E0, gs, *_ = O.Sector(m, vals, 2, 2).lanczos_gs()
    zs = 1j * np.array([0.1, 0.5, 1.3, 4.0]) + 0.05
    _, Gref = O.dense_gf(m, vals, 2, 2, list(range(5)), zs)
    for qr in (False, True):
        err = np.abs(_mcf_gf(m, vals, 2, 2, gs, E0, zs, qr) - Gref).max()
        assert 1e-10 < err < 2e-8


def test_block_lanczos_p1_is_scalar_lanczos():
    """Block size 1: A_j and B_j are the scalar Lanczos coefficients a_j, b_{j+1}
    (src_ed/Hamiltonian/Lanczos.hpp scalar recurrence), for both block factorizations."""
    m, vals = TM.chain(6, U=4.0, mu=2.0)
    S = O.Sector(m, vals, 3, 3)
    E0, gs, *_ = S.lanczos_gs()
    T = O.Sector(m, vals, 4, 3)
    psi = O.create_destroy(6, 3, 3, 0, 1, gs)
    psi = psi / np.linalg.norm(psi)
    a, b = T.lanczos_green(psi, accur_cf=1e-30, max_iter=12)
    for qr in (False, True):
        A, B = O.block_lanczos(T.hv, [psi], qr=qr, M0=12)
        k = min(len(A), len(a)) - 1
        np.testing.assert_allclose([x[0, 0] for x in A[:k]], a[:k], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose([x[0, 0] for x in B[:k]], np.abs(b[1:k + 1]), rtol=1e-10)


def _chain_with_filled_site():
    """4-site chain plus a fifth, isolated site pushed 10 below: doubly occupied in the
    N=6 ground state, so c+_4 |psi0> = 0 and the particle block deflates at its last
    column (a 4 x 5 weight W in the matrix continued fraction)."""
    m, vals = O.hubbard_cluster(5, [(0, 1), (1, 2), (2, 3)], t=1.0, U=3.0, mu=1.2)
    m.add("e4", "one-body", [(4, 4, 1.0), (9, 9, 1.0)])
    vals["e4"] = -10.0
    return m, vals


@pytest.mark.parametrize("qr", [False, True])
def test_mcf_with_deflated_start_block(qr):
    """Particle block: W is 4 x 5 and the fraction is exact. Hole block: c_4|psi0> is an
    eigenvector of H (the site is decoupled), so the first residual block is singular and
    the recurrence stops after one floor, as the reference's does
    (src_ed/Hamiltonian/Lanczos.hpp:542-546, 703-712): G is then truncated there."""
    m, vals = _chain_with_filled_site()
    E0, gs, *_ = O.Sector(m, vals, 3, 3).lanczos_gs()
    zs = 1j * np.array([0.1, 0.7, 3.0]) + 0.02
    for pm in (1, -1):
        T = O.Sector(m, vals, 3 + pm, 3)
        phi = [O.create_destroy(5, 3, 3, a, pm, gs) for a in range(5)]
        q, W = O.mgs_weight(phi)
        A, B = O.block_lanczos(T.hv, q, qr=qr)
        if pm == 1:
            assert len(q) == 4 and W.shape == (4, 5)
            e, Q = O.dense_qmatrix(T.hv, T.dim, phi)
            A = O.mcf_shift(A, E0, True)
            for z in zs:
                assert np.abs(O.mcf_evaluate(A, B, W, z) - O.qmatrix_gf(e - E0, Q, z)).max() < 1e-11
        else:
            assert len(q) == 5 and len(A) == 1


def test_anomalous_dense_term_matches_bdg_closed_form():
    """The oracle's anomalous term, v (c_r c_c + h.c.) for an element (r, c, v) as the user
    gives it (the binding's 1/2 and the HS operator's 2 restated), on a noninteracting
    chain with singlet pairing: the ground-state energy of the S = 0 blocks equals the
    Bogoliubov-de Gennes closed form tr(h) + sum of the negative eigenvalues of
    [[h, D], [D^T, -h]] with D_ij = -v_ij for the element (i up, j down); a mistake of the
    factor 2 or of the operator's form changes it."""
    n = 4
    m, vals = hubbard_like_chain(n)
    L = m.L
    d, s = 0.45, 0.3
    m.add("D", "anomalous", [(i, i + 1 + L, 1.0) for i in range(n - 1)] + [(i + 1, i + L, 1.0) for i in range(n - 1)])
    m.add("S", "anomalous", [(i, i + L, 1.0) for i in range(n)])
    vals.update({"D": d, "S": s})
    st = O.block_states(L, [(k, k) for k in range(L + 1)])
    H, _ = O.dense_hamiltonian(m, vals, 0, 0, states=st)
    E0 = np.linalg.eigvalsh(H)[0]
    h = np.zeros((n, n))
    for (r, c, v) in m.operators["t"].elements:
        if r < L and c < L:
            h[r, c] += vals["t"] * v
            h[c, r] += vals["t"] * v if r != c else 0.0
    for (r, c, v) in m.operators["mu"].elements:
        if r < L:
            h[r, r] += vals["mu"] * v
    D = np.zeros((n, n))
    for name, z in (("D", d), ("S", s)):
        for (r, c, v) in m.operators[name].elements:
            D[r, c - L] += -z * v
    M = np.block([[h, D], [D.T, -h]])
    e = np.linalg.eigvalsh(M)
    closed = np.trace(h) + e[e < 0].sum()
    assert abs(E0 - closed) < 1e-12, (E0, closed)


def hubbard_like_chain(n):
    return O.hubbard_cluster(n, [(i, i + 1) for i in range(n - 1)], t=1.0, U=0.0, mu=0.7)


def test_dense_spinor_gf_reduces_to_dense_gf():
    """The spinor Green function with normal components on one (N_up, N_down) block is
    the plain dense G, and a Nambu component c^+_(i down) gives G^N_ii(z) = -G_down,ii(-z)
    (exchanging c and c^+ in both Lehmann terms); the down-spin G of this spin-symmetric
    chain is the up-spin one."""
    m, vals = O.hubbard_cluster(4, [(0, 1), (1, 2), (2, 3)], t=1.0, U=3.0, mu=1.5)
    L = m.L
    zs = np.array([0.4j, 1.3j, 0.5 + 0.3j])
    E0, G = O.dense_gf(m, vals, 2, 2, list(range(4)), zs)
    st0 = O.block_states(L, [(2, 2)])
    E1, Gs = O.dense_spinor_gf(m, vals, st0, [(i, False) for i in range(4)], O.block_states(L, [(1, 2)]),
                               O.block_states(L, [(3, 2)]), zs)
    assert abs(E0 - E1) < 1e-12
    np.testing.assert_allclose(Gs, G, atol=1e-12)
    _, Gn = O.dense_spinor_gf(m, vals, st0, [(i + L, True) for i in range(4)], O.block_states(L, [(2, 3)]),
                              O.block_states(L, [(2, 1)]), zs)
    _, Gm = O.dense_gf(m, vals, 2, 2, list(range(4)), -zs)
    for i in range(len(zs)):
        np.testing.assert_allclose(np.diag(Gn[i]), -np.diag(Gm[i]), atol=1e-12)
