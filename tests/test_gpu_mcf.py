"""Block Lanczos and the matrix continued fraction (GF_method 'M') on the GPU, against the
oracle's restatement and against brute-force Lehmann sums."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM
from tests.test_gpu_ed import register, with_sites

pytestmark = pytest.mark.gpu


def _device_block(gpu_lib, vecs):
    from pyqcm_amd._lib import DeviceArray
    flat = np.ascontiguousarray(np.concatenate(vecs))
    return DeviceArray.from_host(flat)


@pytest.mark.parametrize("qr", [0, 1])
def test_block_lanczos_vs_oracle(gpu_lib, qr):
    """8-site chain, particle block c+_i|psi0> (p = 8, D = 3920): floors against the oracle
    (same recurrence, same stopping test). The first floors agree to round-off; later ones
    drift as orthogonality is lost in both, so the test compares 8 floors and then the
    Green functions built from all floors."""
    from pyqcm_amd.sector import Sector
    m, vals = TM.chain(8, tp=-0.3, U=3.0, mu=1.1)
    E0, gs, *_ = O.Sector(m, vals, 4, 4).lanczos_gs()
    To = O.Sector(m, vals, 5, 4)
    phi = [O.create_destroy(8, 4, 4, a, 1, gs) for a in range(8)]
    q, W = O.mgs_weight(phi)
    A_o, B_o = O.block_lanczos(To.hv, q, qr=bool(qr))
    S = Sector(8, 5, 4, m.flat(vals))
    p, cap = len(q), 700
    dq = _device_block(gpu_lib, q)
    A = np.zeros(cap * p * p)
    B = np.zeros(cap * p * p)
    M = ctypes.c_int()
    gpu_lib.call("qed_block_lanczos", S.h, dq.ptr, p, 0, 600, 1e-7, 1e-12, qr, gpu_lib.dp(A), gpu_lib.dp(B),
                 cap, ctypes.byref(M))
    dq.free()
    M = M.value
    print("floors: gpu %d, oracle %d" % (M, len(A_o)))
    assert abs(M - len(A_o)) <= p
    A = A[:M * p * p].reshape(M, p, p).transpose(0, 2, 1)  # column-major blocks
    B = B[:M * p * p].reshape(M, p, p).transpose(0, 2, 1)
    for j in range(8):
        np.testing.assert_allclose(A[j], A_o[j], atol=1e-10)
        np.testing.assert_allclose(B[j], B_o[j], atol=1e-10)
    zs = 1j * np.array([0.1, 0.5, 2.0])
    As, Ao = O.mcf_shift(list(A), E0, True), O.mcf_shift(A_o, E0, True)
    for z in zs:
        g = O.mcf_evaluate(As, list(B), W, z)
        go = O.mcf_evaluate(Ao, B_o, W, z)
        assert np.abs(g - go).max() < 1e-10


def test_mcf_eval_kernel_vs_oracle(gpu_lib):
    """Device evaluation of W^T F_0 W on random blocks (non-symmetric B, non-square W,
    transposed accumulation) against the oracle's recursion."""
    rng = np.random.default_rng(3)
    for pa, q, M in ((3, 5, 7), (8, 8, 40), (13, 16, 25)):
        A = [0.5 * (x + x.T) for x in rng.standard_normal((M, pa, pa))]
        B = [np.triu(x) for x in rng.standard_normal((M, pa, pa))]
        W = rng.standard_normal((pa, q))
        zs = np.array([0.3j, 1.0 + 0.5j, -2.0 + 0.05j, 4j])
        Ac = np.ascontiguousarray(np.array(A).transpose(0, 2, 1)).ravel()
        Bc = np.ascontiguousarray(np.array(B).transpose(0, 2, 1)).ravel()
        Wc = np.ascontiguousarray(W.T).ravel()
        for tr in (0, 1):
            base = rng.standard_normal((len(zs), q, q)) + 1j * rng.standard_normal((len(zs), q, q))
            G = np.ascontiguousarray(base.transpose(0, 2, 1)).astype(np.complex128)
            zz = np.ascontiguousarray(zs.astype(np.complex128))
            gpu_lib.call("qed_mcf_eval", pa, q, M, gpu_lib.dp(Ac), gpu_lib.dp(Bc), gpu_lib.dp(Wc), len(zs),
                         gpu_lib.dp(zz.view(np.float64)), tr, gpu_lib.dp(G.view(np.float64)))
            got = G.transpose(0, 2, 1)
            for i, z in enumerate(zs):
                ref = O.mcf_evaluate(A, B, W, z)
                ref = base[i] + (ref.T if tr else ref)
                err = np.abs(got[i] - ref).max() / max(1.0, np.abs(ref).max())
                assert err < 1e-11, (pa, q, M, tr, err)


def _with_method(qcm, qr):
    qcm.set_global_parameter("GF_method", "M")
    qcm.set_global_parameter("block_Lanczos_QR", bool(qr))


def _restore(qcm):
    qcm.set_global_parameter("GF_method", "L")
    qcm.set_global_parameter("block_Lanczos_QR", False)


@pytest.mark.parametrize("qr", [0, 1])
def test_gf_method_M_exhaustive_matches_bruteforce(gpu_lib, qr):
    """4-site chain: the block Krylov spaces are exhausted, so the matrix continued
    fraction is the exact Green function."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, tp=-0.3, U=3.0, mu=1.2)
    _with_method(qcm, qr)
    try:
        lab = register(with_sites(m, 4), vals, "R0:N4:S0")
        zs = np.array([0.05j, 0.5j, 2j, 1.0 + 0.1j, -2.0 + 0.3j])
        G = qcm.Green_function(zs, 0, lab)
        with pytest.raises(Exception):
            qcm.qmatrix(lab)
    finally:
        _restore(qcm)
    _, Gref = O.dense_gf(m, vals, 2, 2, [0, 1, 2, 3], zs)
    np.testing.assert_allclose(G, Gref, atol=1e-11)


@pytest.mark.parametrize("qr", [0, 1])
def test_gf_method_M_matches_band_lanczos(gpu_lib, qr):
    """3x3 Hubbard (N=9, S=1: spin up and down parts): the matrix continued fraction and
    the band-Lanczos Lehmann form of the same instance agree to the convergence of both
    recurrences (stopping test 1e-12 on the lowest eigenvalue: ~1e-8 in G)."""
    from pyqcm_amd import qcm
    m, vals = TM.square(3, 3, tp=-0.3, U=8.0, mu=4.0)
    zs = 1j * np.array([0.05, 0.2, 0.7, 2.0, 6.0])
    lab_l = register(with_sites(m, 9), vals, "R0:N9:S1")
    GL = [qcm.Green_function(zs, s, lab_l) for s in (0, 1)]
    _with_method(qcm, qr)
    try:
        lab_m = register(with_sites(m, 9), vals, "R0:N9:S1")
        GM = [qcm.Green_function(zs, s, lab_m) for s in (0, 1)]
    finally:
        _restore(qcm)
    for s in (0, 1):
        err = np.abs(GM[s] - GL[s]).max()
        print("spin %d: max |G_M - G_L| = %.3g" % (s, err))
        assert err < 5e-8


@pytest.mark.parametrize("qr", [0, 1])
def test_gf_method_M_deflated_blocks_match_oracle(gpu_lib, qr):
    """4-site chain plus a decoupled, doubly occupied fifth site: the particle start block
    deflates (W is 4 x 5) and the hole recurrence stops after one floor (a decoupled
    eigenvector makes the first residual block singular, as in the reference). The
    instance-level G against the oracle's restatement of the same steps."""
    from pyqcm_amd import qcm
    from tests.test_oracle import _chain_with_filled_site, _mcf_gf
    m, vals = _chain_with_filled_site()
    zs = 1j * np.array([0.1, 0.7, 3.0]) + 0.02
    _with_method(qcm, qr)
    try:
        lab = register(with_sites(m, 5), vals, "R0:N6:S0")
        G = qcm.Green_function(zs, 0, lab)
    finally:
        _restore(qcm)
    E0, gs, *_ = O.Sector(m, vals, 3, 3).lanczos_gs()
    Go = _mcf_gf(m, vals, 3, 3, gs, E0, zs, bool(qr))
    err = np.abs(G - Go).max()
    print("max |G_gpu - G_oracle| = %.3g" % err)
    assert err < 1e-10


def _combine(qcm, method, qr, on=True):
    qcm.set_global_parameter("GF_method", method)
    qcm.set_global_parameter("block_Lanczos_QR", bool(qr))
    qcm.set_global_parameter("combine_mcf", bool(on))


def _uncombine(qcm):
    qcm.set_global_parameter("GF_method", "L")
    qcm.set_global_parameter("block_Lanczos_QR", False)
    qcm.set_global_parameter("combine_mcf", False)


@pytest.mark.parametrize("qr", [0, 1])
def test_combine_mcf_exhaustive_matches_bruteforce(gpu_lib, qr):
    """4-site chain, GF_method 'M' with combine_mcf: the particle and hole fractions are
    exact (exhausted Krylov spaces) and so is their combination by block Lanczos on the
    direct sum (combine_via_lanczos); the exported (W, A, B) evaluates to the same G."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, tp=-0.3, U=3.0, mu=1.2)
    zs = np.array([0.05j, 0.5j, 2j, 1.0 + 0.1j, -2.0 + 0.3j])
    _combine(qcm, "M", qr)
    try:
        lab = register(with_sites(m, 4), vals, "R0:N4:S0")
        G = qcm.Green_function(zs, 0, lab)
        W, A, B = qcm.combined_mcf(lab)
    finally:
        _uncombine(qcm)
    _, Gref = O.dense_gf(m, vals, 2, 2, [0, 1, 2, 3], zs)
    print("combined MCF: %d floors of %d x %d, max |G - G_exact| = %.3g"
          % (len(A), A[0].shape[0], A[0].shape[1], np.abs(G - Gref).max()))
    np.testing.assert_allclose(G, Gref, atol=1e-10)
    for i, z in enumerate(zs):
        g = O.mcf_evaluate([a.real for a in A], [b.real for b in B], W.real, z)
        np.testing.assert_allclose(g, G[i], atol=1e-12)


@pytest.mark.parametrize("qr", [0, 1])
def test_combine_mcf_from_qmatrix_matches_lehmann(gpu_lib, qr):
    """3x3 Hubbard, N = 9, S = 1, GF_method 'L' with combine_mcf (Q_matrix_to_mcf): the
    Lehmann form rewritten as one matrix continued fraction by block Lanczos on diag(e).
    The recurrence stops once the lowest eigenvalue of its projection has converged, as
    the reference's does, so the rewriting is truncated: against the oracle's restatement
    of the same steps on the same Lehmann form to 1e-8, against the Lehmann G to 1e-5."""
    from pyqcm_amd import qcm
    m, vals = TM.square(3, 3, tp=-0.3, U=8.0, mu=4.0)
    zs = 1j * np.array([0.05, 0.2, 0.7, 2.0, 6.0]) + 0.1
    lab_l = register(with_sites(m, 9), vals, "R0:N9:S1")
    GL = [qcm.Green_function(zs, s, lab_l) for s in (0, 1)]
    _combine(qcm, "L", qr)
    try:
        lab_c = register(with_sites(m, 9), vals, "R0:N9:S1")
        GC = [qcm.Green_function(zs, s, lab_c) for s in (0, 1)]
        W, A, B = qcm.combined_mcf(lab_c)
        with pytest.raises(Exception, match="not available"):
            qcm.qmatrix(lab_c)
    finally:
        _uncombine(qcm)
    for s in (0, 1):
        err = np.abs(GC[s] - GL[s]).max()
        e, Q = qcm.qmatrix(lab_l, s)
        q, W = O.mgs_weight([Q[:, a].real for a in range(9)])
        A_o, B_o = O.block_lanczos(lambda x: e * x, q, qr=bool(qr), M0=len(e))
        Go = np.array([O.mcf_evaluate(A_o, B_o, W, z) for z in zs])
        err_o = np.abs(GC[s] - Go).max()
        print("spin %d: %d floors (oracle %d), max |G_combined - G_L| = %.3g, vs oracle %.3g"
              % (s, len(A), len(A_o), err, err_o))
        assert err_o < 1e-8
        assert err < 1e-5


def test_combined_mcf_errors(gpu_lib):
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=3.0, mu=1.5)
    lab = register(with_sites(m, 4), vals, "R0:N4:S0")
    with pytest.raises(Exception, match="does not use the MCF format"):
        qcm.combined_mcf(lab)
    qcm.set_global_parameter("GF_method", "M")
    try:
        lab2 = register(with_sites(m, 4), vals, "R0:N4:S0")
        with pytest.raises(Exception, match="is not built unless the global parameter 'combine_mcf'"):
            qcm.combined_mcf(lab2)
    finally:
        _uncombine(qcm)
