"""Parity of the HIP sector path (basis, H.v, create/destroy, Lanczos, band Lanczos,
continued-fraction Lanczos) against the CPU oracle on the same inputs."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu


def gpu_sector(model, vals, nup, ndn):
    from pyqcm_amd.sector import Sector
    return Sector(model.L, nup, ndn, model.flat(vals))


# ------------------------------------------------------------- basis (bit-exact)

@pytest.mark.parametrize("L", [1, 4, 9, 12, 16, 18, 20])
def test_halfbasis_bit_exact(gpu_lib, L):
    from pyqcm_amd import sector as S
    for N in sorted({0, 1, L // 2, L // 2 + 1, L - 1, L}):
        if N < 0 or N > L:
            continue
        g = S.halfbasis(L, N)
        o = O.halfbasis(L, N)
        assert g.dtype == np.uint32 and np.array_equal(g, o)
        r = S.halfbasis_rank(L, N, o)
        assert np.array_equal(r, np.arange(len(o), dtype=np.uint32))
    # states with the wrong particle number or stray bits rank to dim (as ED_halfbasis::index)
    bad = np.array([0, (1 << L) - 1 if L < 32 else 0xFFFFFFFF, 1 << min(L, 31)], np.uint32)
    N = L // 2
    dim = len(O.halfbasis(L, N))
    r = S.halfbasis_rank(L, N, bad)
    for s, rk in zip(bad.tolist(), r.tolist()):
        valid = bin(s).count("1") == N and s < (1 << L)
        assert rk == (o_index(L, N, s) if valid else dim)


def o_index(L, N, s):
    b = O.halfbasis(L, N)
    i = np.searchsorted(b, s)
    return int(i) if i < len(b) and b[i] == s else len(b)


# ------------------------------------------------------------- H.v

HV_MODELS = {
    "chain4": (lambda: TM.chain(4), [(2, 2), (3, 1)]),
    "square3x3_tp": (lambda: TM.square(3, 3, tp=-0.3), [(5, 4), (4, 4)]),
    "square3x4_C1": (lambda: TM.square(3, 4, tp=0.0, U=8.0, mu=4.0), [(6, 6), (5, 6)]),
    "plaquette_8bath": (lambda: TM.plaquette_bath(
        {"U": 7.0, "mu": 3.5, **{"eb%d" % x: 0.4 * x - 1.8 for x in range(1, 9)},
         **{"tb%d" % x: 0.15 + 0.1 * x for x in range(1, 9)}}), [(6, 6), (5, 4)]),
}


@pytest.mark.parametrize("variant", [0, 1, 4, 5, 20])
@pytest.mark.parametrize("name", sorted(HV_MODELS))
def test_hv_matches_oracle(gpu_lib, name, variant):
    make, sectors = HV_MODELS[name]
    m, vals = make()
    rng = np.random.default_rng(42)
    for (nup, ndn) in sectors:
        So = O.Sector(m, vals, nup, ndn)
        Sg = gpu_sector(m, vals, nup, ndn)
        assert (Sg.dim, Sg.du, Sg.dd) == (So.dim, So.du, So.dd)
        x = rng.standard_normal(So.dim)
        y0 = rng.standard_normal(So.dim)
        ref = So.hv(x, y0)
        got = Sg.hv(x, y0, variant=variant)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-13 * scale


def test_hv_16_sites_half_filling(gpu_lib):
    """The 4x4, t'=-0.3, U=8, mu=4 Hamiltonian of the headline benchmark on its N=16, Sz=0
    sector (D = 165,636,900): the fused kernel against the oracle's reference loop order."""
    m, vals = TM.square(4, 4, tp=-0.3, U=8.0, mu=4.0)
    Sg = gpu_sector(m, vals, 8, 8)
    So = O.Sector(m, vals, 8, 8)
    assert Sg.dim == 165636900
    rng = np.random.default_rng(20261202)
    x = rng.standard_normal(Sg.dim)
    ref = So.hv(x)
    from pyqcm_amd._lib import DeviceArray
    dx = DeviceArray.from_host(x)
    outs = {}
    for variant in (0, 4, 5, 12, 20):
        dy = DeviceArray(Sg.dim).zeros()
        Sg.hv_dev(dx, dy, variant)
        got = dy.to_host()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (variant, err)
        outs[variant] = got
        dy.free()
    # the tiled row order changes only the order in which rows are produced
    assert np.array_equal(outs[12], outs[5])


# ------------------------------------------------------------- create / destroy

@pytest.mark.parametrize("L,nup,ndn", [(4, 2, 2), (6, 3, 2), (12, 6, 6)])
def test_apply_c_matches_oracle(gpu_lib, L, nup, ndn):
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import apply_c
    rng = np.random.default_rng(L)
    D = len(O.halfbasis(L, nup)) * len(O.halfbasis(L, ndn))
    x = rng.standard_normal(D)
    dx = DeviceArray.from_host(x)
    for a in range(2 * L):
        for pm in (-1, 1):
            tup, tdn = (nup + pm, ndn) if a < L else (nup, ndn + pm)
            if not (0 <= tup <= L and 0 <= tdn <= L):
                continue
            ref = O.create_destroy(L, nup, ndn, a, pm, x, z=0.7)
            dy = DeviceArray(len(ref)).zeros()
            apply_c(L, nup, ndn, a, pm, dx, dy, z=0.7)
            assert np.array_equal(dy.to_host(), ref)


# ------------------------------------------------------------- Lanczos ground state

@pytest.mark.parametrize("name,nup,ndn", [("chain4", 2, 2), ("square3x3_tp", 5, 4),
                                          ("square3x4_C1", 6, 6)])
def test_lanczos_e0_matches_oracle(gpu_lib, name, nup, ndn):
    """Same start vector on both sides (both normal streams reset: the library keeps one
    process-wide stream whose saved deviate carries over between draws, as the reference
    does), so the iteration counts agree up to the round-off-level stopping decision."""
    from pyqcm_amd import _lib
    m, vals = HV_MODELS[name][0]()
    So = O.Sector(m, vals, nup, ndn)
    Sg = gpu_sector(m, vals, nup, ndn)
    O.lib().or_reset_normal()
    _lib.call("qed_reset_start_vector_state")
    E0o, gso, no, *_ = So.lanczos_gs()
    E0g, psi, ng, a, b = Sg.lanczos_gs()
    assert abs(E0g - E0o) <= 1e-10 * abs(E0o)
    g = psi.to_host()
    assert abs(np.linalg.norm(g) - 1) < 1e-9
    assert abs(abs(g @ gso) - 1) < 1e-8
    assert abs(ng - no) <= 3


def test_lanczos_e0_reference_chain4(gpu_lib, golden):
    row = golden["chain4"]["rows"][0]
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    E0 = gpu_sector(m, vals, 2, 2).lanczos_gs()[0]
    assert abs(E0 - float(row["E0_1"])) < 1e-7


def test_lanczos_e0_reference_plaquette(gpu_lib, golden):
    g = golden["plaquette_8bath"]
    m, vals = TM.plaquette_bath(g["params"])
    E0 = gpu_sector(m, vals, 6, 6).lanczos_gs()[0]
    assert abs(E0 - g["E0"]) < 6e-5


def test_start_vector_is_reference_stream(gpu_lib):
    """The device start vector is the reference's libstdc++ stream (oracle is pinned to it)."""
    from pyqcm_amd import _lib
    m, vals = TM.chain(5)
    So = O.Sector(m, vals, 2, 3)
    Sg = gpu_sector(m, vals, 2, 3)
    O.lib().or_reset_normal()
    _lib.call("qed_reset_start_vector_state")
    for seed in (0, 3):
        np.testing.assert_allclose(Sg.start_vector(seed).to_host(), So.start_vector(seed),
                                   rtol=1e-13, atol=1e-16)


# ------------------------------------------------------------- Green-function Krylov solvers

def test_lanczos_green_matches_oracle(gpu_lib):
    """Continued fraction of c_0|0> (what build_cf feeds LanczosGreen): same depth, the
    leading coefficients and G(iw) agree with the oracle. (Later coefficients of a
    recurrence without reorthogonalization amplify rounding; G(iw) does not.)"""
    from pyqcm_amd._lib import DeviceArray
    m, vals = TM.square(3, 3, tp=-0.3)
    So0 = O.Sector(m, vals, 5, 4)
    E0, gs, *_ = So0.lanczos_gs()
    So = O.Sector(m, vals, 4, 4)
    Sg = gpu_sector(m, vals, 4, 4)
    psi = O.create_destroy(9, 5, 4, 0, -1, gs)
    nrm = psi @ psi
    psi /= math.sqrt(nrm)
    ao, bo = So.lanczos_green(psi)
    ag, bg = Sg.lanczos_green(DeviceArray.from_host(psi))
    assert abs(len(ao) - len(ag)) <= 1
    k = min(len(ao), len(ag), 12)
    np.testing.assert_allclose(ag[:k], ao[:k], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(bg[:k], bo[:k], rtol=1e-9, atol=1e-9)
    for z in 1j * np.array([0.1, 0.5, 2.0, 8.0]):
        ca, cb = O.cf_coefficients(ao, bo, E0, nrm, False)
        ga, gb = O.cf_coefficients(ag, bg, E0, nrm, False)
        assert abs(O.cf_evaluate(ga, gb, z) - O.cf_evaluate(ca, cb, z)) < 1e-8


def test_band_lanczos_matches_oracle(gpu_lib):
    """Device band Lanczos vs the oracle's restatement on c_i|0> of a 3x3 cluster:
    G(iw) within 1e-8."""
    from pyqcm_amd._lib import DeviceArray
    m, vals = TM.square(3, 3, tp=-0.3)
    nup, ndn = 5, 4
    So = O.Sector(m, vals, nup, ndn)
    E0, gs, *_ = So.lanczos_gs()
    zs = 1j * np.array([0.05, 0.3, 1.0, 3.0, 10.0])
    for pm in (-1, 1):
        To = O.Sector(m, vals, nup + pm, ndn)
        Tg = gpu_sector(m, vals, nup + pm, ndn)
        phi = [O.create_destroy(9, nup, ndn, a, pm, gs) for a in range(9)]
        eo, Qo = O.band_lanczos(To.hv, phi)
        eg, Qg = Tg.band_lanczos(DeviceArray.from_host(np.concatenate(phi)), 9)
        for z in zs:
            Go = O.qmatrix_gf(eo, Qo, z)
            Gg = O.qmatrix_gf(eg, Qg, z)
            assert np.abs(Gg - Go).max() < 1e-8


@pytest.mark.parametrize("seed,n", [(0, 1), (0, 7), (1, 100001), (42, 2 ** 20 + 3)])
def test_parallel_start_vector_stream_equals_standard_library(gpu_lib, seed, n):
    """The start-vector stream restated for parallel generation (polar method attempts
    located by jumps of the linear congruential engine) against std::normal_distribution
    itself, bit for bit, including a deviate carried from one call to the next."""
    a = np.zeros(n)
    b = np.zeros(n)
    gpu_lib.call("qed_util_normal_stream", seed, n, 0, 1, gpu_lib.dp(a))
    gpu_lib.call("qed_util_normal_stream", seed, n, 1, 8, gpu_lib.dp(b))
    np.testing.assert_array_equal(a, b)


# ------------------------------------------------------------- Davidson (several states)

def test_davidson_lowest_states_match_dense_spectrum(gpu_lib):
    """Davidson-Liu for the 4 lowest states of the 8-site chain with t' (N=8, Sz=0,
    D = 4,900) at the reference's default residual tolerance (accur_Davidson = 1e-5)
    against the spectrum of the dense H built independently from Jordan-Wigner matrices
    (eigenvalue error ~ residual^2 / gap); residuals and orthonormality of the vectors."""
    import ctypes
    from pyqcm_amd._lib import DeviceArray
    m, vals = TM.chain(8, tp=-0.3, U=3.0, mu=1.1)
    Sg = gpu_sector(m, vals, 4, 4)
    K, D = 4, Sg.dim
    psi = DeviceArray(K * D).zeros()
    ev = np.zeros(K)
    nit = ctypes.c_int()
    gpu_lib.call("qed_davidson", Sg.h, K, 1e-5, 1000, gpu_lib.dp(ev), psi.ptr, ctypes.byref(nit))
    H = O.dense_hamiltonian(m, vals, 4, 4)[0]
    ref = np.linalg.eigvalsh(H)[:K]
    print("Davidson %d iterations: %s vs dense %s" % (nit.value, ev, ref))
    np.testing.assert_allclose(ev, ref, rtol=0, atol=1e-8)
    V = psi.to_host().reshape(K, D)
    np.testing.assert_allclose(V @ V.T, np.eye(K), atol=1e-10)
    for k in range(K):
        r = H @ V[k] - ev[k] * V[k]
        assert np.linalg.norm(r) <= 1e-5, (k, np.linalg.norm(r))
    psi.free()


def test_davidson_12_sites_four_lowest_states(gpu_lib):
    """Davidson for the 4 lowest states of the 3x4 cluster (N=12, Sz=0, D = 853,776): the
    lowest against the Lanczos ground state (eigenvalue error <= residual^2 / gap), each
    residual |H v - e v| <= accur_Davidson, orthonormal vectors, ascending values."""
    import ctypes
    from pyqcm_amd._lib import DeviceArray
    m, vals = TM.square(3, 4, tp=-0.3, U=8.0, mu=4.0)
    S = gpu_sector(m, vals, 6, 6)
    K, D = 4, S.dim
    psi = DeviceArray(K * D).zeros()
    ev = np.zeros(K)
    nit = ctypes.c_int()
    gpu_lib.call("qed_davidson", S.h, K, 1e-5, 1000, gpu_lib.dp(ev), psi.ptr, ctypes.byref(nit))
    E0 = S.lanczos_gs()[0]
    print("Davidson %d iterations: %s, Lanczos E0 %.12f" % (nit.value, ev, E0))
    assert abs(ev[0] - E0) < 1e-8
    assert np.all(np.diff(ev) >= 0)
    y = DeviceArray(D)
    for k in range(K):
        v = psi.offset(k * D)
        y.zeros()
        gpu_lib.call("qed_hv", S.h, v, y.ptr)
        d = np.zeros(3)
        gpu_lib.call("qed_blas_dot", v, v, D, gpu_lib.dp(d[0:1]))
        gpu_lib.call("qed_blas_dot", v, y.ptr, D, gpu_lib.dp(d[1:2]))
        gpu_lib.call("qed_blas_dot", y.ptr, y.ptr, D, gpu_lib.dp(d[2:3]))
        assert abs(d[0] - 1.0) < 1e-12
        # the residual recomputed from the returned vector: within 10% of accur_Davidson (the
        # solver tests it on the vector before the last basis transform; this formula also
        # cancels to ~1e-12 of e^2)
        assert d[2] - 2 * ev[k] * d[1] + ev[k] ** 2 * d[0] <= (1.1e-5) ** 2
        for j in range(k):
            o = np.zeros(1)
            gpu_lib.call("qed_blas_dot", psi.offset(j * D), v, D, gpu_lib.dp(o))
            assert abs(o[0]) < 1e-12
    y.free()
    psi.free()
