"""Sectors of the spin-flip and anomalous (Nambu) mixings on the device
(qed_general_sector_create), against dense Jordan-Wigner Hamiltonians of the whole Fock
space restricted to the same states (oracle.dense_hamiltonian with the reference's
anomalous convention restated, oracle.dense_spinor_gf):

* the basis (union of the allowed (N_up, N_down) blocks, 64-bit states ascending);
* H.v against the dense matrix;
* the Lanczos ground state;
* the Green function of the spinor (c_up, c_down) for spin flip and (c_up, c^+_down) for
  the Nambu case, by band Lanczos from qed_general_apply start vectors.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu


def _chain_model(n, hx=0.0, d=0.0, s=0.0, U=3.0, mu=1.1, tp=-0.3):
    m, vals = TM.chain(n, tp=tp, U=U, mu=mu)
    L = m.L
    if hx:
        m.add("hx", "one-body", [(i, i + L, 1.0) for i in range(n)])
        vals["hx"] = hx
    if d:
        # singlet pairing on the chain's bonds
        m.add("D", "anomalous", [(i, i + 1 + L, 1.0) for i in range(n - 1)] + [(i + 1, i + L, 1.0) for i in range(n - 1)])
        vals["D"] = d
    if s:
        m.add("S", "anomalous", [(i, i + L, 1.0) for i in range(n)])
        vals["S"] = s
    return m, vals


def _device_flat(m, vals):
    """flat arrays as the device receives them: anomalous elements halved as by the
    binding (src_python/qcm_ED_wrap.hpp:179)"""
    from pyqcm_amd.sector import flatten_ops
    ops = []
    for name, op in m.operators.items():
        el = [(r, c, 0.5 * v if op.kind == "anomalous" else v) for (r, c, v) in op.elements]
        ops.append((name, op.kind, el))
    return flatten_ops(ops, vals)


def _labels(states, L):
    """64-bit states (up | down << 32) -> oracle Fock labels (up | down << L)"""
    st = np.asarray(states, np.uint64)
    return ((st & np.uint64(0xffffffff)) | ((st >> np.uint64(32)) << np.uint64(L))).astype(np.int64)


def _pairs(L, kind, value):
    if kind == "N":      # spin flip: N fixed
        return [(nu, value - nu) for nu in range(L + 1) if 0 <= value - nu <= L]
    if kind == "S":      # anomalous: S = N_up - N_down fixed
        return [(nu, nu - value) for nu in range(L + 1) if 0 <= nu - value <= L]
    # full: N parity
    return [(nu, nd) for nu in range(L + 1) for nd in range(L + 1) if (nu + nd) % 2 == value]


CASES = {
    "spin_flip": (dict(n=4, hx=0.4), "N", 4),
    "spin_flip_odd": (dict(n=4, hx=0.3, U=4.0), "N", 3),
    "nambu": (dict(n=4, d=0.5), "S", 0),
    "nambu_onsite": (dict(n=4, s=0.4, d=0.2, U=2.0), "S", 0),
    "full": (dict(n=4, hx=0.3, d=0.4), "P", 0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_general_sector_basis_hv_and_ground_state(gpu_lib, case):
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import GeneralSector
    kw, kind, value = CASES[case]
    m, vals = _chain_model(**kw)
    L = m.L
    pairs = _pairs(L, kind, value)
    S = GeneralSector(L, pairs, _device_flat(m, vals))
    ref = O.block_states(L, pairs)
    got = _labels(S.basis(), L)
    np.testing.assert_array_equal(got, ref)
    H, _ = O.dense_hamiltonian(m, vals, 0, 0, states=ref)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(S.dim)
    y0 = rng.standard_normal(S.dim)
    hy = S.hv(x, y0)
    err = np.abs(hy - (y0 + H @ x)).max() / np.abs(H @ x).max()
    E0, psi, *_ = S.lanczos_gs(accur=1e-12)
    e = np.linalg.eigvalsh(H)
    print(case, "dim", S.dim, "H.v rel err %.1e" % err, "E0 %.12f dense %.12f" % (E0, e[0]))
    assert err < 1e-13
    assert abs(E0 - e[0]) <= 1e-10 * abs(e[0])
    psi.free()
    S.close()


def _spinor(L, n_sites, kind):
    if kind == "N":  # spin flip: (c_i up, c_i down)
        return [(i, False) for i in range(n_sites)] + [(i + L, False) for i in range(n_sites)]
    # Nambu: (c_i up, c^+_i down)
    return [(i, False) for i in range(n_sites)] + [(i + L, True) for i in range(n_sites)]


@pytest.mark.parametrize("case", ["spin_flip", "nambu", "nambu_onsite"])
def test_general_sector_spinor_green_function(gpu_lib, case):
    """G of the spinor from band Lanczos on the target sectors (one particle less or more
    for spin flip, S - 1 and S + 1 for the Nambu spinor) against the dense G."""
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import GeneralSector
    kw, kind, value = CASES[case]
    m, vals = _chain_model(**kw)
    L = m.L
    n = kw["n"]
    flat = _device_flat(m, vals)
    B = GeneralSector(L, _pairs(L, kind, value), flat)
    e0 = np.linalg.eigvalsh(O.dense_hamiltonian(m, vals, 0, 0, states=O.block_states(L, _pairs(L, kind, value)))[0])
    assert e0[1] - e0[0] > 1e-6
    E0, psi, *_ = B.lanczos_gs(accur=1e-12)
    spinor = _spinor(L, n, kind)
    zs = np.array([0.3j, 1.0j, 2.5j, 0.6 + 0.2j])
    G = np.zeros((len(zs), len(spinor), len(spinor)), complex)
    tgt_pairs = {}
    for pm in (-1, 1):
        tp = _pairs(L, kind, value + pm) if kind != "P" else None
        tgt_pairs[pm] = tp
        T = GeneralSector(L, tp, flat)
        phis = DeviceArray(len(spinor) * T.dim).zeros()
        for k, (a, nambu) in enumerate(spinor):
            # the Nambu component's destruction is a creation (model::create_or_destroy)
            B.apply(T, a, -pm if nambu else pm, psi, phis.offset(k * T.dim))
        e, Q = T.band_lanczos(phis, len(spinor))
        e = e - E0
        if pm < 0:
            e = -e
        for i, z in enumerate(zs):
            G[i] += (Q / (z - e)) @ Q.T
        phis.free()
        T.close()
    E0d, Gd = O.dense_spinor_gf(m, vals, O.block_states(L, _pairs(L, kind, value)), spinor,
                                O.block_states(L, tgt_pairs[-1]), O.block_states(L, tgt_pairs[1]), zs)
    print(case, "E0 %.12f dense %.12f" % (E0, E0d), "max |dG| %.2e" % np.abs(G - Gd).max())
    assert abs(E0 - E0d) <= 1e-10 * abs(E0d)
    np.testing.assert_allclose(G, Gd, atol=1e-9)
    psi.free()
    B.close()


@pytest.mark.parametrize("case", ["nambu_bath", "spin_flip_bath", "full_bath"])
def test_general_lookup_tables_equal_binary_search(gpu_lib, case):
    """2x2 plaquette + 4 bath orbitals (L = 8) with pairing or spin flip: the product that
    locates targets through the rank tables is bit-identical to the one that binary
    searches the sorted basis, and the Lanczos E0 of the Nambu sector is the lowest of the
    blocks' union (the pairing-free limit: the factorized sectors' lowest)."""
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import GeneralSector
    p = TM.random_bath_params(np.random.default_rng(4), n_bath=4)
    m, vals = TM.plaquette_bath(p, n_bath=4)
    L = m.L
    if case in ("nambu_bath", "full_bath"):
        m.add("D", "anomalous", [(0, 1 + L, 1.0), (1, 0 + L, 1.0), (2, 3 + L, 1.0), (3, 2 + L, 1.0),
                                 (0, 2 + L, -1.0), (2, 0 + L, -1.0), (1, 3 + L, -1.0), (3, 1 + L, -1.0)])
        vals["D"] = 0.3
    if case in ("spin_flip_bath", "full_bath"):
        m.add("hx", "one-body", [(i, i + L, 1.0) for i in range(4)])
        vals["hx"] = 0.25
    kind, value = {"nambu_bath": ("S", 0), "spin_flip_bath": ("N", 8), "full_bath": ("P", 0)}[case]
    S = GeneralSector(L, _pairs(L, kind, value), _device_flat(m, vals))
    rng = np.random.default_rng(9)
    x = DeviceArray.from_host(rng.standard_normal(S.dim))
    y0 = rng.standard_normal(S.dim)
    ya, yb = DeviceArray.from_host(y0), DeviceArray.from_host(y0)
    S.hv_dev(x, ya)
    _lib.call("qed_hv_variant", S.h, 0, x.ptr, yb.ptr)
    a, b = ya.to_host(), yb.to_host()
    print(case, "dim", S.dim, "max |table - search| %.1e" % np.abs(a - b).max())
    np.testing.assert_array_equal(a, b)
    for v in (x, ya, yb):
        v.free()
    S.close()


@pytest.mark.parametrize("case", ["spin_flip", "nambu_onsite", "full"])
def test_general_sector_diagonal_and_davidson(gpu_lib, case):
    """The diagonal of H on a mixed-number sector (one-body diagonal and density terms)
    against the dense Jordan-Wigner matrix, and Davidson-Liu's 4 lowest states against its
    spectrum: values, residuals <= accur_Davidson, orthonormal vectors."""
    import ctypes
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import GeneralSector
    kw, kind, value = CASES[case]
    m, vals = _chain_model(**kw)
    L = m.L
    pairs = _pairs(L, kind, value)
    S = GeneralSector(L, pairs, _device_flat(m, vals))
    H, _ = O.dense_hamiltonian(m, vals, 0, 0, states=O.block_states(L, pairs))
    np.testing.assert_allclose(S.diag(), np.diag(H), rtol=0, atol=1e-12 * np.abs(H).max())
    K = 4
    psi = DeviceArray(K * S.dim).zeros()
    ev = np.zeros(K)
    nit = ctypes.c_int()
    gpu_lib.call("qed_davidson", S.h, K, 1e-5, 1000, gpu_lib.dp(ev), psi.ptr, ctypes.byref(nit))
    V = psi.to_host().reshape(K, S.dim)
    e = np.linalg.eigvalsh(H)[:K]
    print(case, "dim", S.dim, "Davidson %d iterations" % nit.value, ev, "dense", e)
    np.testing.assert_allclose(ev, e, rtol=0, atol=1e-8)
    np.testing.assert_allclose(V @ V.T, np.eye(K), atol=1e-10)
    for k in range(K):
        assert np.linalg.norm(H @ V[k] - ev[k] * V[k]) <= 1.1e-5
    psi.free()
    S.close()


def test_general_nambu_plaquette_8bath_sector(gpu_lib):
    """The 2x2 plaquette + 8 bath orbitals with d-wave pairing, Nambu sector S = 0
    (D = C(24, 12) = 2,704,156): the rank-table product is bit-identical to the binary
    search at full size, H is symmetric (<a|Hb> = <Ha|b> to round-off), and the Lanczos
    ground state has residual |H psi - E0 psi| <= 1e-6 |E0|; 20 further products are
    timed for the profile of k_hv_general."""
    import time
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import GeneralSector
    p = TM.random_bath_params(np.random.default_rng(4), n_bath=8)
    m, vals = TM.plaquette_bath(p, n_bath=8)
    L = m.L
    m.add("D", "anomalous", [(0, 1 + L, 1.0), (1, 0 + L, 1.0), (2, 3 + L, 1.0), (3, 2 + L, 1.0),
                             (0, 2 + L, -1.0), (2, 0 + L, -1.0), (1, 3 + L, -1.0), (3, 1 + L, -1.0)])
    vals["D"] = 0.3
    S = GeneralSector(L, _pairs(L, "S", 0), _device_flat(m, vals))
    assert S.dim == 2704156
    rng = np.random.default_rng(11)
    xa, xb = rng.standard_normal(S.dim), rng.standard_normal(S.dim)
    ha, hb = S.hv(xa), S.hv(xb)
    sym = abs(xa @ hb - ha @ xb) / abs(xa @ hb)
    x = DeviceArray.from_host(xa)
    y0, y1 = DeviceArray(S.dim).zeros(), DeviceArray(S.dim).zeros()
    S.hv_dev(x, y0)
    _lib.call("qed_hv_variant", S.h, 0, x.ptr, y1.ptr)
    np.testing.assert_array_equal(y0.to_host(), y1.to_host())
    E0, psi, nit, *_ = S.lanczos_gs(accur=1e-12)
    v = psi.to_host()
    r = np.linalg.norm(S.hv(v) - E0 * v)
    t0 = time.perf_counter()
    for _ in range(20):
        S.hv_dev(x, y0)
    _lib.call("qed_sync")
    ms = (time.perf_counter() - t0) / 20 * 1e3
    print("Nambu plaquette+8 bath: dim %d, symmetry %.1e, E0 %.10f in %d iterations, residual %.1e, "
          "%.3f ms per product (host clock)" % (S.dim, sym, E0, nit, r, ms))
    assert sym < 1e-12
    assert r <= 1e-6 * abs(E0)
    for a in (x, y0, y1, psi):
        a.free()
    S.close()
