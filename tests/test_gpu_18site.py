"""18-site cluster (3x6 Hubbard, t=1, U=4, half filling N=18, Sz=0): D = 48620^2 =
2,363,904,400 amplitudes, 18.9 GB per FP64 vector. A row (389 KB) does not fit in LDS,
so the H.v runs its global-memory form.

* H.v of a product vector u (x) w against the closed form assembled from one-spin
  oracle products: H (u x w) = u (x) (H_up w) + (H_dn u) (x) w + U n_up.n_dn (u x w),
  where H_up, H_dn are the oracle's Hamiltonians of the 9-electron one-spin sectors;
* the Lanczos ground state: residual |H psi - E0 psi| and the Rayleigh-quotient bound."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

L, NE = 18, 9


def chain_bonds():
    nn, _ = O.square_bonds(6, 3)  # 6 x 3 sites, index x + 6 y
    return nn


@pytest.fixture(scope="module")
def chain18(gpu_lib):
    from pyqcm_amd.sector import Sector
    model, vals = O.hubbard_cluster(L, chain_bonds(), t=1.0, U=4.0, mu=2.0)
    S = Sector(L, NE, NE, model.flat(vals))
    assert S.dim == 48620 ** 2
    return model, vals, S


def test_18site_hv_product_vector_closed_form(gpu_lib, chain18):
    from pyqcm_amd._lib import DeviceArray
    model, vals, S = chain18
    n = 48620
    rng = np.random.default_rng(18)
    u = rng.standard_normal(n)
    w = rng.standard_normal(n)
    print("18 sites: filling the product vector (%.1f GB)" % (n * n * 8 / 1e9), flush=True)
    X = DeviceArray(n * n)
    chunk = 1024
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        blk = np.ascontiguousarray(np.outer(u[r0:r1], w))
        gpu_lib.call("qed_memcpy_h2d", X.offset(r0 * n), blk.ctypes.data_as(ctypes.c_void_p), blk.nbytes)
    Y = DeviceArray(n * n)
    Y.zeros()
    S.hv_dev(X, Y)
    ms = np.zeros(1)
    Z = DeviceArray(n * n)
    gpu_lib.call("qed_hv_timed", S.h, -1, X.ptr, Z.ptr, 1, 3, gpu_lib.dp(ms))
    Z.free()
    print("18 sites: H.v %.1f ms (%.0f GB/s of 24 D bytes)" % (ms[0], 24.0 * n * n / ms[0] / 1e6), flush=True)
    # closed form from one-spin oracle sectors
    Hup = O.Sector(model, vals, NE, 0)
    Hdn = O.Sector(model, vals, 0, NE)
    hw = Hup.hv(w)
    hu = Hdn.hv(u)
    b = O.halfbasis(L, NE).astype(np.uint64)
    popc = np.vectorize(lambda v: bin(int(v)).count("1"))
    idx = rng.integers(0, n, size=(300, 2))
    got = np.zeros(len(idx))
    one = np.zeros(1)
    for k, (r, c) in enumerate(idx):
        gpu_lib.call("qed_memcpy_d2h", one.ctypes.data_as(ctypes.c_void_p), Y.offset(int(r) * n + int(c)), 8)
        got[k] = one[0]
    r, c = idx[:, 0], idx[:, 1]
    ref = u[r] * hw[c] + hu[r] * w[c] + 4.0 * popc(b[r] & b[c]) * u[r] * w[c]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("18 sites: sampled relative error %.2e" % err, flush=True)
    assert err < 1e-12
    X.free()
    Y.free()


def test_18site_lanczos_ground_state(gpu_lib, chain18):
    from pyqcm_amd._lib import DeviceArray
    model, vals, S = chain18
    D = S.dim
    psi = DeviceArray(D)
    gpu_lib.call("qed_start_vector", S.h, 0, psi.ptr)
    E0 = np.zeros(1)
    nit = ctypes.c_int()
    gpu_lib.call("qed_lanczos_gs", S.h, 1e-12, 1000, psi.ptr, gpu_lib.dp(E0), ctypes.byref(nit), None, None)
    y = DeviceArray(D)
    y.zeros()
    S.hv_dev(psi, y)
    dots = np.zeros(3)
    gpu_lib.call("qed_blas_dot", psi.ptr, psi.ptr, D, gpu_lib.dp(dots[0:1]))
    gpu_lib.call("qed_blas_dot", psi.ptr, y.ptr, D, gpu_lib.dp(dots[1:2]))
    gpu_lib.call("qed_blas_dot", y.ptr, y.ptr, D, gpu_lib.dp(dots[2:3]))
    nn, ph, yy = dots
    e = E0[0]
    res2 = yy - 2 * e * ph + e * e * nn
    print("18 sites: E0 = %.10f after %d iterations, <psi|psi> = %.15f, Rayleigh %.10f, |r|^2 = %.2e"
          % (e, nit.value, nn, ph / nn, res2), flush=True)
    assert abs(nn - 1.0) < 1e-10
    assert abs(ph / nn - e) < 1e-9 * abs(e)
    assert res2 < 1e-10 * e * e
    psi.free()
    y.free()
