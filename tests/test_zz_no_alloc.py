"""qed_lanczos_gs_noalloc, the Lanczos ground state that allocates no device memory (its
vectors and partial sums live in a fixed device array of the library), against
qed_lanczos_gs on the same sectors and start vectors: three calls in a row on sectors of
decreasing size (the array keeps the previous call's contents), one of them on the fused
Lanczos product; and the error for a sector beyond the array's limit."""
import numpy as np
import pytest

from tests import models as TM

pytestmark = pytest.mark.gpu

LIMIT = 1 << 21


def _run(entry, S, x0, max_iter=400):
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    psi = DeviceArray.from_host(x0)
    E0 = np.zeros(1)
    nit = np.zeros(1, np.int32)
    a = np.zeros(max_iter + 2)
    b = np.zeros(max_iter + 2)
    _lib.call(entry, S.h, 1e-12, max_iter, psi.ptr, _lib.dp(E0), _lib.ip(nit), _lib.dp(a), _lib.dp(b))
    n = int(nit[0])
    out = (float(E0[0]), n, a[:n].copy(), b[:n].copy(), psi.to_host())
    psi.free()
    return out


def test_noalloc_lanczos_matches_lanczos_gs(gpu_lib):
    from pyqcm_amd.sector import Sector
    cases = [
        # 3x4 Hubbard, N=12 Sz=0: D = 924^2 = 853,776 (rows shorter than 4096: ELL kernel)
        (TM.square(3, 4, tp=-0.3, U=8.0, mu=4.0), 6, 6),
        # 4x4 Hubbard, 8 up + 1 down: D = 12870 * 16 = 205,920 (fused Lanczos product)
        (TM.square(4, 4, tp=-0.3, U=8.0, mu=4.0), 8, 1),
        # 3x3 Hubbard, 5 up + 4 down: D = 126^2 = 15,876
        (TM.square(3, 3, tp=-0.3, U=8.0, mu=4.0), 5, 4),
    ]
    rng = np.random.default_rng(41)
    for (m, vals), nup, ndn in cases:
        S = Sector(m.L, nup, ndn, m.flat(vals))
        assert S.dim <= LIMIT
        x0 = rng.standard_normal(S.dim)
        x0 /= np.linalg.norm(x0)
        ref = _run("qed_lanczos_gs", S, x0)
        got = _run("qed_lanczos_gs_noalloc", S, x0)
        print("L=%d (%d, %d) D=%d: E0 %.14f / %.14f, iterations %d / %d, max|dpsi| %.1e"
              % (m.L, nup, ndn, S.dim, ref[0], got[0], ref[1], got[1], np.abs(ref[4] - got[4]).max()),
              flush=True)
        assert got[1] == ref[1]
        assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0])
        np.testing.assert_allclose(got[2], ref[2], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(got[3], ref[3], rtol=1e-12, atol=1e-12)
        assert np.abs(got[4] - ref[4]).max() <= 1e-10
        S.close()


def test_noalloc_lanczos_rejects_large_sector(gpu_lib):
    from pyqcm_amd import _lib
    from pyqcm_amd._lib import DeviceArray
    from pyqcm_amd.sector import Sector
    m, vals = TM.square(4, 4, tp=-0.3, U=8.0, mu=4.0)
    S = Sector(m.L, 7, 6, m.flat(vals))  # D = 11440 * 8008 > 2^21
    assert S.dim > LIMIT
    psi = DeviceArray(8).zeros()  # never touched: the size check comes first
    E0 = np.zeros(1)
    nit = np.zeros(1, np.int32)
    with pytest.raises(RuntimeError, match="exceeds the scratch limit of 2097152"):
        _lib.call("qed_lanczos_gs_noalloc", S.h, 1e-12, 10, psi.ptr, _lib.dp(E0), _lib.ip(nit), None, None)
    psi.free()
    S.close()
