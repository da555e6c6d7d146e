"""Instance files (ED::write_instance_hdf5 / read_instance_hdf5): a solved instance written
to HDF5 and read into a fresh instance gives the same Green function, self-energy and
Lehmann representation, bit for bit; the file's layout and values, read back with HDF5's
own h5dump, are those of the reference's writers (src_ed/model_instance.hpp:1339-1377,
src_ed/state.hpp:46-69, src_ed/Q_matrix.hpp:320-331, src_ed/continued_fraction_set.cpp:
136-153, src_ed/mcf_set.hpp:162-176, src_ed/matrix_continued_fraction.hpp:511-527)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from tests import models as TM
from tests.test_gpu_ed import register, with_sites

pytestmark = pytest.mark.gpu

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
ZS = np.array([0.05j, 0.4j, 2.0j, 1.0 + 0.2j, -1.5 + 0.1j])


def _method(qcm, m):
    qcm.set_global_parameter("GF_method", m)


def _dump_values(path, obj, attr=False):
    """Values of a dataset (or, with attr, an attribute), through h5dump (an independent
    reader)."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-w", "0", "-a" if attr else "-d", obj, path],
                         capture_output=True, text=True, check=True).stdout
    body = out[out.index("DATA {") + 6:]
    if '"' in body:
        return re.findall(r'"([^"]*)"', body)[0]
    body = re.sub(r"\(\d+\):", " ", body.split("}")[0])
    return np.array([float(x) for x in re.findall(r"[-+]?\d[\d.]*(?:[eE][-+]?\d+)?", body)])


def _attr(path, obj):
    return _dump_values(path, obj, attr=True)


def _solve_write_read(qcm, tmp_path, m, vals, n_sites, sectors, method):
    _method(qcm, method)
    try:
        a = register(with_sites(m, n_sites), vals, sectors)
        E0, sec = qcm.ground_state_solve(a)
        spins = (0, 1) if "S1" in sectors or "S-1" in sectors else (0,)
        G = [qcm.Green_function(ZS, s, a) for s in spins]
        S = [qcm.self_energyC(ZS, s, a) for s in spins]
        path = str(tmp_path / "instance.h5")
        qcm.write_instance_to_hdf5(path, "cluster_0", a)
        b = register(with_sites(m, n_sites), vals, sectors)
        qcm.read_instance_from_hdf5(path, "cluster_0", b)
        Gb = [qcm.Green_function(ZS, s, b) for s in spins]
        Sb = [qcm.self_energyC(ZS, s, b) for s in spins]
        E0b, secb = qcm.ground_state_solve(b)
        assert qcm.Potthoff_functional(b) == qcm.Potthoff_functional(a)
    finally:
        _method(qcm, "L")
    for x, y in zip(G + S, Gb + Sb):
        np.testing.assert_array_equal(x, y)
    assert secb == sec
    assert abs(E0b - E0) <= 1e-12 * abs(E0)
    return a, b, path, E0, sec


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_round_trip_plaquette_8bath(gpu_lib, tmp_path, method):
    """2x2 plaquette + 8 bath orbitals, N = 12, S = 0 (D = 853,776): G(z) and Sigma(z) of
    the instance read from the file equal those of the solved instance, bit for bit."""
    from pyqcm_amd import qcm
    rng = np.random.default_rng(12)
    m, vals = TM.plaquette_bath(TM.random_bath_params(rng))
    a, b, path, E0, sec = _solve_write_read(qcm, tmp_path, m, vals, 4, "R0:N12:S0", method)
    if method == "L":
        ea, Qa = qcm.qmatrix(a)
        eb, Qb = qcm.qmatrix(b)
        np.testing.assert_array_equal(ea, eb)
        np.testing.assert_array_equal(Qa, Qb)


@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_round_trip_two_states_spin_mixing(gpu_lib, tmp_path, method):
    """5-site chain, sectors N = 5, S = +-1 at temperature 0.1: two degenerate states of
    weight 1/2 and spin-up and spin-down Green functions (mixing 4): both states and both
    spins survive the round trip."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(5, U=3.0, mu=1.1, tp=-0.3)
    qcm.set_global_parameter("temperature", 0.1)
    try:
        a, b, path, E0, sec = _solve_write_read(qcm, tmp_path, m, vals, 5, "R0:N5:S1/R0:N5:S-1", method)
    finally:
        qcm.set_global_parameter("temperature", 0.0)
    assert sec.count("/") == 1
    if H5DUMP:
        assert int(_attr(path, "/cluster_0/nstates")[0]) == 2
        assert int(_attr(path, "/cluster_0/mixing")[0]) == 4
        for k in (0, 1):
            w = _attr(path, "/cluster_0/state_%d/weight" % k)[0]
            assert abs(w - 0.5) < 1e-6
        out = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True).stdout
        assert "/cluster_0/state_1/gf_down" in out


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
@pytest.mark.parametrize("method", ["L", "F", "M"])
def test_layout_and_values_match_reference_writers(gpu_lib, tmp_path, method):
    """The object paths and attributes of the reference's writers, and the values, read by
    h5dump: the ground-state energy, the Lehmann poles and weights against qcm.qmatrix,
    the diagonal continued fraction against G(z) at one frequency, the MCF weight matrix
    against the spectral sum rule (W^T W of particles + holes = 1 for the Green function's
    anticommutator)."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    a, b, path, E0, sec = _solve_write_read(qcm, tmp_path, m, vals, 4, "R0:N4:S0", method)
    fmt = {"L": "bl", "F": "cf", "M": "mcf"}[method]
    g = "/cluster_0"
    assert abs(_attr(path, g + "/GS_energy")[0] - E0) < 1e-14
    assert _attr(path, g + "/GF_format") == fmt
    assert _attr(path, g + "/GS_sectors") == sec
    assert _attr(path, g + "/state_0/sec") == "R0:N4:S0"
    assert _attr(path, g + "/state_0/gf_format") == fmt
    assert abs(_attr(path, g + "/state_0/energy")[0] - E0) < 1e-14
    assert int(_attr(path, g + "/complex_HS")[0]) == 0
    assert _attr(path, g + "/params/U")[0] == 4.0
    gf = g + "/state_0/gf"
    if fmt == "bl":
        e, Q = qcm.qmatrix(a)
        M = len(e)
        assert int(_attr(path, gf + "/nblocks")[0]) == 1
        assert int(_attr(path, gf + "/block_0/L")[0]) == 4
        assert int(_attr(path, gf + "/block_0/M")[0]) == M
        np.testing.assert_array_equal(_dump_values(path, gf + "/block_0/e"), e)
        v = _dump_values(path, gf + "/block_0/v/data").reshape(M, 4).T  # column-major L x M
        np.testing.assert_array_equal(v, Q.real.T)
    elif fmt == "cf":
        z = ZS[1]
        Gz = qcm.Green_function(np.array([z]), 0, a)[0]
        for o in range(4):
            tot = 0.0
            for part in ("e", "h"):
                aa = _dump_values(path, gf + "/r0_a%d_b%d/%s/a" % (o, o, part))
                bb = _dump_values(path, gf + "/r0_a%d_b%d/%s/b" % (o, o, part))
                f = 0.0
                for i in range(len(aa) - 1, -1, -1):
                    f = bb[i] / (z - aa[i] - f)
                tot += f
            assert abs(tot - Gz[o, o]) < 1e-13
        for part in ("e", "h"):
            _dump_values(path, gf + "/r0_a3_b2/%s/a" % part)
    else:
        assert int(_attr(path, gf + "/nblocks")[0]) == 1
        S = np.zeros((4, 4))
        for part in ("e", "h"):
            p = int(_attr(path, gf + "/block_0/%s/p" % part)[0])
            r = int(_attr(path, gf + "/block_0/%s/W/data/r" % part)[0])
            c = int(_attr(path, gf + "/block_0/%s/W/data/c" % part)[0])
            assert (r, c) == (p, 4)
            W = _dump_values(path, gf + "/block_0/%s/W/data/re" % part).reshape(c, r).T
            assert not _dump_values(path, gf + "/block_0/%s/W/data/im" % part).any()
            A0 = _dump_values(path, gf + "/block_0/%s/A_0/data/data" % part).reshape(p, p)
            np.testing.assert_allclose(A0, A0.T, atol=1e-12)
            S += W.T @ W
        np.testing.assert_allclose(S, np.eye(4), atol=1e-10)


def test_overwrite_parameter_warning_and_errors(gpu_lib, tmp_path, capfd):
    """Writing a group twice replaces it; reading into an instance whose parameters differ
    prints the reference's warning; a missing file, a missing group and an unknown label
    raise with the reference's messages; a read instance has no ground-state vectors, so its
    averages come from the Green function and the susceptibility is refused."""
    from pyqcm_amd import qcm
    m, vals = TM.chain(4, U=4.0, mu=2.0)
    a = register(with_sites(m, 4), vals, "R0:N4:S0")
    path = str(tmp_path / "i.h5")
    qcm.write_instance_to_hdf5(path, "cluster_0", a)
    qcm.write_instance_to_hdf5(path, "cluster_0", a)
    qcm.write_instance_to_hdf5(path, "cluster_1", a)
    other = dict(vals, U=5.0)
    b = register(with_sites(m, 4), other, "R0:N4:S0")
    capfd.readouterr()
    qcm.read_instance_from_hdf5(path, "cluster_1", b)
    out = capfd.readouterr().out
    assert "WARNING : The value of U from the HDF5 solution (4) differs from the expected value (5)." in out
    np.testing.assert_array_equal(qcm.Green_function(ZS, 0, a), qcm.Green_function(ZS, 0, b))
    with pytest.raises(Exception, match="HDF5 error while reading instance"):
        qcm.read_instance_from_hdf5(str(tmp_path / "absent.h5"), "cluster_0", b)
    with pytest.raises(Exception, match="HDF5 error while reading instance"):
        qcm.read_instance_from_hdf5(path, "cluster_7", b)
    with pytest.raises(Exception, match="The label 987654 is out of range."):
        qcm.write_instance_to_hdf5(path, "cluster_0", 987654)
    ave = qcm.cluster_averages(b)  # from the Green function average: no interaction average
    assert np.isnan(ave["U"][0]) and not np.isnan(ave["mu"][0])
    with pytest.raises(Exception, match="ground-state vectors"):
        qcm.susceptibility("mu", [1j], b)
