"""Failure paths of the solver, forced from the host through the reference's global
parameters (every kernel still runs to its normal end): the Lanczos iteration limit
(src_ed/Hamiltonian/Lanczos.hpp:92), the Q-matrix unitarity check
(src_ed/Q_matrix.hpp:235-260) and the band-Lanczos degeneracy guard no_degenerate_BL
(src_ed/Hamiltonian/Lanczos.hpp:368-370), in the sequential and the batched solvers."""
import contextlib
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu

_counter = itertools.count(5000)


def register(model, values, sectors, n_sites):
    from pyqcm_amd import qcm
    name = "err%d" % next(_counter)
    qcm.new_model(name, n_sites, model.L - n_sites)
    for opname, op in model.operators.items():
        qcm.new_operator(name, opname, op.kind, [(r + 1, c + 1, v) for (r, c, v) in op.elements])
    label = next(_counter)
    qcm.new_model_instanceC(name, values, sectors, label)
    return label


@contextlib.contextmanager
def global_value(name, value, default):
    from pyqcm_amd import qcm
    qcm.set_global_parameter(name, value)
    try:
        yield
    finally:
        qcm.set_global_parameter(name, default)


def two_chains():
    """Two identical, uncoupled 4-site chains: every one-particle excitation is doubly
    degenerate, so the band Lanczos of the Green function meets a degenerate lowest pair."""
    m = O.ClusterModel(8)
    L = 8
    bonds = [(0, 1), (1, 2), (2, 3), (4, 5), (5, 6), (6, 7)]
    m.add("t", "one-body", [(a, b, -1.0) for a, b in bonds] + [(a + L, b + L, -1.0) for a, b in bonds])
    m.add("U", "interaction", [(i, i + L, 1.0) for i in range(8)])
    m.add("mu", "one-body", [(i, i, -1.0) for i in range(8)] + [(i + L, i + L, -1.0) for i in range(8)])
    return m, {"t": 1.0, "U": 4.0, "mu": 2.0}


def test_lanczos_iteration_limit(gpu_lib):
    from pyqcm_amd import qcm
    from pyqcm_amd.sector import Sector
    m, vals = TM.square(3, 3, tp=-0.3)
    S = Sector(m.L, 5, 4, m.flat(vals))
    with pytest.raises(RuntimeError, match="Lanczos procedure exceeded 3 iterations"):
        S.lanczos_gs(max_iter=3)
    lab = register(m, vals, "R0:N9:S1", 9)
    with global_value("max_iter_lanczos", 3, 1000):
        with pytest.raises(RuntimeError, match="Lanczos procedure exceeded 3 iterations"):
            qcm.ground_state_solve(lab)


def test_qmatrix_unitarity_check(gpu_lib):
    """A threshold below the rounding of the sum rule makes the check fail with the
    reference's message; the default threshold passes."""
    from pyqcm_amd import qcm
    m, vals = TM.square(3, 3, tp=-0.3)
    lab = register(m, vals, "R0:N9:S1", 9)
    qcm.ground_state_solve(lab)
    with global_value("accur_Q_matrix", 1e-300, 1e-5):
        with pytest.raises(RuntimeError, match="Q-matrix does not satisfy unitary condition by"):
            qcm.Green_function(np.array([0.5j]), 0, lab)
    lab2 = register(m, vals, "R0:N9:S1", 9)
    G = qcm.Green_function(np.array([0.5j]), 0, lab2)
    assert np.all(np.isfinite(G))


@pytest.mark.parametrize("batched", [False, True])
def test_band_lanczos_degeneracy_guard(gpu_lib, batched):
    from pyqcm_amd import qcm
    m, vals = two_chains()
    with global_value("max_dim_full", 0, 256):
        lab = register(m, vals, "R0:N8:S0", 8)
        # without the guard the degenerate spectrum is accepted
        if batched:
            qcm.batch_solve([lab])
        G = qcm.Green_function(np.array([0.5j]), 0, lab)
        assert np.all(np.isfinite(G))
        with global_value("no_degenerate_BL", True, False):
            lab2 = register(m, vals, "R0:N8:S0", 8)
            with pytest.raises(RuntimeError,
                               match="band Lanczos: the gap between the first two eigenvalues is smaller than"):
                if batched:
                    qcm.batch_solve([lab2])
                else:
                    qcm.ground_state_solve(lab2)
                    qcm.Green_function(np.array([0.5j]), 0, lab2)
