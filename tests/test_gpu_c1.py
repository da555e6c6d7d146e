"""The plain ED + CPT workload's cluster: the 3x4 Hubbard cluster (t=1, t'=0, U=8, mu=4) on
its N=12, Sz=0 sector (D = 924^2 = 853,776). Its ground state and its band-Lanczos Green
function with 12 start vectors per excitation kind (N=11 and N=13 sectors, D = 731,808
each) against the oracle's restatement of the same steps (src_ed/Hamiltonian/Lanczos.hpp:
41-143, 214-395) from the oracle's own ground state."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tests import models as TM

pytestmark = pytest.mark.gpu

_names = itertools.count(1)


def test_green_function_3x4_matches_oracle(gpu_lib):
    from pyqcm_amd import qcm
    m, vals = TM.square(3, 4, tp=0.0, U=8.0, mu=4.0)
    name = "c1_3x4_%d" % next(_names)
    qcm.new_model(name, 12, 0)
    for opname, op in m.operators.items():
        qcm.new_operator(name, opname, op.kind, [(r + 1, c + 1, v) for (r, c, v) in op.elements])
    lab = 7700 + next(_names)
    qcm.new_model_instanceC(name, vals, "R0:N12:S0", lab)
    E0g, sec = qcm.ground_state_solve(lab)
    So = O.Sector(m, vals, 6, 6)
    E0o, gs, *_ = So.lanczos_gs()
    print("E0 device %.12f oracle %.12f" % (E0g, E0o))
    assert abs(E0g - E0o) <= 1e-10 * abs(E0o)
    zs = 1j * np.array([0.1, 0.5, 2.0, 6.0])
    es, Qs = [], []
    for pm in (-1, 1):
        To = O.Sector(m, vals, 6 + pm, 6)
        phi = [O.create_destroy(12, 6, 6, a, pm, gs) for a in range(12)]
        e, Q = O.band_lanczos(lambda x: To.hv_omp(x, threads=16), phi)
        e = e - E0o
        if pm == -1:
            e = -e
        e, Q = O.streamline(e, Q)
        es.append(e)
        Qs.append(Q)
    e = np.concatenate(es)
    Q = np.hstack(Qs)
    G = qcm.Green_function(zs, 0, lab)
    for i, z in enumerate(zs):
        ref = O.qmatrix_gf(e, Q, z)
        err = np.abs(G[i] - ref).max()
        print("Im z %g: max |G_gpu - G_oracle| = %.2e (|G| <= %.3f; poles gpu %d, oracle %d)"
              % (z.imag, err, np.abs(ref).max(), len(qcm.qmatrix(lab)[0]), len(e)))
        # the two recurrences may stop a block apart (a round-off-level decision): 5e-8
        assert err <= 5e-8 * max(1.0, np.abs(ref).max())
