"""Writes tests/golden/c2_lanczos.json: the oracle's Lanczos ground state of the 4x4
Hubbard cluster (t=1, t'=-0.3, U=8, mu=4) on its N=16, Sz=0 sector (D = 165,636,900),
from the reference start vector (libstdc++ default_random_engine(0) + normal_distribution,
src_ed/Hamiltonian/Lanczos.cpp:7), accuracy 1e-12 and the reference's stopping test
(src_ed/Hamiltonian/Lanczos.hpp:41-95). Products use the oracle's OpenMP-over-rows form.

Run in the build container (about nine minutes on 8 cores, 5 GB of memory):
    python tests/golden/make_c2_lanczos.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    nn, nnn = O.square_bonds(4, 4)
    m, vals = O.hubbard_cluster(16, nn, tp_bonds=nnn, tp=-0.3, U=8.0, mu=4.0)
    S = O.Sector(m, vals, 8, 8)
    assert S.dim == 165636900
    t0 = time.time()
    O.lib().or_reset_normal()
    gs = S.start_vector(0)
    accur, max_iter = 1e-12, 600
    E0 = np.zeros(1)
    nit = np.zeros(1, np.int32)
    a = np.zeros(max_iter + 2)
    b = np.zeros(max_iter + 2)
    rc = O.lib().or_lanczos_gs_threads(S.h, accur, max_iter, O._dp(gs), O._dp(E0), O._ip(nit), O._dp(a),
                                       O._dp(b), os.cpu_count() or 1, 0)
    if rc != 0:
        raise RuntimeError(O.lib().or_last_error().decode())
    n = int(nit[0])
    out = {
        "source": "oracle Lanczos (tests/golden/make_c2_lanczos.py), 4x4 t=1 t'=-0.3 U=8 mu=4, N=16 Sz=0",
        "dim": int(S.dim), "seed": 0, "accur": accur, "E0": float(E0[0]), "iterations": n,
        "alpha": [float(v) for v in a[:min(n, 20)]], "beta": [float(v) for v in b[:min(n, 20)]],
        "seconds": round(time.time() - t0, 1),
    }
    with open(os.path.join(ROOT, "tests", "golden", "c2_lanczos.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
