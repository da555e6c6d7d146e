"""Cluster models shared by the tests, built with the oracle's model description."""
import numpy as np

from oracle import oracle as O

PLAQ_SEQ = {1: [1, 1, 1, 1], 2: [1, 1, 1, 1], 3: [1, 1, -1, -1], 4: [1, 1, -1, -1],
            5: [1, -1, 1, -1], 6: [1, -1, 1, -1], 7: [1, -1, -1, 1], 8: [1, -1, -1, 1]}


def chain(n, t=1.0, U=4.0, mu=2.0, tp=0.0):
    bonds = [(i, i + 1) for i in range(n - 1)]
    tpb = [(i, i + 2) for i in range(n - 2)] if tp else ()
    return O.hubbard_cluster(n, bonds, t=t, U=U, mu=mu, tp_bonds=tpb, tp=tp)


def square(nx, ny, t=1.0, tp=0.0, U=8.0, mu=4.0):
    nn, nnn = O.square_bonds(nx, ny)
    return O.hubbard_cluster(nx * ny, nn, t=t, U=U, mu=mu, tp_bonds=nnn if tp else (), tp=tp)


def plaquette_bath(params, n_bath=8):
    """2x2 plaquette + bath orbitals, operators as in the reference's
    tests/test_diagonalization/cluster_2x2_2C_8b_C2v_L.py (U, mu on the 4 sites;
    eb_x bath energies; tb_x site-bath hybridizations with C2v sign patterns)."""
    L = 4 + n_bath
    m = O.ClusterModel(L)
    m.add("U", "interaction", [(i, i + L, 1.0) for i in range(4)])
    m.add("mu", "one-body", [(i, i, -1.0) for i in range(4)] + [(i + L, i + L, -1.0) for i in range(4)])
    vals = {"U": params["U"], "mu": params["mu"]}
    for x in range(1, n_bath + 1):
        b = 3 + x
        seq = PLAQ_SEQ[(x - 1) % 8 + 1]
        m.add("eb%d" % x, "one-body", [(b, b, 1.0), (b + L, b + L, 1.0)])
        m.add("tb%d" % x, "one-body", [(i, b, seq[i]) for i in range(4)] + [(i + L, b + L, seq[i]) for i in range(4)])
        vals["eb%d" % x] = params.get("eb%d_1" % x, params.get("eb%d" % x))
        vals["tb%d" % x] = params.get("tb%d_1" % x, params.get("tb%d" % x))
    return m, vals


def random_bath_params(rng, n_bath=8):
    U = rng.uniform(0, 12)
    p = {"U": U, "mu": rng.uniform(0, U)}
    for x in range(1, n_bath + 1):
        p["eb%d" % x] = rng.uniform(-3, 3)
        p["tb%d" % x] = rng.uniform(0.1, 1.0)
    return p


def ops_list(model):
    return [(name, op.kind, op.elements) for name, op in model.operators.items()]
