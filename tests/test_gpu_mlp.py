"""Warm-start MLP on the GPU: forward and gradient against the oracle, BF16 matrix-core
inference against FP32, Adam against the oracle's update, training, and the warm start's
gain on a CDMFT problem."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _data(rng, n, ni, no):
    x = rng.normal(size=(n, ni)).astype(np.float32)
    A = rng.normal(size=(ni, no))
    t = (np.sin(x @ A) + 0.1 * (x @ A) ** 2).astype(np.float32)
    return x, t


def test_forward_and_gradient_match_oracle(gpu_lib):
    from pyqcm_amd.warmstart import MLP
    rng = np.random.default_rng(0)
    net = MLP(3, 128, 16, seed=7)
    assert net.n_params == 19088
    p = net.get_params().astype(np.float64)
    x, t = _data(rng, 37, 3, 16)  # not a multiple of the 16-sample tile
    y = net.forward_raw(x)
    yr, _ = MO.forward(p, x.astype(np.float64), 3, 128, 16)
    np.testing.assert_allclose(y, yr, rtol=1e-5, atol=1e-5)
    loss, g = net.loss_grad_raw(x, t)
    lr_, gr = MO.loss_grad(p, x.astype(np.float64), t.astype(np.float64), 3, 128, 16)
    assert abs(loss - lr_) <= 1e-5 * abs(lr_)
    assert np.linalg.norm(g - gr) <= 1e-5 * np.linalg.norm(gr)


def test_bf16_matrix_core_inference_close_to_fp32(gpu_lib):
    from pyqcm_amd.warmstart import MLP
    rng = np.random.default_rng(1)
    net = MLP(3, 128, 16, seed=3)
    x = rng.normal(size=(100, 3)).astype(np.float32)
    y32 = net.forward_raw(x)
    y16 = net.forward_raw(x, bf16=True)
    err = np.abs(y16 - y32).max() / np.abs(y32).max()
    print("bf16 vs fp32 relative max error %.3g" % err)
    assert err < 2e-2
    # and it is not the FP32 path in disguise
    assert np.abs(y16 - y32).max() > 0


def test_one_adam_step_matches_oracle(gpu_lib):
    from pyqcm_amd.warmstart import MLP
    rng = np.random.default_rng(2)
    net = MLP(3, 128, 16, seed=5)
    p0 = net.get_params().astype(np.float64)
    x, t = _data(rng, 64, 3, 16)
    _, g = net.loss_grad_raw(x, t)
    net.train_raw(x, t, epochs=1, batch=64, lr=1e-3)
    p1 = net.get_params().astype(np.float64)
    pr, _, _ = MO.adam_step(p0, np.zeros_like(p0), np.zeros_like(p0), g.astype(np.float64), 1, 1e-3)
    np.testing.assert_allclose(p1, pr, atol=2e-6)


def test_training_reduces_loss(gpu_lib):
    from pyqcm_amd.warmstart import MLP
    rng = np.random.default_rng(3)
    net = MLP(3, 128, 16, seed=1)
    x, t = _data(rng, 512, 3, 16)
    l0, _ = net.loss_grad_raw(x, t)
    hist = net.train_raw(x, t, epochs=200, batch=64, lr=2e-3)
    l1, _ = net.loss_grad_raw(x, t)
    print("loss %.4g -> %.4g" % (l0, l1))
    assert l1 < 0.05 * l0 and hist[-1] < hist[0]


def test_warm_start_reduces_cdmft_iterations(gpu_lib):
    """Trained on fixed points of the plain loop (2-site + 4-bath chain, U and mu varied),
    the network's predictions start held-out CDMFT runs closer to their fixed points:
    fewer outer iterations, same converged bath."""
    from pyqcm_amd import cdmft as C
    from pyqcm_amd import lattice as LT
    from pyqcm_amd.warmstart import MLP
    M = LT.LatticeModel("L2_4b_ws", LT.chain_tiling(2), n_bath=4)
    M.add_cluster_operator("tb1", [(1, 3, -1.0), (2, 4, -1.0), (7, 9, -1.0), (8, 10, -1.0)])
    M.add_cluster_operator("tb2", [(1, 5, -1.0), (2, 6, -1.0), (7, 11, -1.0), (8, 12, -1.0)])
    M.add_cluster_operator("eb1", [(3, 3, 1.0), (4, 4, 1.0), (9, 9, 1.0), (10, 10, 1.0)])
    M.add_cluster_operator("eb2", [(5, 5, 1.0), (6, 6, 1.0), (11, 11, 1.0), (12, 12, 1.0)])
    P = C.CDMFTProblem(M, ["eb1", "eb2", "tb1", "tb2"], "R0:N6:S0",
                       grid=C.frequency_grid("matsubara", (2, 50)))
    rng = np.random.default_rng(4)
    cold = np.array([1.0, -1.0, 0.5, 0.5])

    def points(n):
        U = rng.uniform(2.0, 6.0, n)
        return [{"t": 1.0, "U": u, "mu": u / 2} for u in U]

    train = points(40)
    B = C.BatchedCDMFT(P, train, [cold] * len(train))
    ok = B.converged & ~B.failed
    X = np.array([[v["U"], v["mu"], 0.0] for v in train])[ok]
    Y = np.zeros((ok.sum(), 16))
    Y[:, :4] = np.array(B.x)[ok]
    net = MLP(3, 128, 16, seed=0)
    net.fit(X, Y, epochs=600, batch=16, lr=2e-3)
    test = points(6)
    Xt = np.array([[v["U"], v["mu"], 0.0] for v in test])
    warm = net.predict(Xt)[:, :4]
    Bc = C.BatchedCDMFT(P, test, [cold] * len(test))
    Bw = C.BatchedCDMFT(P, test, list(warm))
    print("cold iterations", Bc.niter, "warm iterations", Bw.niter)
    assert Bc.converged.all() and Bw.converged.all()
    np.testing.assert_allclose(np.array(Bw.x), np.array(Bc.x), atol=2e-3)
    assert Bw.n_solves.sum() < Bc.n_solves.sum()
