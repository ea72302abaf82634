"""Direction network: the FP64 oracle's gradient against central finite
differences (CPU), and the GPU's forward pass, loss, gradient and training
against the oracle (GPU)."""
import numpy as np
import pytest

import dirnet_oracle as DO


def _net(rng, L, H, O):
    n = H * L + H + H * H + H + O * H + O
    return rng.standard_normal(n) * 0.3


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    L, H, O, B = 3, 6, 4, 5
    p = _net(rng, L, H, O)
    Z = rng.standard_normal((B, L))
    A = rng.standard_normal((O, O))
    C = A @ A.T / O
    _, g = DO.loss_grad(p, Z, C, L, H, O)
    h = 1e-6
    num = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (DO.loss(p + e, Z, C, L, H, O) - DO.loss(p - e, Z, C, L, H, O)) / (2 * h)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)


def test_oracle_map_is_odd():
    rng = np.random.default_rng(1)
    L, H, O = 3, 5, 2
    p = _net(rng, L, H, O)
    Z = rng.standard_normal((7, L))
    V, _ = DO.forward(p, Z, L, H, O)
    Vm, _ = DO.forward(p, -Z, L, H, O)
    np.testing.assert_allclose(V, -Vm, atol=1e-12)


@pytest.mark.gpu
def test_gpu_forward_loss_and_gradient_match_oracle():
    import achr
    rng = np.random.default_rng(2)
    L, H, O, B = 16, 64, 12, 96
    print(f"dirnet L={L} H={H} O={O} B={B}", flush=True)
    net = achr.DirNet(L, H, O, seed=3)
    p = net.params().astype(np.float64)
    A = rng.standard_normal((O, O))
    C = (A @ A.T / O).astype(np.float32)
    net.set_target(C)
    Z = rng.standard_normal((B, L)).astype(np.float32)
    V = net.sample(Z, prec=0)
    Vo, _ = DO.forward(p, Z.astype(np.float64), L, H, O)
    assert np.linalg.norm(V - Vo) / np.linalg.norm(Vo) < 1e-5
    Vb = net.sample(Z, prec=2)
    assert np.linalg.norm(Vb - Vo) / np.linalg.norm(Vo) < 3e-2
    np.testing.assert_array_equal(net.sample(-Z, prec=0), -V)  # odd, row by row
    loss, g = net.loss_grad(Z)
    lo, go = DO.loss_grad(p, Z.astype(np.float64), C.astype(np.float64), L, H, O)
    assert abs(loss - lo) / lo < 1e-4
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4


@pytest.mark.gpu
def test_gpu_training_fits_a_target_covariance():
    import achr
    rng = np.random.default_rng(4)
    O = 24
    lam = np.exp(rng.uniform(-3, 3, O))
    Q, _ = np.linalg.qr(rng.standard_normal((O, O)))
    C = ((Q * lam) @ Q.T).astype(np.float32)
    net = achr.DirNet(O, 128, O, seed=5)
    net.set_target(C)
    print("dirnet training", flush=True)
    losses = net.train(300, 1024, 3e-3, seed=1)
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    Z = rng.standard_normal((20000, O)).astype(np.float32)
    V = net.sample(Z, prec=2).astype(np.float64)
    emp = V.T @ V / len(V)
    assert np.linalg.norm(emp - C) / np.linalg.norm(C) < 0.15


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle_at_width_512():
    """The network of the learned workload (hidden width 512, a rank-512
    principal subspace in and out): GPU forward, loss and gradient against
    the FP64 oracle, whose gradient the finite-difference test pins."""
    import achr
    rng = np.random.default_rng(6)
    L, H, O, B = 512, 512, 512, 256
    print(f"dirnet L={L} H={H} O={O} B={B}", flush=True)
    net = achr.DirNet(L, H, O, seed=9)
    p = net.params().astype(np.float64)
    A = rng.standard_normal((O, O))
    C = (A @ A.T / O).astype(np.float32)
    net.set_target(C)
    Z = rng.standard_normal((B, L)).astype(np.float32)
    V = net.sample(Z, prec=0)
    Vo, _ = DO.forward(p, Z.astype(np.float64), L, H, O)
    assert np.linalg.norm(V - Vo) / np.linalg.norm(Vo) < 1e-5
    np.testing.assert_array_equal(net.sample(-Z, prec=0), -V)
    loss, g = net.loss_grad(Z)
    lo, go = DO.loss_grad(p, Z.astype(np.float64), C.astype(np.float64), L, H, O)
    print(f"loss {loss} vs {lo}; grad rel err {np.linalg.norm(g - go) / np.linalg.norm(go):.3g}", flush=True)
    assert abs(loss - lo) / lo < 1e-4
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4
