"""GPU parity of the tile-cost MLP kernels (chai_hip_mlp_*) against the CPU
oracle (oracle/mlp.c, double precision from the same f32 parameters).

Tolerances: the GPU evaluates in FP32 with FMA; each layer is a 64-term dot
product, so |y_gpu - y_cpu| <= 1e-4 * (1 + |y|) and gradients agree to
||g_gpu - g_cpu|| <= 1e-4 ||g_cpu|| (plus 1e-4 * max|g| per element).  The SGD
step is bit-exact (p - lr * g in f32, no contraction)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def data(F, B, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, F)).astype(np.float32)
    Y = (np.sin(X[:, 0]) + 0.5 * X[:, 1 % F] ** 2 - X[:, 2 % F] * X[:, 3 % F]).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("F,B", [(16, 1000), (5, 37), (32, 300), (1, 1)])
def test_forward_matches_oracle(chai, oracle, F, B):
    from chai_amd.tuner import MLP
    m = MLP(F, seed=F)
    X, _ = data(F, B, 1)
    y = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    ref = oracle.mlp_forward(m.params(), X)
    assert np.all(np.abs(y - ref) <= 1e-4 * (1 + np.abs(ref))), float(np.abs(y - ref).max())


@pytest.mark.parametrize("F,B", [(16, 1000), (7, 64), (32, 3), (16, 5000)])
def test_backward_matches_oracle(chai, oracle, F, B):
    from chai_amd.tuner import MLP
    m = MLP(F, seed=100 + F)
    X, Y = data(F, B, 2)
    g = m.backward(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()).cpu().numpy().astype(np.float64)
    loss, ref = oracle.mlp_gradient(m.params(), X, Y)
    err = np.abs(g - ref)
    assert np.linalg.norm(g - ref) <= 1e-4 * np.linalg.norm(ref), (np.linalg.norm(g - ref), np.linalg.norm(ref))
    assert np.all(err <= 1e-4 * np.abs(ref).max() + 1e-7)
    # the gradient call does not move the parameters
    np.testing.assert_array_equal(m.params(), m.params())


def test_backward_is_deterministic(chai):
    from chai_amd.tuner import MLP
    m = MLP(16, seed=3)
    X, Y = data(16, 4000, 3)
    Xd, Yd = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    g1 = m.backward(Xd, Yd).cpu()
    g2 = m.backward(Xd, Yd).cpu()
    assert torch.equal(g1, g2)


def test_sgd_update_is_exact(chai):
    from chai_amd.tuner import MLP
    m = MLP(16, seed=4)
    X, Y = data(16, 200, 4)
    p0 = m.params()
    g = m.backward(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda())
    m.update(g, 0.05, adam=False)
    torch.cuda.synchronize()
    expect = p0 - np.float32(0.05) * g.cpu().numpy()
    np.testing.assert_array_equal(m.params(), expect)


def test_training_reduces_loss(chai, oracle):
    from chai_amd.tuner import MLP
    m = MLP(8, seed=5)
    X, Y = data(8, 4096, 5)
    Xd, Yd = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    l0, _ = oracle.mlp_gradient(m.params(), X, Y)
    m.fit(Xd, Yd, epochs=60, batch=256, lr=3e-3)
    l1, _ = oracle.mlp_gradient(m.params(), X, Y)
    print("loss", l0, "->", l1)
    assert l1 < 0.1 * l0


def test_predict_variant_records_choice(chai):
    from chai_amd import tuner
    m = tuner.MLP(tuner.NFEAT, seed=6)
    shape = (1024, 512, 256, 0, 0, 4)
    v, pred, secs = tuner.predict(m, shape)
    assert 0 <= v < 8 and pred.shape == (8,) and v == int(np.argmin(pred))
    X = torch.from_numpy(np.stack([tuner.features(shape, k) for k in range(8)])).cuda()
    np.testing.assert_allclose(m.forward(X).cpu().numpy(), pred, rtol=0, atol=0)


def test_small_learned_tuner_end_to_end(chai):
    """Tiny version of the learned-autotune workload: time every variant on 60
    shapes (M, N, K <= 2048), train, predict 20 unseen shapes; the model's
    regret must beat a uniformly random variant choice (the full-size
    comparison with the exhaustive search runs in bench.py)."""
    from chai_amd import tuner
    ws = tuner.Workspace(hi=2048)
    train = tuner.random_shapes(60, seed=11, hi=2048)
    test = tuner.random_shapes(20, seed=12, hi=2048)
    t_train = np.array([tuner.time_variants(ws, s, trials=2)[0] for s in train])
    t_test = np.array([tuner.time_variants(ws, s, trials=2)[0] for s in test])
    assert np.all(t_train > 0) and np.all(t_test > 0)
    X, Y = tuner.dataset(train, t_train)
    mu, sd = Y.mean(), Y.std()
    m = tuner.MLP(tuner.NFEAT, seed=0)
    m.fit(torch.from_numpy(X).cuda(), torch.from_numpy((Y - mu) / sd).cuda(), epochs=150, batch=64, lr=3e-3)
    choice = [tuner.predict(m, s)[0] for s in test]
    r_model = tuner.regret(choice, t_test).mean()
    fixed = int(np.argmin(np.log(t_train).sum(0)))
    r_fixed = tuner.regret([fixed] * len(test), t_test).mean()
    r_random = np.mean([tuner.regret([v] * len(test), t_test).mean() for v in range(8)])
    print("regret model %.3f fixed-variant %.3f random %.3f" % (r_model, r_fixed, r_random))
    assert r_model < r_random
