"""Cost model on the GPU: forward and backward kernels against the float64
CPU reference, training convergence, and learned dispatch end to end."""
import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    ctx = ab.Context(ix)
    return k, v, ix, ctx


def _rand_model(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=oracle.mlp_nparams()).astype(np.float32)
    mu = rng.normal(size=12).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, size=12).astype(np.float32)
    return p, mu, sd


def test_forward_matches_reference(small):
    _, _, _, ctx = small
    p, mu, sd = _rand_model(1)
    ctx.set_model(p, mu, sd)
    x = np.random.default_rng(2).normal(size=(300, 12)).astype(np.float32)
    y, sel = ctx.predict(x)
    xs = (x.astype(np.float64) - mu) / sd
    want = oracle.mlp_forward(p.astype(np.float64), xs)
    np.testing.assert_allclose(y, want, rtol=1e-4, atol=1e-4)
    assert np.array_equal(sel == ab.AB_PATH_DENSE, y[:, 0] < y[:, 1])
    ctx.clear_model()


@pytest.mark.parametrize("n", [1, 37, 256, 1000])
def test_gradient_matches_reference(small, n):
    _, _, _, ctx = small
    p, mu, sd = _rand_model(3 + n)
    ctx.set_model(p, mu, sd)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 12)).astype(np.float32)
    t = rng.normal(size=(n, 2)).astype(np.float32)
    loss, g = ctx.model_grad(x, t)
    xs = (x.astype(np.float64) - mu) / sd
    wl, wg = oracle.mlp_loss_grad(p.astype(np.float64), xs, t.astype(np.float64))
    assert loss == pytest.approx(wl, rel=1e-4)
    # float32 accumulation over n samples: relative to the gradient's scale
    np.testing.assert_allclose(g, wg, rtol=2e-3, atol=2e-5 * max(1.0, np.abs(wg).max()))
    ctx.clear_model()


def test_training_converges(small):
    _, _, _, ctx = small
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4000, 12)).astype(np.float32)
    w = rng.normal(size=(12, 2))
    t = (np.tanh(x @ w) + 0.5 * x[:, :2] ** 2).astype(np.float32)
    ctx.clear_model()
    curve = ctx.train_model(x, t, epochs=40, batch=256, lr=3e-3, seed=3)
    assert np.isfinite(curve).all()
    assert curve[-1] < 0.1 * curve[0], curve
    y, _ = ctx.predict(x[:500])
    assert np.mean((y - t[:500]) ** 2) < 0.05
    ctx.clear_model()


def test_learned_dispatch_end_to_end(small):
    k, v, ix, ctx = small
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    qk, qo = ab.gen_queries(0xADAB0401, configs.C1.nkeys, 24, size_dist=1, p1=80.0, p2=1.2, lo=1, hi=2000,
                            key_dist=1, zipf_s=1.0, perm_seed=configs.C1.seed)
    # a model whose dense output is the matched-record feature and sparse a
    # constant: dispatch must follow the sign and results stay bit-exact
    p = np.zeros(oracle.mlp_nparams(), np.float32)
    p[0 * 12 + 2] = 1.0               # h1[0] = relu(f2 standardized)
    p[64 * 12 + 64 + 0] = 1.0          # h2[0] = relu(h1[0])
    o3 = 64 * 12 + 64 + 64 * 64 + 64
    p[o3 + 0] = -1.0                   # y_dense = -h2[0] + b
    p[o3 + 64 * 2] = 0.0
    p[o3 + 64 * 2 + 1] = -8.0          # y_sparse = -8: dense iff log1p(M) > 8
    ctx.set_model(p, np.zeros(12, np.float32), np.ones(12, np.float32), slot=ab.MODEL_BATCH)
    try:
        bms, ids, counts, paths = ctx.query_batch(qk, qo, ab.AB_PATH_LEARNED)
    finally:
        ctx.clear_model()
    f = np.stack([ix.features_host(qk[qo[i]:qo[i + 1]])[0] for i in range(24)])
    want_dense = -np.maximum(f[:, 2], 0) < -8.0
    assert np.array_equal(paths == ab.AB_PATH_DENSE, want_dense)
    assert want_dense.any() and (~want_dense).any()
    for i in range(24):
        wb, wi = csr.union(qk[qo[i]:qo[i + 1]])
        assert np.array_equal(bms[i], wb) and np.array_equal(ids[i], wi)
