"""What the trained cost model gains over the threshold rule, on held-out
timed queries it never saw (tests/golden/cost_heldout.npz), and that the
GPU inference kernel takes the same decisions as a float64 forward pass."""
import json
import os

import numpy as np
import pytest

import adabool_amd as ab
from adabool_amd import costmodel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json")
HELD = os.path.join(ROOT, "tests", "golden", "cost_heldout.npz")


def _forward(m, x):
    """float64 forward of a committed model (12 -> 64 -> 64 -> 2, ReLU)."""
    p = np.asarray(m["params"], np.float64)
    F, H1, H2, O = 12, 64, 64, 2
    o = 0
    W1 = p[o:o + H1 * F].reshape(H1, F); o += H1 * F
    b1 = p[o:o + H1]; o += H1
    W2 = p[o:o + H2 * H1].reshape(H2, H1); o += H2 * H1
    b2 = p[o:o + H2]; o += H2
    W3 = p[o:o + O * H2].reshape(O, H2); o += O * H2
    b3 = p[o:o + O]
    xs = (x.astype(np.float64) - np.asarray(m["mean"])) / np.asarray(m["std"])
    h1 = np.maximum(xs @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


@pytest.fixture(scope="module")
def data():
    return json.load(open(MODEL)), dict(np.load(HELD))


def test_model_file_shape(data):
    model, _ = data
    for name in ("single", "batch"):
        m = model[name]
        assert len(m["params"]) == 5122 and len(m["mean"]) == 12 and len(m["std"]) == 12
        assert all(s > 0 for s in m["std"])
        assert m["meta"]["tau"] >= 0
    assert model["meta"]["train_queries"] + model["meta"]["heldout_queries"] >= 200_000


def test_learned_beats_threshold_single_query(data):
    """Single-query dispatch: the network's regret vs the oracle-best path is
    below the threshold rule's (same timings, tau fit on the training set)."""
    model, d = data
    t = d["t_single_ms"].astype(np.float64)
    learned = _forward(model["single"], d["features"])
    choice = learned[:, 0] < learned[:, 1]
    tau = model["single"]["meta"]["tau"]
    thr = d["M"].astype(np.float64) >= tau * d["R"].astype(np.float64)
    r_learned = costmodel.regret(choice, t)
    r_thr = costmodel.regret(thr, t)
    assert r_learned < r_thr, (r_learned, r_thr)
    # and both are far better than always taking one path
    assert r_learned < costmodel.regret(np.zeros_like(choice), t)
    assert r_learned < costmodel.regret(np.ones_like(choice), t)


def test_learned_beats_threshold_batch(data):
    """Batch dispatch, on held-out batch-share labels (an index the network
    never saw, early-stopped on validation regret): the network's regret is
    at most the threshold rule's (tau fit on the same training timings) and
    within 1% of the per-query best, and both rules beat sending every query
    one way (every query dense costs about a third more, every query sparse
    many times more)."""
    model, d = data
    t = d["t_batch_ms"].astype(np.float64)
    learned = _forward(model["batch"], d["features"])
    choice = learned[:, 0] < learned[:, 1]
    tau = model["batch"]["meta"]["tau"]
    thr = d["M"].astype(np.float64) >= tau * d["R"].astype(np.float64)
    r_learned = costmodel.regret(choice, t)
    r_thr = costmodel.regret(thr, t)
    all_dense = costmodel.regret(np.ones_like(choice), t)
    all_sparse = costmodel.regret(np.zeros_like(choice), t)
    assert r_learned <= r_thr, (r_learned, r_thr)
    assert r_learned < 0.01 and r_learned < all_dense / 10 and r_learned < all_sparse / 10, (r_learned, all_dense)
    assert r_thr < all_dense and r_thr < all_sparse, (r_thr, all_dense, all_sparse)
    # the committed model's own record says the same of the whole held-out index
    meta = model["batch"]["meta"]
    assert meta["heldout_regret_learned"] <= meta["heldout_regret_threshold"]
    assert meta["validation_regret_learned"] <= meta["validation_regret_threshold"]


def test_regret_definition():
    t = np.array([[1.0, 2.0], [3.0, 1.0]])
    assert costmodel.regret(np.array([True, False]), t) == 0.0
    assert costmodel.regret(np.array([False, True]), t) == pytest.approx((2 + 3) / 2 - 1)
    M = np.array([1, 10, 100], np.uint64)
    R = np.full(3, 1000, np.uint64)
    tt = np.array([[5.0, 1.0], [5.0, 6.0], [5.0, 50.0]])
    tau = costmodel.fit_threshold(M, R, tt)
    assert 0.001 < tau <= 0.01  # dense from M = 10 on


@pytest.mark.gpu
def test_gpu_inference_matches_float64(data):
    model, d = data
    k, v = ab.gen_pairs(3, 5000, 500, lo=5, hi=15)
    ctx = ab.Context(ab.Index(k, v))
    ctx.load_model(MODEL)
    for slot, name in ((ab.MODEL_SINGLE, "single"), (ab.MODEL_BATCH, "batch")):
        y, sel = ctx.predict(d["features"], slot=slot)
        want = _forward(model[name], d["features"])
        np.testing.assert_allclose(y, want, rtol=1e-3, atol=1e-3)
        margin = np.abs(want[:, 0] - want[:, 1]) > 1e-3
        assert np.array_equal((sel == ab.AB_PATH_DENSE)[margin], (want[:, 0] < want[:, 1])[margin])
