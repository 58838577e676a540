"""Batch modes of ab_query_batch against the CPU oracle: host dispatch (0),
device dispatch with one persistent union launch (1), and the same replayed
from a cached hipGraph (2).  Graph replays are checked one by one, with other
device work and synchronizations between them, and after every setting that
changes launch parameters."""
import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu

MODES = (ab.Context.BATCH_HOST, ab.Context.BATCH_FUSED, ab.Context.BATCH_GRAPH)


@pytest.fixture(scope="module")
def c1():
    import torch

    assert torch.cuda.is_available()
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    return ix, csr


def _mixed(seed, nq, nkeys):
    return ab.gen_queries(seed, nkeys, nq, size_dist=1, p1=60.0, p2=1.3, lo=0, hi=3000, key_dist=1, zipf_s=1.0,
                          perm_seed=configs.C1.seed)


def _verify(csr, qk, qo, res, what):
    bms, ids, counts, paths = res
    for i in range(qo.size - 1):
        want_bm, want_ids = csr.union(qk[qo[i]:qo[i + 1]])
        assert np.array_equal(bms[i], want_bm), f"{what} q{i} path {paths[i]}: bitmap differs"
        assert counts[i] == want_ids.size, f"{what} q{i}: count"
        assert np.array_equal(ids[i], want_ids), f"{what} q{i}: ids differ"


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("path", [ab.AB_PATH_THRESHOLD, ab.AB_PATH_DENSE, ab.AB_PATH_SPARSE])
def test_modes_match_oracle(c1, mode, path):
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(mode)
    ctx.set_batch_threshold(0.01)
    qk, qo = _mixed(0xADAB0201, 48, configs.C1.nkeys)
    print(f"mode={mode} path={path}", flush=True)
    res = ctx.query_batch(qk, qo, path)
    if path == ab.AB_PATH_THRESHOLD:
        assert (res[3] == ab.AB_PATH_DENSE).any() and (res[3] == ab.AB_PATH_SPARSE).any(), res[3]
    else:
        assert (res[3] == path).all()
    _verify(csr, qk, qo, res, f"mode {mode} path {path}")
    ctx.close()


def test_graph_replays_with_eager_work_between(c1):
    """Same batch shape (query count and key total) with new keys each time:
    the cached graph is replayed and must read the new keys and offsets."""
    import torch

    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(ab.Context.BATCH_GRAPH)
    ctx.set_batch_threshold(0.01)
    qk0, qo0 = _mixed(0xADAB0202, 32, configs.C1.nkeys)
    rng = np.random.default_rng(5)
    x = torch.arange(1 << 20, device="cuda", dtype=torch.float32)
    for rep in range(5):
        # same offsets, permuted query order and fresh key values of the same total
        qk = qk0.copy()
        if rep:
            qk = rng.integers(0, configs.C1.nkeys, size=qk0.size, dtype=np.uint64).astype(np.uint32)
            if rep % 2:
                qk[: qk.size // 2] = qk0[: qk.size // 2]
        print(f"replay {rep}", flush=True)
        res = ctx.query_batch(qk, qo0, ab.AB_PATH_THRESHOLD)
        _verify(csr, qk, qo0, res, f"replay {rep}")
        # unrelated device work and host synchronization between replays
        y = (x * (rep + 1)).sum()
        torch.cuda.synchronize()
        assert float(y) > 0
    ctx.close()


def test_graph_recapture_on_changes(c1):
    """Shape, setting and model changes each give a new graph; results stay exact."""
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(ab.Context.BATCH_GRAPH)
    qk, qo = _mixed(0xADAB0203, 40, configs.C1.nkeys)
    steps = [
        ("base", lambda: None),
        ("tau", lambda: ctx.set_batch_threshold(0.002)),
        ("group 7", lambda: ctx.set_dense_group(7)),
        ("group 33", lambda: ctx.set_dense_group(33)),
        ("group 128", lambda: ctx.set_dense_group(128)),
        ("threads 256", lambda: ctx.set_dense_threads(256)),
        ("no columns", lambda: ctx.set_dense_columns(False)),
        ("chunk 256", lambda: ctx.set_sparse_chunk(256)),
        ("threads 128 (host dispatch)", lambda: ctx.set_dense_threads(128)),
    ]
    for name, act in steps:
        act()
        # a shape runs as launches the first time and from a graph when it
        # comes again; a smaller batch in between changes the shape
        for n in (40, 40, 40, 17, 17, 40):
            sub_qo = qo[: n + 1]
            sub_qk = qk[: int(sub_qo[-1])]
            print(f"recapture {name} n={n}", flush=True)
            res = ctx.query_batch(sub_qk, sub_qo, ab.AB_PATH_THRESHOLD)
            _verify(csr, sub_qk, sub_qo, res, f"{name} n={n}")
    ctx.close()


def test_graph_reuse_across_key_counts(c1):
    """One graph serves batches of the same query count whose key counts round
    to the same capacity, with keys at a different device address each time:
    the keys are staged and the key count is read on the device."""
    import torch

    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(ab.Context.BATCH_GRAPH)
    ctx.set_batch_threshold(0.01)
    nq = 24
    nv = configs.C1.nvalues
    ids = torch.full((nq * nv,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.full((nq,), -1, dtype=torch.int32, device="cuda")
    rng = np.random.default_rng(11)
    for rep in range(8):
        sizes = rng.integers(1, 300, nq)
        if rep == 5:
            sizes[:] = 0  # no keys at all
        if rep == 6:
            sizes[0] = 5000  # a larger key count: another capacity
        qo = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
        qk = rng.integers(0, configs.C1.nkeys + 10, int(qo[-1])).astype(np.uint32)
        pad = int(rng.integers(0, 64))  # the keys start at a new address
        dk = torch.from_numpy(np.concatenate([np.zeros(pad, np.uint32), qk]).view(np.int32)).cuda()
        print(f"reuse rep={rep} K={int(qo[-1])}", flush=True)
        ctx.query_batch_device(dk[pad:] if qk.size else dk, qo, ab.AB_PATH_THRESHOLD, None, ids, nv, cnt,
                               want_paths=False)
        ctx.sync()
        h = ids.cpu().numpy().view(np.uint32)
        c = cnt.cpu().numpy().view(np.uint32)
        for i in range(nq):
            want = csr.union(qk[qo[i]:qo[i + 1]])[1]
            assert c[i] == want.size and np.array_equal(h[i * nv:i * nv + want.size], want), f"rep {rep} q{i}"
    ctx.close()


def test_device_buffers_and_timing(c1):
    """query_batch_device into caller buffers with per-kernel timing on (mode 2
    then runs as mode 1 and reports the fused union class, and the sparse
    route launch that precedes it, with no per-group dense launches)."""
    import torch

    ix, csr = c1
    ctx = ab.Context(ix)
    qk, qo = _mixed(0xADAB0204, 24, configs.C1.nkeys)
    nq = qo.size - 1
    sw = ix.stride_words
    dk = torch.from_numpy(qk.view(np.int32)).cuda()
    bm = torch.empty(nq * sw, dtype=torch.int32, device="cuda")
    for timing in (False, True):
        ctx.set_timing(timing)
        bm.fill_(-1)
        torch.cuda.synchronize()
        ctx.query_batch_device(dk, qo, ab.AB_PATH_THRESHOLD, bm)
        ctx.sync()
        got = bm.cpu().numpy().view(np.uint32).reshape(nq, sw)
        for i in range(nq):
            want_bm, _ = csr.union(qk[qo[i]:qo[i + 1]])
            assert np.array_equal(got[i], want_bm), f"timing={timing} q{i}"
        if timing:
            t = ctx.timing(reset=True)
            assert t["fused_union"][1] >= 1 and t["fused_union"][0] > 0, t
            assert t["dense_tiles"][1] == 0 and t["sparse_union"][1] == 1, t
    ctx.set_timing(False)
    ctx.close()


@pytest.mark.parametrize("ts", [8, 13])
def test_fused_edge_filter(ts):
    from .test_gpu_parity import _edge_filter

    k, v, nvalues = _edge_filter()
    ix = ab.Index(k, v, 12, nvalues, ts)
    csr = oracle.CSR(k, v, 12, nvalues)
    rng = np.random.default_rng(ts + 1)
    qs = [rng.choice(12, size=rng.integers(0, 8), replace=False).astype(np.uint32) for _ in range(70)]
    qs[0] = np.arange(12, dtype=np.uint32)
    qk = np.concatenate(qs)
    qo = np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64)
    for mode in (ab.Context.BATCH_FUSED, ab.Context.BATCH_GRAPH):
        ctx = ab.Context(ix)
        ctx.set_batch_mode(mode)
        for group in (1, 2, 32, 64, 100, 128, 256):
            ctx.set_dense_group(group)
            for path in (ab.AB_PATH_DENSE, ab.AB_PATH_SPARSE):
                print(f"edge ts={ts} mode={mode} group={group} path={path}", flush=True)
                res = ctx.query_batch(qk, qo, path)
                _verify(csr, qk, qo, res, f"ts={ts} mode={mode} group={group} path={path}")
        ctx.close()


def test_host_dispatch_caps_group_width(c1):
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(ab.Context.BATCH_HOST)
    ctx.set_dense_group(128)
    qk, qo = _mixed(0xADAB0205, 150, configs.C1.nkeys)
    res = ctx.query_batch(qk, qo, ab.AB_PATH_DENSE)
    _verify(csr, qk, qo, res, "host dispatch, group 128")
    ctx.close()


def test_autotune_winner_matches_oracle(c1):
    """ab_ctx_autotune on a small C1 batch in graph mode: every candidate of
    the grid reproduces the entry settings' bitmaps, the context is left on
    the winner, and the winner's results (left in place, and applied to a
    fresh context) equal the oracle's."""
    import torch

    ix, csr = c1
    qk, qo = _mixed(0xADAB0207, 40, configs.C1.nkeys)
    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    ctx = ab.Context(ix)
    ctx.set_batch_mode(ab.Context.BATCH_GRAPH)
    ctx.set_batch_threshold(0.01)
    print("autotune on 40 C1 queries", flush=True)
    res = ctx.autotune(d_keys, qo, ab.AB_PATH_THRESHOLD, repeats=1)
    assert len(res) == 9 * 2 * 2 * 2
    assert all(r["mismatched_words"] == 0 for r in res), [r for r in res if r["mismatched_words"]][:4]
    assert all(r["ms"] > 0 for r in res) and [r["ms"] for r in res] == sorted(r["ms"] for r in res)
    best = res[0]
    out = ctx.query_batch(qk, qo, ab.AB_PATH_THRESHOLD)
    assert (out[3] == ab.AB_PATH_DENSE).any() and (out[3] == ab.AB_PATH_SPARSE).any(), out[3]
    _verify(csr, qk, qo, out, f"autotuned {best}")
    ctx2 = ab.Context(ix)
    ctx2.set_batch_mode(ab.Context.BATCH_GRAPH)
    ctx2.set_batch_threshold(0.01)
    ctx2.apply_tuning(best)
    _verify(csr, qk, qo, ctx2.query_batch(qk, qo, ab.AB_PATH_THRESHOLD), f"applied {best}")
    ctx2.close()
    ctx.close()


def test_batch_tuning_table(c1):
    """ab_ctx_set_batch_tuning: each batch applies the entry for its size (a
    batch below the first entry keeps the settings in force), results equal
    the oracle's at every switch in both device modes, bad tables are refused,
    and the autotuner's candidates are not overridden by the table."""
    ix, csr = c1
    small = {"min_batch": 2, "group": 8, "dense_threads": 256, "sparse_chunk": 256, "dense_columns": False}
    large = {"min_batch": 40, "group": 128, "dense_threads": 512, "sparse_chunk": 8192, "dense_columns": True}
    for mode in (ab.Context.BATCH_FUSED, ab.Context.BATCH_GRAPH):
        ctx = ab.Context(ix)
        ctx.set_batch_mode(mode)
        ctx.set_batch_threshold(0.01)
        ctx.set_batch_tuning([small, large])
        for nq, want in ((12, small), (60, large), (1, large), (39, small), (150, large)):
            qk, qo = _mixed(0xADAB0208 + nq, nq, configs.C1.nkeys)
            print(f"tuning table mode={mode} nq={nq}", flush=True)
            res = ctx.query_batch(qk, qo, ab.AB_PATH_THRESHOLD)
            got = ctx.batch_settings()
            assert got == {k: v for k, v in want.items() if k != "min_batch"}, (nq, got)
            _verify(csr, qk, qo, res, f"table mode {mode} nq {nq}")
        ctx.close()
    ctx = ab.Context(ix)
    with pytest.raises(ab.AdaBoolError):
        ctx.set_batch_tuning([large, small])  # not ascending
    with pytest.raises(ab.AdaBoolError):
        ctx.set_batch_tuning([dict(small, dense_threads=100)])
    ctx.set_batch_tuning([dict(large, min_batch=1)])
    qk, qo = _mixed(0xADAB0209, 24, configs.C1.nkeys)
    import torch

    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    res = ctx.autotune(d_keys, qo, ab.AB_PATH_THRESHOLD, repeats=1)
    # each candidate ran with its own group width (with the table in force
    # every one would have recorded the entry's)
    assert len({r["group"] for r in res}) > 1, res
    ctx.set_batch_tuning([])
    ctx.close()
