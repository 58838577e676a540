"""Serving loop and packed batches against the CPU oracle.

* packed ids: ab_query_batch_packed writes each query's ids at its own offset
  (aligned and unaligned offsets, every batch mode);
* alone == in mix: a request's ids are the same whether it is served on its
  own or inside a stream of others, and equal the oracle's;
* admission: a small result arena makes requests wait for memory, and one
  that can never fit fails with AB_OOM while the rest complete;
* open-loop Poisson replay: every count equals the oracle's."""
import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    import torch

    assert torch.cuda.is_available()
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    return ix, csr


def _mix(seed, nq):
    """C3-like sizes (power law 1.5 on [1, 3000]) with Zipf-weighted keys, plus
    a few empty and out-of-range queries."""
    qk, qo = ab.gen_queries(seed, configs.C1.nkeys, nq, size_dist=2, p1=1.5, lo=1, hi=3000, key_dist=1, zipf_s=1.0,
                            perm_seed=configs.C1.seed)
    qs = [qk[qo[i]:qo[i + 1]] for i in range(nq)]
    qs[1] = np.zeros(0, np.uint32)
    qs[2] = np.array([configs.C1.nkeys + 5, 2**32 - 1], np.uint32)
    return qs


def _flat(qs):
    return (np.concatenate(qs) if qs else np.zeros(0, np.uint32)), \
        np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("aligned", [True, False])
def test_packed_batch(c1, mode, aligned):
    import torch

    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_mode(mode)
    ctx.set_batch_threshold(0.01)
    qs = _mix(0xADAB0501, 40)
    qk, qo = _flat(qs)
    want = [csr.union(q)[1] for q in qs]
    # offsets: each query gets exactly its count (+ a gap), unaligned if asked
    off, pos = [], 0 if aligned else 1
    for w in want:
        off.append(pos)
        pos += w.size + (4 - w.size % 4) % 4 + 4 if aligned else w.size + 3
    dk = torch.from_numpy(qk.view(np.int32)).cuda()
    ids = torch.full((pos + 8,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.empty(len(qs), dtype=torch.int32, device="cuda")
    print(f"packed mode={mode} aligned={aligned}", flush=True)
    ctx.query_batch_packed(dk, qo, ab.AB_PATH_THRESHOLD, ids, off, None, cnt)
    ctx.sync()
    h = ids.cpu().numpy().view(np.uint32)
    c = cnt.cpu().numpy().view(np.uint32)
    for i, w in enumerate(want):
        assert c[i] == w.size, f"q{i}"
        assert np.array_equal(h[off[i]:off[i] + w.size], w), f"q{i} ids"
        # nothing written past the query's ids
        end = off[i + 1] if i + 1 < len(off) else pos
        assert (h[off[i] + w.size:end] == 0xFFFFFFFF).all(), f"q{i} wrote past its range"
    ctx.close()


def test_batch_beyond_grid_limit(c1):
    """70000 queries in one packed batch: more than a grid's y dimension holds
    (65535), so the per-query launches (zero-fill, group masks, popcount, id
    writer) go in slices of queries, device-dispatched in graph mode and
    host-dispatched in mode 0."""
    import torch

    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_threshold(0.01)
    rng = np.random.default_rng(0xADAB0503)
    nq = 70000
    qs = [rng.choice(configs.C1.nkeys, size=int(rng.integers(1, 3)), replace=False).astype(np.uint32)
          for _ in range(nq)]
    for i in range(0, nq, 5000):  # some large queries, which go dense
        qs[i] = rng.choice(configs.C1.nkeys, size=300, replace=False).astype(np.uint32)
    qs[nq - 1] = np.zeros(0, np.uint32)
    qk, qo = _flat(qs)
    want = [csr.union_ids(q) for q in qs]
    counts = np.array([w.size for w in want], np.int64)
    off = np.concatenate([[0], np.cumsum((counts + 3) // 4 * 4 + 4)[:-1]]).astype(np.uint64)
    pos = int(off[-1] + (counts[-1] + 3) // 4 * 4 + 4)
    dk = torch.from_numpy(qk.view(np.int32)).cuda()
    ids = torch.full((pos + 8,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.empty(nq, dtype=torch.int32, device="cuda")
    for mode in (2, 0):
        ctx.set_batch_mode(mode)
        ids.fill_(-1)
        cnt.fill_(-1)
        print(f"packed batch of {nq}, mode {mode}", flush=True)
        paths = ctx.query_batch_packed(dk, qo, ab.AB_PATH_THRESHOLD, ids, off, None, cnt)
        ctx.sync()
        assert (paths == ab.AB_PATH_DENSE).any() and (paths == ab.AB_PATH_SPARSE).any()
        c = cnt.cpu().numpy().view(np.uint32)
        assert np.array_equal(c.astype(np.int64), counts)
        h = ids.cpu().numpy().view(np.uint32)
        for i in list(range(0, nq, 97)) + list(range(0, nq, 5000)) + [65534, 65535, 65536, nq - 2, nq - 1]:
            a, n = int(off[i]), int(counts[i])
            assert np.array_equal(h[a:a + n], want[i]), f"mode {mode} q{i} ids"
            assert (h[a + n:a + n + 4] == 0xFFFFFFFF).all(), f"mode {mode} q{i} wrote past its range"
    ctx.close()


def test_alone_equals_in_mix(c1):
    ix, csr = c1
    qs = _mix(0xADAB0502, 300)
    srv = ab.Server(ix, max_batch=64, max_wait_us=100, inflight=2, deliver=ab.Server.DELIVER_HOST,
                    arena_bytes=64 << 20)
    try:
        for ctx in srv.contexts:
            ctx.set_batch_threshold(0.01)
        # in mix: everything submitted at once, served in continuous batches
        tickets = [srv.submit(q) for q in qs]
        got = {}
        while len(got) < len(qs):
            for r in srv.poll(cap=128, wait_ms=100):
                assert r["status"] == 0
                got[r["ticket"]] = r
                srv.release(r["ticket"])
        mix_ids = [got[t]["ids"] for t in tickets]
        st = srv.stats()
        assert st["batches"] >= 2 and st["mean_batch"] > 1, st
        # alone: one request at a time
        for i in (0, 1, 2, 3, 7, 50, 120, 299):
            t = srv.submit(qs[i])
            rs = []
            while not rs:
                rs = srv.poll(cap=4, wait_ms=200)
            assert len(rs) == 1 and rs[0]["ticket"] == t
            alone = rs[0]["ids"]
            srv.release(t)
            assert np.array_equal(alone, mix_ids[i]), f"q{i}: alone != in mix"
        for i, q in enumerate(qs):
            assert np.array_equal(mix_ids[i], csr.union(q)[1]), f"q{i} differs from the oracle"
    finally:
        srv.close()


def test_admission_waits_for_memory(c1):
    ix, csr = c1
    qs = _mix(0xADAB0503, 120)
    nv = configs.C1.nvalues
    # arena of 1.5 full results: large requests must wait for earlier releases
    srv = ab.Server(ix, max_batch=16, inflight=2, deliver=ab.Server.DELIVER_HOST, arena_bytes=int(nv * 4 * 1.5))
    try:
        big = np.arange(configs.C1.nkeys, dtype=np.uint32)  # every key: upper bound nvalues, fits
        order = qs[:60] + [big] + qs[60:]
        tickets = [srv.submit(q) for q in order]
        got = {}
        while len(got) < len(order):
            for r in srv.poll(cap=64, wait_ms=100):
                assert r["status"] == 0, r
                got[r["ticket"]] = r["ids"]
                srv.release(r["ticket"])
        for t, q in zip(tickets, order):
            assert np.array_equal(got[t], csr.union(q)[1])
        assert srv.stats()["arena_used_bytes"] == 0
    finally:
        srv.close()
    # an arena smaller than one request's upper bound: that request fails, others pass
    srv = ab.Server(ix, max_batch=8, inflight=1, deliver=ab.Server.DELIVER_HOST, arena_bytes=4096 * 4)
    try:
        small = np.array([0], np.uint32)
        t_small = srv.submit(small)
        t_big = srv.submit(np.arange(configs.C1.nkeys, dtype=np.uint32))
        t_small2 = srv.submit(small)
        res = {}
        while len(res) < 3:
            for r in srv.poll(cap=8, wait_ms=100):
                res[r["ticket"]] = r
                if r["status"] == 0:
                    srv.release(r["ticket"])
        assert res[t_big]["status"] == 2  # AB_OOM
        for t in (t_small, t_small2):
            assert res[t]["status"] == 0 and np.array_equal(res[t]["ids"], csr.union(small)[1])
    finally:
        srv.close()


@pytest.mark.parametrize("rate", [0.0, 20000.0])
def test_replay_counts(c1, rate):
    ix, csr = c1
    qs = _mix(0xADAB0504, 400)
    qk, qo = _flat(qs)
    srv = ab.Server(ix, max_batch=128, max_wait_us=200, inflight=2, arena_bytes=256 << 20)
    try:
        counts, lat, wall = srv.replay(qk, qo, rate_qps=rate, seed=3)
        want = np.array([csr.union(q)[1].size for q in qs])
        assert np.array_equal(counts.astype(np.int64), want)
        assert (lat > 0).all() and wall > 0
        st = srv.stats()
        assert st["completed"] == len(qs) and st["lat_p99_ms"] >= st["lat_p50_ms"] > 0
        assert st["arena_used_bytes"] == 0
    finally:
        srv.close()
