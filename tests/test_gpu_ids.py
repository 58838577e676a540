"""Id-only batches (ab_query_batch_ids: no result bitmap) against the CPU oracle.

* C1 mix: ids and counts equal the oracle's sort-based union (and its bitmap
  walk), at aligned and unaligned offsets, with nothing written outside a
  query's range, and equal the bitmap path's ids;
* every record type at window and tile edges;
* small expansion budgets: the batch runs in many parts, and a query larger
  than the budget runs as a part of its own;
* values next to 2^32 - 1 (nvalues = 2^32 - 1), where a bitmap would take
  512 MB per query, on host- and device-built indices;
* the serving loop with ids_only: alone == in mix == oracle, and replay
  counts equal the oracle's;
* both methods (by sort, by value slice) and the automatic choice; a filter
  of 1.06M values (five value slices, the last partial) with a heavy key over
  the whole value range, at aligned and unaligned offsets, in one part and in
  many."""
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
    empty, out-of-range, repeated-key and all-key queries."""
    qk, qo = ab.gen_queries(seed, configs.C1.nkeys, nq, size_dist=2, p1=1.5, lo=1, hi=3000, key_dist=1, zipf_s=1.0,
                            perm_seed=configs.C1.seed)
    qs = [qk[qo[i]:qo[i + 1]] for i in range(nq)]
    qs[1] = np.zeros(0, np.uint32)
    qs[2] = np.array([configs.C1.nkeys + 5, 2**32 - 1], np.uint32)
    qs[3] = np.concatenate([qs[0], qs[0], qs[0][:1]])
    qs[4] = np.arange(configs.C1.nkeys, dtype=np.uint32)
    return qs


def _flat(qs):
    return (np.concatenate(qs) if qs else np.zeros(0, np.uint32)), \
        np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64)


def _run_ids(ctx, qs, aligned=True, room=None):
    """Runs one id-only batch; every query gets its id bound (or `room`) plus a
    sentinel gap.  Returns (ids per query, counts, host buffer, offsets, end)."""
    import torch

    qk, qo = _flat(qs)
    off, pos = [], 0 if aligned else 1
    for q in qs:
        off.append(pos)
        n = ctx.index.id_bound(q) if room is None else room
        pos += (n + 3) // 4 * 4 + 4 if aligned else n + 3
    dk = torch.from_numpy(qk.view(np.int32)).cuda() if qk.size else torch.zeros(1, dtype=torch.int32, device="cuda")
    ids = torch.full((pos + 8,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.full((max(len(qs), 1),), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    ctx.query_batch_ids_device(dk, qo, ids, off, cnt)
    ctx.sync()
    h = ids.cpu().numpy().view(np.uint32)
    c = cnt.cpu().numpy().view(np.uint32)[: len(qs)]
    return [h[off[i]: off[i] + int(c[i])] for i in range(len(qs))], c, h, off, pos


def _check_ranges(h, off, pos, counts, what):
    for i in range(len(off)):
        end = off[i + 1] if i + 1 < len(off) else pos
        assert (h[off[i] + int(counts[i]):end] == 0xFFFFFFFF).all(), f"{what} q{i} wrote past its range"


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("aligned", [True, False])
def test_ids_c1_mix(c1, aligned, method):
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_id_method(method)
    qs = _mix(0xADAB0601, 60)
    print(f"ids c1 mix aligned={aligned} method={method}", flush=True)
    got, counts, h, off, pos = _run_ids(ctx, qs, aligned)
    for i, q in enumerate(qs):
        want = csr.union_ids(q)
        assert counts[i] == want.size, f"q{i}"
        assert np.array_equal(got[i], want), f"q{i} ids"
        assert np.array_equal(want, csr.union(q)[1])
    _check_ranges(h, off, pos, counts, f"aligned={aligned}")
    ctx.close()


def test_ids_equal_bitmap_path(c1):
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_batch_threshold(0.01)
    qs = _mix(0xADAB0602, 40)
    qk, qo = _flat(qs)
    print("ids vs bitmap path", flush=True)
    _, bm_ids, bm_counts, paths = ctx.query_batch(qk, qo, ab.AB_PATH_THRESHOLD)
    assert (paths == ab.AB_PATH_DENSE).any() and (paths == ab.AB_PATH_SPARSE).any()
    ids, counts = ctx.query_batch_ids(qk, qo)
    assert np.array_equal(counts, bm_counts)
    for i in range(len(qs)):
        assert np.array_equal(ids[i], bm_ids[i]), f"q{i}"
    # the same context keeps working on the bitmap path afterwards
    _, again, _, _ = ctx.query_batch(qk, qo, ab.AB_PATH_THRESHOLD)
    for i in range(len(qs)):
        assert np.array_equal(again[i], bm_ids[i])
    ctx.close()


def _edge_filter():
    nvalues = 3 * 8192 + 77
    keys, vals = [], []
    spans = [[0], [255, 256], [256, 257, 300, 511], list(range(4088, 4096)), list(range(8192, 8206)),
             list(range(512, 542)), list(range(1024, 1224)), list(range(0, 256, 2)) + [nvalues - 1],
             [nvalues - 1, nvalues - 2, 8191, 8192], list(range(16384, 16640)),
             [5, 5, 5, 9000]]  # a repeated pair
    for key, vs in enumerate(spans):
        keys += [key] * len(vs)
        vals += vs
    return np.array(keys, np.uint32), np.array(vals, np.uint32), nvalues


@pytest.mark.parametrize("ts", [8, 13])
def test_ids_edge_filter(ts):
    k, v, nvalues = _edge_filter()
    ix = ab.Index(k, v, 12, nvalues, ts)
    ctx = ab.Context(ix)
    csr = oracle.CSR(k, v, 12, nvalues)
    qs = [np.array([i], np.uint32) for i in range(12)] + [np.arange(12, dtype=np.uint32),
                                                           np.array([6, 9, 11, 6], np.uint32),
                                                           np.zeros(0, np.uint32)]
    print(f"ids edge filter ts={ts}", flush=True)
    got, counts, h, off, pos = _run_ids(ctx, qs)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i], csr.union(q)[1]), f"ts={ts} q{i}"
    _check_ranges(h, off, pos, counts, f"ts={ts}")
    ctx.close()


@pytest.mark.parametrize("budget", [4096, 50000])
def test_ids_budget_parts(c1, budget):
    ix, csr = c1
    ctx = ab.Context(ix)
    ctx.set_id_budget(budget)
    qs = _mix(0xADAB0603, 80)  # q4 (every key, ~1M values) exceeds either budget
    print(f"ids budget={budget}", flush=True)
    got, counts, h, off, pos = _run_ids(ctx, qs)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i], csr.union_ids(q)), f"budget={budget} q{i}"
    _check_ranges(h, off, pos, counts, f"budget={budget}")
    with pytest.raises(ab.AdaBoolError):
        ctx.set_id_budget(100)
    with pytest.raises(ab.AdaBoolError):
        ctx.set_id_method(3)
    ctx.close()


@pytest.mark.parametrize("device_build", [False, True])
def test_ids_top_of_u32(device_build):
    import torch

    nvalues = 2**32 - 1
    top = nvalues - 1
    rng = np.random.default_rng(0xADAB0604)
    nk, n = 300, 60000
    k = rng.integers(0, nk, n).astype(np.uint32)
    v = (top - rng.integers(0, 200000, n)).astype(np.uint32)
    v[: n // 20] = rng.integers(0, 5000, n // 20)  # low ids too
    k[:300] = 7  # a heavy key -> bitsets
    v[:300] = top - np.arange(300)
    v[300], v[301] = top, 255
    if device_build:
        ix = ab.Index.build_device(torch.from_numpy(k.view(np.int32)).cuda(), torch.from_numpy(v.view(np.int32)).cuda(),
                                   nk, nvalues)
    else:
        ix = ab.Index(k, v, nk, nvalues)
    assert ix.info["nvalues"] == nvalues
    csr = oracle.CSR(k, v, nk, nvalues)
    ctx = ab.Context(ix)
    qs = [rng.choice(nk, size=int(rng.integers(1, 40)), replace=False).astype(np.uint32) for _ in range(50)]
    qs[0] = np.array([7], np.uint32)
    qs[1] = np.arange(nk, dtype=np.uint32)
    print(f"ids top of u32 device_build={device_build}", flush=True)
    got, counts, h, off, pos = _run_ids(ctx, qs)
    for i, q in enumerate(qs):
        assert np.array_equal(got[i], csr.union_ids(q)), f"q{i}"
    assert got[0][-1] == top
    _check_ranges(h, off, pos, counts, "top")
    ctx.close()
    ix.close()


def test_server_ids_only_alone_equals_in_mix(c1):
    ix, csr = c1
    qs = _mix(0xADAB0605, 200)
    srv = ab.Server(ix, max_batch=64, max_wait_us=100, inflight=2, deliver=ab.Server.DELIVER_HOST,
                    arena_bytes=64 << 20, ids_only=1)
    try:
        print("server ids_only mix", flush=True)
        tickets = [srv.submit(q) for q in qs]
        got = {}
        while len(got) < len(qs):
            for r in srv.poll(cap=128, wait_ms=100):
                assert r["status"] == 0 and r["path"] == ab.AB_PATH_IDS
                got[r["ticket"]] = r["ids"]
                srv.release(r["ticket"])
        mix = [got[t] for t in tickets]
        assert srv.stats()["batches"] >= 2
        for i in (0, 1, 2, 3, 4, 50, 199):
            t = srv.submit(qs[i])
            rs = []
            while not rs:
                rs = srv.poll(cap=4, wait_ms=200)
            assert len(rs) == 1 and rs[0]["ticket"] == t
            assert np.array_equal(rs[0]["ids"], mix[i]), f"q{i}: alone != in mix"
            srv.release(t)
        for i, q in enumerate(qs):
            assert np.array_equal(mix[i], csr.union_ids(q)), f"q{i} differs from the oracle"
        qk, qo = _flat(qs)
        counts, lat, wall = srv.replay(qk, qo, rate_qps=20000.0, seed=5)
        assert np.array_equal(counts.astype(np.int64), np.array([m.size for m in mix]))
        assert srv.stats()["arena_used_bytes"] == 0
    finally:
        srv.close()


@pytest.fixture(scope="module")
def wide():
    """1.06M values (four value slices of 2^18 and a partial fifth) over 20k
    keys, ~2.4M pairs, Zipf-weighted keys; a heavy key spans every slice."""
    nvalues, nkeys = 4 * (1 << 18) + 12345, 20000
    k, v = ab.gen_pairs(0xADAB0607, nvalues, nkeys, lo=1, hi=3, zipf_s=1.0)
    hv = np.arange(0, nvalues, 4, dtype=np.uint32)
    k = np.concatenate([k, np.full(hv.size, 17, np.uint32)])
    v = np.concatenate([v, hv])
    ix = ab.Index(k, v, nkeys, nvalues)
    csr = oracle.CSR(k, v, nkeys, nvalues)
    qk, qo = ab.gen_queries(0xADAB0608, nkeys, 120, size_dist=2, p1=1.5, lo=1, hi=6000, key_dist=1, zipf_s=1.0)
    qs = [qk[qo[i]:qo[i + 1]] for i in range(120)]
    qs[0] = np.array([17], np.uint32)
    qs[1] = np.zeros(0, np.uint32)
    qs[2] = np.arange(nkeys, dtype=np.uint32)
    qs[3] = np.array([nkeys + 1, 17, 17, 3], np.uint32)
    return ix, csr, qs


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("budget", [0, 60000])
@pytest.mark.parametrize("aligned", [True, False])
def test_ids_wide_filter(wide, aligned, budget, method):
    """Both methods (by sort, by value slice) on five value slices, the last
    partial, in one part and in many."""
    ix, csr, qs = wide
    ctx = ab.Context(ix)
    ctx.set_id_method(method)
    if budget:
        ctx.set_id_budget(budget)
    print(f"ids wide filter aligned={aligned} budget={budget} method={method}", flush=True)
    ids, counts, h, off, pos = _run_ids(ctx, qs, aligned)
    _check_ranges(h, off, pos, counts, f"aligned={aligned} budget={budget}")
    for i, q in enumerate(qs):
        want = csr.union_ids(q)
        assert counts[i] == want.size, f"q{i}"
        assert np.array_equal(ids[i], want), f"q{i} ids"
    assert counts[2] > 1 << 20  # the all-key query spans every slice
    ctx.close()
