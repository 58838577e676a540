"""Full-size workloads against the oracle, every query checked.

* C2 at the timed configurations: the 4M x 1M filter device-built at value
  tile 2^10 with 512-thread work-groups and at 2^9 with 256, 256-query dense
  groups with column masks, learned and threshold dispatch, graph mode; three
  calls of the same 1024-query batch (the first runs as launches, the second
  is captured, the third replays).
* C3 on the C2 filter: power-law batches of 1, 64 and 4096 queries in graph
  mode, with both paths firing.
* 65,536 C3 queries in ONE device-dispatched batch (above a grid's y limit of
  65535), on the C2 distributions at 1/16 of the size so that every id fits
  in memory twice.
* C5 served: the 16M x 4M filter device-built and served to C3-mix requests:
  alone == in mix == oracle.

Each query is compared by count, bitmap digest and id digest (oracle.h
defines them; the oracle computes them from its own union, the checker from
the GPU's bitmaps and ids), so results of tens of GB need no second copy."""
import os

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json")
NT = min(16, os.cpu_count() or 1)
CHUNK_BYTES = 1 << 31  # host copies of at most 2 GiB at a time


def _digest_bitmaps(d_bm, nq, sw):
    out = []
    rows = max(1, CHUNK_BYTES // (sw * 4))
    for q0 in range(0, nq, rows):
        q1 = min(nq, q0 + rows)
        h = d_bm[q0 * sw:q1 * sw].cpu().numpy().view(np.uint32).reshape(q1 - q0, sw)
        out.append(oracle.bitmap_digest(h, sw, NT))
    return np.concatenate(out) if out else np.zeros(0, np.uint64)


def _digest_ids(d_ids, off, counts):
    """Digests of every query's ids in a packed device buffer, in host chunks."""
    nq = off.size
    dig = np.zeros(nq, np.uint64)
    asc = np.zeros(nq, bool)
    q0 = 0
    while q0 < nq:
        q1 = q0 + 1
        while q1 < nq and int(off[q1] + counts[q1] - off[q0]) * 4 <= CHUNK_BYTES:
            q1 += 1
        a, b = int(off[q0]), int(off[q1 - 1] + counts[q1 - 1])
        h = d_ids[a:b].cpu().numpy().view(np.uint32) if b > a else np.zeros(0, np.uint32)
        d, s = oracle.ids_digest(h, off[q0:q1] - a, counts[q0:q1], NT)
        dig[q0:q1], asc[q0:q1] = d, s
        q0 = q1
    return dig, asc


def _packed_offsets(counts):
    room = (counts.astype(np.uint64) + 3) // 4 * 4 + 4
    off = np.concatenate([[0], np.cumsum(room)[:-1]]).astype(np.uint64)
    return off, int(off[-1] + room[-1]) if counts.size else 0


def _run_checked(ctx, qk, qo, path, want, calls=3, ids_on=(2,), bitmaps_on=(0, 1, 2), what=""):
    """Runs the same batch `calls` times into sentinel-filled outputs; checks
    counts on every call, bitmaps and ids on the calls named; returns paths."""
    import torch

    wc, wb, wi = want
    nq = qo.size - 1
    sw = ctx.index.stride_words
    off, total = _packed_offsets(wc)
    dk = torch.from_numpy(np.ascontiguousarray(qk).view(np.int32)).cuda()
    d_bm = torch.empty(max(nq * sw, 1), dtype=torch.int32, device="cuda")
    d_ids = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
    d_cnt = torch.empty(max(nq, 1), dtype=torch.int32, device="cuda")
    paths = None
    for call in range(calls):
        d_bm.fill_(-1)
        d_ids.fill_(-1)
        d_cnt.fill_(-1)
        torch.cuda.synchronize()
        print(f"[workload] {what}: call {call} of {nq} queries", flush=True)
        paths = ctx.query_batch_packed(dk, qo, path, d_ids, off, d_bm, d_cnt)
        ctx.sync()
        c = d_cnt[:nq].cpu().numpy().view(np.uint32).astype(np.uint64)
        bad = np.flatnonzero(c != wc)
        assert bad.size == 0, f"{what} call {call}: counts differ at queries {bad[:8]}"
        if call in bitmaps_on:
            bd = _digest_bitmaps(d_bm, nq, sw)
            bad = np.flatnonzero(bd != wb)
            assert bad.size == 0, f"{what} call {call}: bitmaps differ at queries {bad[:8]}"
        if call in ids_on:
            idg, asc = _digest_ids(d_ids, off, wc)
            assert asc.all(), f"{what} call {call}: ids not ascending at {np.flatnonzero(~asc)[:8]}"
            bad = np.flatnonzero(idg != wi)
            assert bad.size == 0, f"{what} call {call}: ids differ at queries {bad[:8]}"
    return paths


@pytest.fixture(scope="module")
def c2():
    import torch

    assert torch.cuda.is_available()
    k, v = configs.make_filter(configs.C2)
    csr = oracle.CSR(k, v, configs.C2.nkeys, configs.C2.nvalues)
    dk = torch.from_numpy(k.view(np.int32)).cuda()
    dv = torch.from_numpy(v.view(np.int32)).cuda()
    ix = ab.Index.build_device(dk, dv, configs.C2.nkeys, configs.C2.nvalues, 9)
    yield ix, csr, (dk, dv)
    ix.close()
    del dk, dv
    torch.cuda.empty_cache()


def _ctx(ix, group, threads, columns=True):
    ctx = ab.Context(ix)
    if os.path.exists(MODEL):
        d = ctx.load_model(MODEL)
        ctx.set_threshold(d["single"]["meta"]["tau"])
        ctx.set_batch_threshold(d["batch"]["meta"]["tau"])
    ctx.set_dense_group(group)
    ctx.set_dense_threads(threads)
    ctx.set_dense_columns(columns)
    ctx.set_batch_mode(2)
    return ctx


@pytest.mark.parametrize("tile,threads", [(10, 512), (9, 256)])
def test_c2_at_the_timed_configuration(c2, tile, threads):
    """The batch the benchmark times, with the settings its autotuner picks
    (tile 2^10, 512 threads, 256-query groups; 2^9 and 256 threads before):
    every query's count, bitmap and ids equal the oracle's on launch, capture
    and replay."""
    ix, csr, (dk, dv) = c2
    qk, qo = configs.make_queries(configs.C2, configs.C2_Q)
    want = csr.union_batch_digest(qk, qo, NT)
    tix = ix if tile == 9 else ab.Index.build_device(dk, dv, configs.C2.nkeys, configs.C2.nvalues, tile)
    assert tix.info["tile_shift"] == tile
    ctx = _ctx(tix, 256, threads)
    for path in (ab.AB_PATH_AUTO, ab.AB_PATH_THRESHOLD):
        _run_checked(ctx, qk, qo, path, want, calls=3, ids_on=(0, 2), bitmaps_on=(0, 1, 2),
                     what=f"C2 tile {tile} threads {threads} path {path}")
    ctx.close()
    if tix is not ix:
        tix.close()


@pytest.mark.parametrize("batch", [1, 64, 4096])
def test_c3_batches_on_c2_filter(c2, batch):
    """C3 power-law queries (1..50000 Zipf-weighted keys) in graph mode at
    batch sizes 1, 64 and 4096; batches of one run as separate calls."""
    ix, csr, _ = c2
    qk, qo = configs.make_queries(configs.C2, configs.C3_Q, 4096)
    ctx = _ctx(ix, 256 if batch >= 1024 else 32, 256)
    if batch == 1:
        for i in (0, 1, 2, 17, 100):
            q = qk[qo[i]:qo[i + 1]]
            o = np.array([0, q.size], np.uint64)
            want = csr.union_batch_digest(q, o, 1)
            _run_checked(ctx, q, o, ab.AB_PATH_AUTO, want, calls=2, ids_on=(0, 1), bitmaps_on=(0, 1),
                         what=f"C3 batch 1 query {i}")
    else:
        k0 = qk[: int(qo[batch])]
        o0 = qo[: batch + 1]
        want = csr.union_batch_digest(k0, o0, NT)
        paths = _run_checked(ctx, k0, o0, ab.AB_PATH_AUTO, want, calls=3,
                             ids_on=(2,) if batch > 64 else (0, 1, 2), what=f"C3 batch {batch}")
        assert (paths == ab.AB_PATH_DENSE).any() and (paths == ab.AB_PATH_SPARSE).any(), "both paths fire"
    ctx.close()


def test_c3_65536_queries_one_batch():
    """65,536 C3 queries in one graph-mode batch: above the 65535 of a grid's
    y dimension, still device-dispatched.  C2's distributions at 1/16 of the
    size (250k values x 62.5k keys), so the 65,536 id lists (about 40 GB)
    fit in memory."""
    import torch

    spec = configs.scaled(configs.C2, 1 / 16)
    k, v = configs.make_filter(spec)
    qk, qo = configs.make_queries(spec, configs.C3_Q, 65536)
    csr = oracle.CSR(k, v, spec.nkeys, spec.nvalues)
    want = csr.union_batch_digest(qk, qo, NT)
    ix = ab.Index(k, v, spec.nkeys, spec.nvalues)
    ctx = _ctx(ix, 256, 256)
    paths = _run_checked(ctx, qk, qo, ab.AB_PATH_AUTO, want, calls=3, ids_on=(2,), bitmaps_on=(0, 2),
                         what="C3 65536 queries")
    assert (paths == ab.AB_PATH_DENSE).any() and (paths == ab.AB_PATH_SPARSE).any()
    ctx.close()
    ix.close()
    torch.cuda.empty_cache()


def test_c5_served_alone_equals_in_mix_equals_oracle():
    """The C5 filter (16M values x 4M keys, ~400M pairs) built on the device
    and served: 256 C3-mix requests submitted at once (continuous batches,
    two in flight), then a few alone; every result's count and id digest
    equal the oracle's, and a Poisson replay's counts equal them too."""
    import torch

    spec = configs.C5
    k, v = configs.make_filter(spec)
    csr = oracle.CSR(k, v, spec.nkeys, spec.nvalues)
    dk = torch.from_numpy(k.view(np.int32)).cuda()
    dv = torch.from_numpy(v.view(np.int32)).cuda()
    ix = ab.Index.build_device(dk, dv, spec.nkeys, spec.nvalues, 10)
    del dk, dv, k, v
    torch.cuda.empty_cache()
    nq = 256
    qk, qo = configs.make_queries(spec, configs.C3_Q, nq)
    qs = [qk[qo[i]:qo[i + 1]] for i in range(nq)]
    wc, _, wi = csr.union_batch_digest(qk, qo, NT)
    srv = ab.Server(ix, max_batch=64, max_wait_us=200, inflight=2, deliver=ab.Server.DELIVER_HOST,
                    arena_bytes=6 << 30)
    try:
        if os.path.exists(MODEL):
            srv.load_model(MODEL)
        for c in srv.contexts:
            c.set_dense_group(128)
            c.set_dense_threads(512)
        print("[workload] C5 in mix", flush=True)
        tickets = [srv.submit(q) for q in qs]
        got = {}
        while len(got) < nq:
            for r in srv.poll(cap=64, wait_ms=200):
                assert r["status"] == 0, r
                d, asc = oracle.ids_digest(r["ids"], [0], [r["count"]])
                got[r["ticket"]] = (r["count"], int(d[0]), bool(asc[0]), r["path"])
                srv.release(r["ticket"])
        mix = [got[t] for t in tickets]
        for i in range(nq):
            assert mix[i][0] == wc[i] and mix[i][1] == int(wi[i]) and mix[i][2], f"C5 q{i} in mix != oracle"
        used = {m[3] for m in mix}
        assert ab.AB_PATH_DENSE in used and ab.AB_PATH_SPARSE in used
        assert srv.stats()["mean_batch"] > 1
        print("[workload] C5 alone", flush=True)
        for i in (0, 1, 5, 77, 255):
            t = srv.submit(qs[i])
            rs = []
            while not rs:
                rs = srv.poll(cap=4, wait_ms=200)
            assert len(rs) == 1 and rs[0]["ticket"] == t
            d, _ = oracle.ids_digest(rs[0]["ids"], [0], [rs[0]["count"]])
            srv.release(t)
            assert (rs[0]["count"], int(d[0])) == mix[i][:2], f"C5 q{i}: alone != in mix"
        print("[workload] C5 Poisson replay", flush=True)
        counts, lat, wall = srv.replay(qk, qo, rate_qps=20000.0, seed=5)
        assert np.array_equal(counts.astype(np.uint64), wc)
    finally:
        srv.close()
        ix.close()
        torch.cuda.empty_cache()
