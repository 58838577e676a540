"""Id-only batches (ab_query_batch_ids, no result bitmap) beside the bitmap
path.

* C2 filter (4M values x 1M keys, ~80M pairs) with two query sets:
  - small results: queries of 64 distinct uniform keys (the C1 query shape),
    a few thousand ids each against a 512 KB bitmap;
  - the C3 mix (power-law sizes on [1, 50000], Zipf-weighted keys), whose
    lists hold 3.4M values at the median query: the bitmap's ground;
  id-only and bitmap batches (ids + counts packed at each query's id bound).
* C2 wide: the same pairs with every value moved to a distinct place in the
  u32 range (v -> v * 0x9E3779B1 mod 2^32, nvalues = 2^32 - 1), where a result
  bitmap takes 512 MB per query: id-only batches of the 64-key queries, and the
  bitmap path one query at a time, the only form of it that fits.

    python -m adabool_amd.ids_bench [--queries 4096] [--batches 1,16,256,1024,4096]

Counts of every run are checked against each other.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import AB_PATH_AUTO, AB_PATH_THRESHOLD, Context, Index, configs, stats
from .autotune import lookup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WIDE_MUL = 0x9E3779B1


def log(msg):
    print(f"[ids] {msg}", file=sys.stderr, flush=True)


def id_bounds(pk, nkeys, nvalues, qk, qo):
    nnz_key = np.bincount(pk, minlength=nkeys).astype(np.int64)
    valid = qk < nkeys
    per_key = np.where(valid, nnz_key[np.where(valid, qk, 0)], 0)
    cs = np.concatenate([[0], np.cumsum(per_key)])
    return np.minimum(cs[qo[1:].astype(np.int64)] - cs[qo[:-1].astype(np.int64)], nvalues)


def run_batches(ctx, d_keys, qo, ub, b, nrun, fn):
    """Times nrun queries in batches of b through fn(sub_keys, sub_qoff,
    d_ids, offsets, d_counts); returns (q/s, batch latencies, counts)."""
    import torch

    span = 0
    offs = []
    for s in range(0, nrun, b):
        room = (ub[s:s + b] + 3) // 4 * 4
        span = max(span, int(room.sum()))
        offs.append(np.concatenate([[0], np.cumsum(room[:-1])]).astype(np.uint64))
    d_ids = torch.empty(max(span, 4), dtype=torch.int32, device="cuda")
    d_cnt = torch.empty(b, dtype=torch.int32, device="cuda")
    counts = np.zeros(nrun, np.int64)
    nk = d_keys.numel()

    def sub(s):
        return d_keys[int(qo[s]):] if int(qo[s]) < nk else d_keys, (qo[s:s + b + 1] - qo[s]).astype(np.uint64)

    # warm-up over every batch once: the context's scratch grows to the
    # largest batch (and graphs are captured) before the timed pass
    for j, s in enumerate(range(0, nrun, b)):
        ks, qs = sub(s)
        fn(ks, qs, d_ids, offs[j], d_cnt)
    ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for j, s in enumerate(range(0, nrun, b)):
        ks, qs = sub(s)
        s0 = time.perf_counter()
        fn(ks, qs, d_ids, offs[j], d_cnt)
        ctx.sync()
        lat.append((time.perf_counter() - s0) * 1e3)
        counts[s:s + b] = d_cnt[:b].cpu().numpy().view(np.uint32)
    wall = time.perf_counter() - t0
    del d_ids
    torch.cuda.empty_cache()
    return nrun / wall, lat, counts


def bench_set(ctx, ix, tuning, d_keys, qo, ub, batches, min_q, name):
    nq = qo.size - 1
    ref, rows = None, []
    for b in batches:
        if b > nq:
            continue
        if tuning:
            cfg = lookup(tuning, ix, b)
            if cfg:
                ctx.apply_tuning(cfg)
        nrun = min(nq, max(min_q, b * 4))
        nrun -= nrun % b
        bq, bl, bc = run_batches(ctx, d_keys, qo, ub, b, nrun,
                                 lambda k, q, i, o, c: ctx.query_batch_packed(k, q, AB_PATH_AUTO, i, o, None, c))
        iq, il, ic = run_batches(ctx, d_keys, qo, ub, b, nrun,
                                 lambda k, q, i, o, c: ctx.query_batch_ids_device(k, q, i, o, c))
        if ref is None:
            ref = bc
        n = min(ref.size, bc.size)
        ok = bool(np.array_equal(bc, ic)) and bool(np.array_equal(bc[:n], ref[:n]))
        rows.append({"batch": b, "queries": nrun, "bitmap_qps": round(bq, 1), "ids_only_qps": round(iq, 1),
                     "bitmap_batch_p99_ms": stats(bl)["p99"], "ids_only_batch_p99_ms": stats(il)["p99"],
                     "mean_ids": float(bc.mean()), "counts_equal": ok})
        log(f"{name} batch {b}: bitmap {bq:.0f} q/s, id-only {iq:.0f} q/s, mean ids {bc.mean():.0f}, "
            f"counts equal: {ok}")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--queries", type=int, default=4096, help="64-key queries")
    ap.add_argument("--c3-queries", type=int, default=1024)
    ap.add_argument("--batches", default="1,16,256,1024,4096")
    ap.add_argument("--min-queries", type=int, default=2048)
    ap.add_argument("--c3-min-queries", type=int, default=256)
    ap.add_argument("--wide-baseline-queries", type=int, default=256)
    ap.add_argument("--model", default=os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json"))
    ap.add_argument("--tuning", default=os.path.join(ROOT, "adabool_amd", "models", "tuning_c2.json"))
    ap.add_argument("--id-method", type=int, default=0, choices=[0, 1, 2],
                    help="id-only method: 0 automatic, 1 by sort, 2 by value slice where it applies")
    ap.add_argument("--no-wide", action="store_true", help="skip the nvalues = 2^32 - 1 filter")
    a = ap.parse_args(argv)

    import torch

    assert torch.cuda.is_available(), "the id-only benchmark needs a GPU"
    spec = configs.C2
    pk, pv = configs.make_filter(spec)
    batches = [int(x) for x in a.batches.split(",")]
    uq = configs.QuerySpec(0xADAB0006, a.queries, 0, 0.0, 0.0, 64, 64, 0, 1.0)
    qk, qo = configs.make_queries(spec, uq, a.queries)
    ck, co = configs.make_queries(spec, configs.C3_Q, a.c3_queries)
    nq = qo.size - 1
    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    d_ckeys = torch.from_numpy(ck.view(np.int32)).cuda()
    out = {"workload": f"C2 filter ({spec.nvalues} values x {spec.nkeys} keys, {pk.size} pairs); {nq} queries of "
                       f"64 distinct uniform keys, {co.size - 1} C3 queries; ids + counts"}

    # ---- C2: bitmap path vs id-only ------------------------------------------
    tuning = json.load(open(a.tuning)) if os.path.exists(a.tuning) else None
    ts = 0
    if tuning:
        ts = next(iter(next(iter(tuning["entries"].values())).values())).get("tile_shift", 0)
    ix = Index(pk, pv, spec.nkeys, spec.nvalues, ts)
    ctx = Context(ix)
    ctx.set_id_method(a.id_method)
    if os.path.exists(a.model):
        d = ctx.load_model(a.model)
        ctx.set_threshold(d["single"]["meta"]["tau"])
        ctx.set_batch_threshold(d["batch"]["meta"]["tau"])
    out["c2_uniform64"] = bench_set(ctx, ix, tuning, d_keys, qo, id_bounds(pk, spec.nkeys, spec.nvalues, qk, qo),
                                    batches, a.min_queries, "C2 uniform-64")
    out["c2_c3mix"] = bench_set(ctx, ix, tuning, d_ckeys, co, id_bounds(pk, spec.nkeys, spec.nvalues, ck, co),
                                [b for b in batches if b <= a.c3_queries], a.c3_min_queries, "C2 C3-mix")
    ctx.close()
    ix.close()

    # ---- C2 wide: values over the whole u32 range -----------------------------
    if not a.no_wide:
        wv = (pv.astype(np.uint64) * WIDE_MUL % (1 << 32)).astype(np.uint32)
        keep = wv != 0xFFFFFFFF  # the one value that nvalues = 2^32 - 1 cannot hold
        wk, wv = pk[keep], wv[keep]
        nv_wide = 2**32 - 1
        t = time.time()
        ixw = Index.build_device(wk, wv, spec.nkeys, nv_wide)
        log(f"wide index (device build) in {time.time() - t:.1f}s: {ixw.info['nrec_total']} records, "
            f"bitmap {ixw.stride_bytes / 2**20:.0f} MiB per query")
        ubw = id_bounds(wk, spec.nkeys, nv_wide, qk, qo)
        ctx = Context(ixw)
        ctx.set_id_method(a.id_method)
        wide_ref = None
        runs = []
        for b in batches:
            nrun = min(nq, max(a.min_queries, b * 4))
            nrun -= nrun % b
            iq, il, ic = run_batches(ctx, d_keys, qo, ubw, b, nrun,
                                     lambda k, q, i, o, c: ctx.query_batch_ids_device(k, q, i, o, c))
            if wide_ref is None:
                wide_ref = ic
            ok = bool(np.array_equal(ic[: wide_ref.size], wide_ref[: ic.size]))
            runs.append({"batch": b, "queries": nrun, "ids_only_qps": round(iq, 1), "batch_p99_ms": stats(il)["p99"],
                         "counts_equal_batch1": ok})
            log(f"wide batch {b}: id-only {iq:.0f} q/s, counts equal: {ok}")
        # the bitmap path fits one query at a time: 512 MB bitmap, ids at the bound
        nb = min(a.wide_baseline_queries, nq)
        d_bm = torch.empty(ixw.stride_words, dtype=torch.int32, device="cuda")
        d_ids = torch.empty(int(max(ubw[:nb].max(), 4)), dtype=torch.int32, device="cuda")
        d_c = torch.empty(1, dtype=torch.int32, device="cuda")
        ctx.query_device(qk[qo[0]:qo[1]], AB_PATH_THRESHOLD, d_bm, d_ids, d_c)
        ctx.sync()
        lat, cnt = [], np.zeros(nb, np.int64)
        t0 = time.perf_counter()
        for i in range(nb):
            s0 = time.perf_counter()
            ctx.query_device(qk[qo[i]:qo[i + 1]], AB_PATH_THRESHOLD, d_bm, d_ids, d_c)
            ctx.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            cnt[i] = int(d_c.cpu().numpy().view(np.uint32)[0])
        base = nb / (time.perf_counter() - t0)
        st = stats(lat)
        same = bool(np.array_equal(cnt, wide_ref[:nb])) if wide_ref.size >= nb else None
        out["wide"] = {"nvalues": nv_wide, "pairs": int(wk.size), "bitmap_bytes_per_query": ixw.stride_bytes,
                       "runs": runs, "bitmap_per_query": {"qps": round(base, 1), "p50_ms": st["p50"], "p99_ms": st["p99"],
                                                          "queries": nb, "counts_equal_ids_only": same},
                       "best_ids_only_vs_bitmap_per_query": round(max(r["ids_only_qps"] for r in runs) / base, 1)}
        log(f"wide bitmap path per query: {base:.0f} q/s (p50 {st['p50']:.3f} ms), counts equal: {same}")
        ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
