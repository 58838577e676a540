"""C3 batch-size sweep: the C2 filter and C3 queries (power-law sizes on
[1, 50000], Zipf-weighted keys) run through the batched path at batch sizes
1..4096 and through plain per-query synchronous threshold dispatch.

    python -m adabool_amd.sweep [--queries 16384] [--batches 1,4,16,64,256,1024,4096]

Every batch writes each query's ascending ids into a packed result buffer
(offsets from each query's id upper bound) and its count; the counts of every
batch size are checked against the batch-size-1 run.  Prints one JSON line.
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


def log(msg):
    print(f"[sweep] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--queries", type=int, default=16384)
    ap.add_argument("--batches", default="1,4,16,64,256,1024,4096")
    ap.add_argument("--min-queries", type=int, default=2048, help="queries timed per batch size at least")
    ap.add_argument("--baseline-queries", type=int, default=2048)
    ap.add_argument("--model", default=os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json"))
    ap.add_argument("--tuning", default=os.path.join(ROOT, "adabool_amd", "models", "tuning_c2.json"))
    a = ap.parse_args(argv)

    import torch

    assert torch.cuda.is_available(), "the sweep needs a GPU"
    spec = configs.C2
    pk, pv = configs.make_filter(spec)
    # id upper bound per key: its list length
    nnz_key = np.bincount(pk, minlength=spec.nkeys).astype(np.int64)
    tuning = json.load(open(a.tuning)) if os.path.exists(a.tuning) else None
    ts = 0
    if tuning:
        ent = next(iter(tuning["entries"].values()))
        ts = next(iter(ent.values())).get("tile_shift", 0)
    ix = Index(pk, pv, spec.nkeys, spec.nvalues, ts)
    del pk, pv
    qk, qo = configs.make_queries(spec, configs.C3_Q, a.queries)
    nq = qo.size - 1
    sizes = np.diff(qo).astype(np.int64)
    valid = qk < spec.nkeys
    per_key = np.where(valid, nnz_key[np.where(valid, qk, 0)], 0)
    cs = np.concatenate([[0], np.cumsum(per_key)])
    ub = np.minimum(cs[qo[1:].astype(np.int64)] - cs[qo[:-1].astype(np.int64)], spec.nvalues)
    log(f"{nq} queries: |q| median {np.median(sizes):.0f} mean {sizes.mean():.0f}; id bound mean {ub.mean():.0f}")

    ctx = Context(ix)
    if os.path.exists(a.model):
        d = ctx.load_model(a.model)
        ctx.set_threshold(d["single"]["meta"]["tau"])
        ctx.set_batch_threshold(d["batch"]["meta"]["tau"])
    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    out = {"workload": f"C2 filter ({spec.nvalues} values x {spec.nkeys} keys), {nq} C3 queries "
                       f"(power law 1.5 on [1, 50000], Zipf-weighted keys); ids + counts",
           "tile_shift": ix.info["tile_shift"], "runs": []}
    ref_counts = None
    for b in [int(x) for x in a.batches.split(",")]:
        if tuning:
            cfg = lookup(tuning, ix, b)
            if cfg:
                ctx.apply_tuning(cfg)
        nrun = min(nq, max(a.min_queries, b * 4))
        nrun -= nrun % b
        # packed id offsets per batch, 16-byte aligned
        span = 0
        for s in range(0, nrun, b):
            span = max(span, int(((ub[s:s + b] + 3) // 4 * 4).sum()))
        d_ids = torch.empty(max(span, 4), dtype=torch.int32, device="cuda")
        d_cnt = torch.empty(b, dtype=torch.int32, device="cuda")
        counts = np.zeros(nrun, np.int64)
        lat = []
        # warm-up on the first batch (allocation, graph capture)
        offs = []
        for s in range(0, nrun, b):
            o = np.concatenate([[0], np.cumsum((ub[s:s + b - 1] + 3) // 4 * 4)]).astype(np.uint64)
            offs.append(o)
        q0 = qo[: b + 1]
        ctx.query_batch_packed(d_keys, q0, AB_PATH_AUTO, d_ids, offs[0], None, d_cnt)
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j, s in enumerate(range(0, nrun, b)):
            sub = (qo[s:s + b + 1] - qo[s]).astype(np.uint64)
            s0 = time.perf_counter()
            ctx.query_batch_packed(d_keys[int(qo[s]):] if int(qo[s]) < qk.size else d_keys, sub, AB_PATH_AUTO,
                                   d_ids, offs[j], None, d_cnt)
            ctx.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
            counts[s:s + b] = d_cnt[:b].cpu().numpy().view(np.uint32)
        wall = time.perf_counter() - t0
        if ref_counts is None:
            ref_counts = counts
        ok = bool(np.array_equal(counts, ref_counts[:nrun])) if nrun <= ref_counts.size else bool(
            np.array_equal(counts[: ref_counts.size], ref_counts))
        st = stats(lat)
        r = {"batch": b, "queries": nrun, "qps": round(nrun / wall, 1), "batch_p50_ms": st["p50"],
             "batch_p99_ms": st["p99"], "counts_equal_batch1": ok}
        out["runs"].append(r)
        log(f"batch {b}: {r['qps']:.0f} q/s over {nrun} queries, batch p50 {st['p50']:.3f} ms p99 {st['p99']:.3f} ms,"
            f" counts equal: {ok}")
        del d_ids
        torch.cuda.empty_cache()

    # plain per-query synchronous threshold dispatch
    nb = min(a.baseline_queries, nq)
    sw = ix.stride_words
    d_bm = torch.empty(sw, dtype=torch.int32, device="cuda")
    d_ids = torch.empty(spec.nvalues, dtype=torch.int32, device="cuda")
    d_c = torch.empty(1, dtype=torch.int32, device="cuda")
    lat = []
    t0 = time.perf_counter()
    for i in range(nb):
        s0 = time.perf_counter()
        ctx.query_device(qk[qo[i]:qo[i + 1]], AB_PATH_THRESHOLD, d_bm, d_ids, d_c)
        ctx.sync()
        lat.append((time.perf_counter() - s0) * 1e3)
    base = nb / (time.perf_counter() - t0)
    st = stats(lat)
    out["per_query_threshold"] = {"qps": round(base, 1), "p50_ms": st["p50"], "p99_ms": st["p99"], "queries": nb}
    out["best_batched_qps"] = max(r["qps"] for r in out["runs"])
    out["speedup_best_batched_vs_per_query"] = round(out["best_batched_qps"] / base, 2)
    log(f"per-query threshold: {base:.0f} q/s; best batched {out['best_batched_qps']:.0f} q/s")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
