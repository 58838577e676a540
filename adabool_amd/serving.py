"""Serving benchmark: the C5 filter (16M values x 4M keys, ~400M pairs) served
to an open-loop Poisson stream of C3-mix queries (power-law sizes on
[1, 50000], Zipf-weighted keys) through the serving loop (ab_server), beside
the plain per-query synchronous threshold dispatch on the same GPU.

    python -m adabool_amd.serving [--scale 1.0] [--queries 20000]
                                  [--rates 50000,100000,200000,500000]

Prints one JSON line: per offered rate the achieved throughput and latency
percentiles (CalculateStats definition, from each query's scheduled arrival),
the saturation throughput (all queries offered at once) and the per-query
baseline.  Results stay in device memory (a GPU consumer); --deliver host
adds the copy of every id list to pinned host memory.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import AB_PATH_AUTO, AB_PATH_THRESHOLD, Context, Index, Server, configs, stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


TUNING_KEYS = ("group", "dense_threads", "sparse_chunk", "dense_columns")


def tuning_table(results):
    """Launch settings by batch size (Context.set_batch_tuning) from autotune
    winners on samples of ascending size: [(sample queries, winner dict)].
    Each winner applies from the geometric mean of its sample size and the
    previous one (the first from batches of 1); a sample no larger than the
    previous one adds nothing."""
    table, lo = [], 0
    for n, r in results:
        if table and n <= lo:
            continue
        start = 1 if not table else max(int(round((lo * n) ** 0.5)), table[-1]["min_batch"] + 1)
        table.append(dict({k: r[k] for k in TUNING_KEYS}, min_batch=start, sample=int(n), ms=r["ms"]))
        lo = n
    return table


def log(msg):
    print(f"[serving] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0, help="C5 size factor")
    ap.add_argument("--queries", type=int, default=20000)
    ap.add_argument("--rates", default="10000,20000,50000,100000,200000,500000")
    ap.add_argument("--max-batch", type=int, default=1024)
    ap.add_argument("--max-wait-us", type=int, default=200)
    ap.add_argument("--inflight", type=int, default=2)
    ap.add_argument("--batch-mode", type=int, default=1)
    ap.add_argument("--arena-gb", type=float, default=128.0)
    ap.add_argument("--deliver", choices=["device", "host"], default="device")
    ap.add_argument("--baseline-queries", type=int, default=2000)
    ap.add_argument("--model", default=os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json"))
    ap.add_argument("--tile-shifts", default="10", help="value tiles the autotuner also tries (device builds)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--settings", default="",
                    help="group/threads/chunk/tile used as given instead of autotuning, e.g. 128/512/8192/10")
    a = ap.parse_args(argv)

    import torch

    assert torch.cuda.is_available(), "the serving benchmark needs a GPU"
    spec = configs.C5 if a.scale == 1.0 else configs.scaled(configs.C5, a.scale)
    t = time.time()
    pk, pv = configs.make_filter(spec)
    log(f"filter {spec.name}: {pk.size} pairs in {time.time() - t:.1f}s")
    t = time.time()
    ix = Index(pk, pv, spec.nkeys, spec.nvalues)
    t_host = time.time() - t
    info = ix.info
    log(f"index (host build): R={info['nrec_total']} device={info['device_bytes'] / 1e9:.2f} GB in {t_host:.1f}s")
    # the same index built on the GPU from device-resident pairs
    dk = torch.from_numpy(pk.view(np.int32)).cuda()
    dv = torch.from_numpy(pv.view(np.int32)).cuda()
    torch.cuda.synchronize()
    t = time.time()
    ixd = Index.build_device(dk, dv, spec.nkeys, spec.nvalues)
    torch.cuda.synchronize()
    t_dev = time.time() - t
    same = ixd.info["key_stream_bytes"] == info["key_stream_bytes"] and ixd.info["nrec_total"] == info["nrec_total"]
    log(f"index (device build): {t_dev:.2f}s, same shape: {same}")
    ixd.close()
    qk, qo = configs.make_queries(spec, configs.C3_Q, a.queries)
    nq = qo.size - 1
    sizes = np.diff(qo)
    log(f"{nq} queries: |q| median {np.median(sizes):.0f}, mean {sizes.mean():.0f}, max {sizes.max()}")

    # ---- autotune on a sample batch: launch parameters, then other value tiles
    tuned = None
    serve_ix = ix
    if a.settings:
        g, t, ch, ts = (int(x) for x in a.settings.split("/"))
        tuned = {"group": g, "dense_threads": t, "sparse_chunk": ch, "dense_columns": True, "tile_shift": ts,
                 "given": True}
        if ts != info["tile_shift"]:
            serve_ix = Index.build_device(dk, dv, spec.nkeys, spec.nvalues, ts)
        log(f"settings as given: {tuned}")
    elif not a.no_autotune:
        # launch settings by batch size: samples of 32 and 256 queries (light
        # and medium load) and of max_batch (saturation; it also picks the
        # value tile), each applied from the geometric mean of its sample size
        # and the one below
        def sample(ns):
            ns = min(nq, ns)
            return torch.from_numpy(np.ascontiguousarray(qk[: int(qo[ns])]).view(np.int32)).cuda(), qo[: ns + 1]

        def tune_on(index, s):
            c = Context(index)
            if os.path.exists(a.model):
                c.load_model(a.model)
            r = c.autotune(s[0], s[1], AB_PATH_AUTO, repeats=2)[0]
            c.close()
            return r

        big, mid, small = sample(a.max_batch), sample(256), sample(32)
        best = dict(tune_on(ix, big), tile_shift=info["tile_shift"])
        for ts in [int(x) for x in a.tile_shifts.split(",") if x and int(x) != info["tile_shift"]]:
            cand = Index.build_device(dk, dv, spec.nkeys, spec.nvalues, ts)
            r = tune_on(cand, big)
            log(f"autotune tile 2^{ts}: {r['ms']:.3f} ms vs {best['ms']:.3f} ms")
            if r["ms"] < best["ms"]:
                if serve_ix is not ix:
                    serve_ix.close()
                serve_ix, best = cand, dict(r, tile_shift=ts)
            else:
                cand.close()
        sizes = [int(x[1].size - 1) for x in (small, mid, big)]
        results = [(sizes[0], tune_on(serve_ix, small))]
        if sizes[1] > sizes[0]:
            results.append((sizes[1], tune_on(serve_ix, mid)))
        results.append((sizes[2], best))
        tuned = {"tile_shift": best["tile_shift"], "table": tuning_table(results)}
        log(f"autotune: {tuned}")
    del dk, dv, pk, pv
    torch.cuda.empty_cache()

    out = {"workload": f"{spec.name}: {spec.nvalues} values x {spec.nkeys} keys, {info['nnz']} pairs, Zipf "
                       f"{spec.zipf_s}; {nq} C3-mix queries (power law 1.5 on [1, 50000], Zipf-weighted keys), "
                       f"Poisson arrivals",
           "deliver": a.deliver, "max_batch": a.max_batch, "inflight": a.inflight, "max_wait_us": a.max_wait_us,
           "batch_mode": a.batch_mode, "index_build_s": {"host": round(t_host, 2), "device": round(t_dev, 3)},
           "autotune": tuned}

    # ---- per-query synchronous threshold dispatch (the x5 yardstick) -------
    ctx = Context(ix)
    if os.path.exists(a.model):
        d = ctx.load_model(a.model)
        ctx.set_threshold(d["single"]["meta"]["tau"])
    nb = min(a.baseline_queries, nq)
    sw = ix.stride_words
    d_bm = torch.empty(sw, dtype=torch.int32, device="cuda")
    d_ids = torch.empty(spec.nvalues, dtype=torch.int32, device="cuda")
    d_c = torch.empty(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    for i in range(min(16, nb)):
        ctx.query_device(qk[qo[i]:qo[i + 1]], AB_PATH_THRESHOLD, d_bm, d_ids, d_c)
    ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for i in range(nb):
        s0 = time.perf_counter()
        ctx.query_device(qk[qo[i]:qo[i + 1]], AB_PATH_THRESHOLD, d_bm, d_ids, d_c)
        ctx.sync()
        lat.append((time.perf_counter() - s0) * 1e3)
    base_qps = nb / (time.perf_counter() - t0)
    st = stats(lat)
    out["per_query_threshold"] = {"qps": round(base_qps, 1), "p50_ms": st["p50"], "p99_ms": st["p99"],
                                  "queries": nb}
    log(f"per-query threshold: {base_qps:.0f} q/s p50 {st['p50']:.3f} ms p99 {st['p99']:.3f} ms")
    ctx.close()
    del d_bm, d_ids, d_c
    torch.cuda.empty_cache()

    # ---- serving loop --------------------------------------------------------
    srv = Server(serve_ix, max_batch=a.max_batch, max_wait_us=a.max_wait_us, inflight=a.inflight,
                 arena_bytes=int(a.arena_gb * (1 << 30)), path=AB_PATH_AUTO,
                 deliver=Server.DELIVER_HOST if a.deliver == "host" else Server.DELIVER_DEVICE,
                 batch_mode=a.batch_mode)
    if os.path.exists(a.model):
        srv.load_model(a.model)
    if tuned:
        for c in srv.contexts:
            if "table" in tuned:
                c.set_batch_tuning(tuned["table"])
            else:
                c.apply_tuning(tuned)
    # warm-up (allocations, code objects)
    srv.replay(qk[: int(qo[min(nq, 2048)])], qo[: min(nq, 2048) + 1], 0.0, 7)
    srv.stats(reset=True)
    runs = []
    counts_ref = None
    for rate in [0.0] + [float(x) for x in a.rates.split(",") if x]:
        counts, lat, wall = srv.replay(qk, qo, rate, 11)
        st_srv = srv.stats(reset=True)
        ls = stats(lat)
        if counts_ref is None:
            counts_ref = counts.copy()
        same = bool(np.array_equal(counts, counts_ref))
        r = {"offered_qps": rate or None, "achieved_qps": round(nq / wall, 1), "p50_ms": ls["p50"],
             "p90_ms": ls["p90"], "p99_ms": ls["p99"], "max_ms": ls["max"], "batches": st_srv["batches"],
             "mean_batch": round(st_srv["mean_batch"], 1), "dense_queries": st_srv["dense_queries"],
             "counts_equal_saturation_run": same}
        runs.append(r)
        log(f"offered {rate or 'all-at-once'}: {r['achieved_qps']:.0f} q/s, p50 {ls['p50']:.2f} ms "
            f"p99 {ls['p99']:.2f} ms, {st_srv['batches']} batches of {st_srv['mean_batch']:.1f}")
    srv.close()
    out["runs"] = runs
    out["saturation_qps"] = runs[0]["achieved_qps"]
    out["speedup_vs_per_query_threshold"] = round(runs[0]["achieved_qps"] / base_qps, 2)
    out["result_ids_per_query_mean"] = float(counts_ref.mean())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
