#!/usr/bin/env python3
"""Benchmark of the AdaBool OR-filtering path on one MI355X.

A step is one batched pass over the 1024 C2 queries (4M values x 1M keys,
~80M pairs, Zipf s=1.1; LogNormal(ln 200, 1.2) query sizes in [8, 20000] with
Zipf-weighted keys): per-query dispatch (learned cost model when one is
loaded, else the threshold rule), the persistent sparse union and the grouped
dense tile kernels, then bitmap -> ascending-id compaction.  Inputs (index and
query keys) are resident in HBM before timing starts; every step writes every
query's full result bitmap, id list and count.

Prints one JSON line; see the metric definition in the output.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def list_bytes_per_query(ix, keys: np.ndarray, qoff: np.ndarray):
    """Compressed-list bytes (4 + payload per record) each query's keys own."""
    import adabool_amd as ab

    nk = ix.info["nkeys"]
    valid = keys < nk
    kk = np.where(valid, keys, 0).astype(np.int64)
    per_key = np.zeros(keys.size, np.float64)
    for t in range(ab.NTYPES):
        segs = _segs(ix, t)
        cnt = (segs[kk + 1].astype(np.int64) - segs[kk].astype(np.int64)) * valid
        per_key += cnt * (4 + ab.PAYLOAD_BYTES[t])
    cs = np.concatenate([[0.0], np.cumsum(per_key)])
    return cs[qoff[1:].astype(np.int64)] - cs[qoff[:-1].astype(np.int64)]


_SEG_CACHE = {}


def _segs(ix, t):
    key = (id(ix), t)
    if key not in _SEG_CACHE:
        _SEG_CACHE[key] = ix.key_segments(t)
    return _SEG_CACHE[key]


# kernel class -> kernel symbol prefix in rocprofv3 output
KERNEL_SYMBOL = {"fused_union": "ab::k_union_fused", "write_ids": "ab::k_write_ids", "dense_tiles": "ab::k_dense_tiles",
                 "sparse_union": "ab::k_sparse_union", "popcount": "ab::k_popcount", "decide": "ab::k_features"}


def measured_traffic(path: str, kclass: str):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled against a plain
    copy) of the kernel behind `kclass`, from a committed rocprofv3 summary."""
    if not path or not os.path.exists(path) or kclass not in KERNEL_SYMBOL:
        return None, None
    d = json.load(open(path))
    cands = [(v["dispatches"], k, v) for k, v in d.get("kernels", {}).items() if k.startswith(KERNEL_SYMBOL[kclass])]
    if not cands:
        return None, None
    _, name, v = max(cands)
    return float(v["fetch_bytes_median"] + v["write_bytes_median"]), f"{os.path.relpath(path, ROOT)}: {name} median"


def cpu_model() -> str:
    """lscpu's model name, with the CPU family and model numbers (a
    virtualized host may report only the vendor's product line)."""
    name, fam, mod = "unknown", None, None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and name == "unknown":
                name = v
            elif k == "cpu family" and fam is None:
                fam = v
            elif k == "model" and mod is None:
                mod = v
    except OSError:
        pass
    return f"{name} (family {fam}, model {mod})" if fam is not None else name


def host_threads() -> int:
    """Host cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(pairs_k, pairs_v, spec, keys, qoff, counts_gpu, budget_s: float, sample=None, more=()):
    """Oracle (plain C, CSR-by-key union -> bitmap + ids) on the host cores,
    over a bounded sample of the same queries, about budget_s of wall time
    (a prefix, or whole passes over all of them; one thread per core across
    queries, and one thread alone).  As the checker it also compares: the
    GPU counts of those queries; the bitmaps and ids of `sample` (host copies
    of the first queries' GPU results: bitmaps [n, stride], ids, counts) by
    per-query digests; and, for each (name, keys, qoff, counts) in `more`,
    the GPU counts of every query."""
    import oracle

    nthreads = host_threads()
    csr = oracle.CSR(pairs_k, pairs_v, spec.nkeys, spec.nvalues)
    checks = {}
    if sample is not None:
        bms, ids, cnts = sample
        n = bms.shape[0]
        wc, wb, wi = csr.union_batch_digest(keys[: int(qoff[n])], qoff[: n + 1], nthreads)
        off = np.concatenate([[0], np.cumsum(cnts.astype(np.uint64))[:-1]]).astype(np.uint64)
        gi, asc = oracle.ids_digest(ids, off, cnts, nthreads)
        checks["sample_queries"] = n
        checks["sample_bitmaps_match"] = bool(np.array_equal(oracle.bitmap_digest(bms, 0, nthreads), wb))
        checks["sample_ids_match"] = bool(np.array_equal(cnts.astype(np.uint64), wc) and asc.all()
                                          and np.array_equal(gi, wi))
    for name, mk, mo, mc in more:
        _, oc, _, _ = csr.union_batch(mk, mo, nthreads)
        checks[f"{name}_counts_match"] = bool(np.array_equal(oc.astype(np.int64), np.asarray(mc, np.int64)))
        checks[f"{name}_queries"] = int(mo.size - 1)
    nq = qoff.size - 1
    # calibrate on a few queries, then size the sample to ~budget_s of work
    n0 = min(nq, 2 * nthreads)
    t0, c0, _, _ = csr.union_batch(keys, qoff[: n0 + 1], nthreads)
    per_q = max(t0 / n0, 1e-6)
    ns = int(min(nq, max(n0, budget_s / per_q)))
    sub = qoff[: ns + 1]
    t, counts, lat, _ = csr.union_batch(keys[: int(sub[-1])], sub, nthreads)
    ok = bool(np.array_equal(counts.astype(np.int64), counts_gpu[:ns].astype(np.int64)))
    passes = 1
    while ns == nq and t < budget_s * 0.9:  # every query fits: whole passes until the budget is used
        tp, cp, lp, _ = csr.union_batch(keys, qoff, nthreads)
        ok = ok and bool(np.array_equal(cp.astype(np.int64), counts_gpu.astype(np.int64)))
        t += tp
        lat = np.concatenate([lat, lp])
        passes += 1
    # one thread: a smaller prefix of about a quarter of the budget
    n1 = int(min(nq, max(1, budget_s / 4 / (per_q * nthreads))))
    t1, _, lat1, _ = csr.union_batch(keys[: int(qoff[n1])], qoff[: n1 + 1], 1)
    what = f"{passes} passes over all {nq}" if passes > 1 else f"first {ns} of the {nq}"
    coo = coo_scan_c1()
    return {"value": ns * passes / t, "unit": "queries/sec", "cores": nthreads, "kind": "port",
            "sample": f"{what} C2 queries, oracle/oracle.c or_union_batch, {nthreads} threads, "
                      f"{t:.1f} s", "counts_match_gpu": ok, "cpu_model": cpu_model(),
            "host_cpus_visible": os.cpu_count(), "compiled": "gcc -O3 -march=x86-64-v4 (AVX-512)",
            "checks": checks,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "single_thread": {"value": n1 / t1, "queries": n1, "p50_ms": float(np.percentile(lat1, 50)),
                              "p99_ms": float(np.percentile(lat1, 99))},
            "coo_scan_c1": coo}


def coo_scan_c1(reps: int = 20):
    """Sanity figure: the reference's own ground-truth method (a scan of every
    pair per query, test/adsfilter_q_all.cu:196-208, restated as or_scan_coo)
    on C1 (1M pairs, one query of 64 distinct uniform keys), one thread; its
    bitmap is checked against the CSR union."""
    import oracle
    from adabool_amd import configs

    k, v = configs.make_filter(configs.C1)
    qk, qo = configs.make_queries(configs.C1, configs.C1_Q)
    q = qk[qo[0]:qo[1]]
    want = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues).union(q)[0]
    lat = []
    ok = True
    for _ in range(reps):
        t0 = time.perf_counter()
        bm = oracle.scan_coo(k, v, q, configs.C1.nvalues)
        lat.append(time.perf_counter() - t0)
        ok = ok and bool(np.array_equal(bm, want))
    return {"value": 1.0 / float(np.median(lat)), "unit": "queries/sec", "cores": 1, "pairs": int(k.size),
            "queries": reps, "p50_ms": float(np.median(lat)) * 1e3, "matches_csr_union": ok}


def c3_leg(ab, torch, ctx, ix, spec, pk, path, steps: int = 5, nq: int = 4096):
    """C3 mix on the C2 filter: one batch of 4096 power-law queries (1..50000
    Zipf-weighted keys) per step through the same autotuned context (graph
    mode), ids packed at offsets from each query's id bound.  Steps run one
    at a time; the union's roofline counts each query's list bytes and its
    bitmap (no sharing credit) over the sparse route + fused union time.  The
    same batch then runs under the threshold rule and with every query dense
    (other_dispatch), with counts equal to the first rule's."""
    from adabool_amd import configs

    qk, qo = configs.make_queries(spec, configs.C3_Q, nq)
    nnz_key = np.bincount(pk, minlength=spec.nkeys).astype(np.int64)
    valid = qk < spec.nkeys
    per_key = np.where(valid, nnz_key[np.where(valid, qk, 0)], 0)
    cs = np.concatenate([[0], np.cumsum(per_key)])
    ub = np.minimum(cs[qo[1:].astype(np.int64)] - cs[qo[:-1].astype(np.int64)], spec.nvalues)
    room = (ub + 3) // 4 * 4
    off = np.concatenate([[0], np.cumsum(room)[:-1]]).astype(np.uint64)
    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    d_ids = torch.empty(int(room.sum()) + 4, dtype=torch.int32, device="cuda")
    d_cnt = torch.empty(nq, dtype=torch.int32, device="cuda")
    paths = None
    for _ in range(2):  # first call runs as launches, the second is captured
        paths = ctx.query_batch_packed(d_keys, qo, path, d_ids, off, None, d_cnt)
    ctx.sync()
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        ctx.query_batch_packed(d_keys, qo, path, d_ids, off, None, d_cnt)
        ctx.sync()
        lat.append((time.perf_counter() - s0) * 1e3)
    wall = time.perf_counter() - t0
    counts = d_cnt.cpu().numpy().view(np.uint32).astype(np.int64)
    ctx.timing(reset=True)
    ctx.set_timing(True)
    for _ in range(2):
        ctx.query_batch_packed(d_keys, qo, path, d_ids, off, None, d_cnt)
    ctx.sync()
    kt = ctx.timing(reset=True)
    ctx.set_timing(False)
    # the same batch under the other dispatch rules: a mix where the choice
    # matters (same counts on every rule)
    rules = {}
    for name, p in (("threshold", ab.AB_PATH_THRESHOLD), ("all_dense", ab.AB_PATH_DENSE)):
        for _ in range(2):
            pp = ctx.query_batch_packed(d_keys, qo, p, d_ids, off, None, d_cnt)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(3):
            ctx.query_batch_packed(d_keys, qo, p, d_ids, off, None, d_cnt)
            ctx.sync()
        rules[name] = {"qps": round(nq * 3 / (time.perf_counter() - t1), 1),
                       "dense_queries": int((pp == ab.AB_PATH_DENSE).sum()),
                       "counts_equal": bool(np.array_equal(d_cnt.cpu().numpy().view(np.uint32), counts))}
        log(f"C3 mix batch {nq}, {name} dispatch: {rules[name]['qps']:.0f} q/s, dense {rules[name]['dense_queries']}")
    union_ms = (kt["fused_union"][0] + kt["sparse_union"][0]) / 2
    alg = float((list_bytes_per_query(ix, qk, qo) + ix.stride_bytes).sum())
    out = {"workload": f"C2 filter, {nq} C3-mix queries (power law 1.5 on [1, 50000], Zipf-weighted keys) per "
                       f"batch; ids + counts", "queries_per_step": nq, "steps": steps,
           "qps": round(nq * steps / wall, 1), "batch_p50_ms": float(np.percentile(lat, 50)),
           "batch_p99_ms": float(np.percentile(lat, 99)), "dense_queries": int((paths == ab.AB_PATH_DENSE).sum()),
           "ids_per_query_mean": float(counts.mean()),
           "kernel_ms_per_step": {k: round(v[0] / 2, 4) for k, v in kt.items() if v[1]},
           "other_dispatch": rules,
           "union_roofline": {"bound": "hbm", "achieved": round(alg / (union_ms * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(alg / (union_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "bytes_per_step": alg, "union_ms": round(union_ms, 4)}}
    log(f"C3 mix batch {nq}: {out['qps']:.0f} q/s, batch p50 {out['batch_p50_ms']:.2f} ms, "
        f"dense {out['dense_queries']}, union {union_ms:.2f} ms")
    del d_ids, d_keys, d_cnt
    torch.cuda.empty_cache()
    return out, (qk, qo, counts)


# Serving launch settings by batch size on the C5 filter (value tile 2^10):
# 64-query groups and 256 threads keep light and medium load (batches of up
# to a few hundred queries) fastest, 256-query groups and 512 threads the
# near-full batches of saturation.
C5_TUNING = [{"min_batch": 1, "group": 64, "dense_threads": 256, "sparse_chunk": 4096, "dense_columns": True},
             {"min_batch": 512, "group": 256, "dense_threads": 512, "sparse_chunk": 4096, "dense_columns": True}]


def c5_leg(ab, torch, model_path: str, nq: int = 8192, rates=(0.0, 10000.0, 50000.0)):
    """C5 serving: the 16M x 4M filter (~400M pairs) built on the device and
    resident, C3-mix requests replayed through the serving loop (1024-query
    batches, two in flight, memory-based admission into a 128 GiB result
    arena, learned dispatch, launch settings by batch size): saturation (all
    at once) and Poisson arrivals at 10k and 50k queries/s, with p50/p99
    latency from each arrival."""
    from adabool_amd import configs

    spec = configs.C5
    t = time.time()
    pk, pv = configs.make_filter(spec)
    dk = torch.from_numpy(pk.view(np.int32)).cuda()
    dv = torch.from_numpy(pv.view(np.int32)).cuda()
    del pk, pv
    ix = ab.Index.build_device(dk, dv, spec.nkeys, spec.nvalues, 10)
    torch.cuda.synchronize()
    del dk, dv
    torch.cuda.empty_cache()
    info = ix.info
    qk, qo = configs.make_queries(spec, configs.C3_Q, nq)
    log(f"C5 index: {info['nnz']} pairs, device {info['device_bytes'] / 1e9:.2f} GB ({time.time() - t:.1f}s)")
    srv = ab.Server(ix, max_batch=1024, max_wait_us=200, inflight=2, arena_bytes=128 << 30, batch_mode=1)
    runs = []
    try:
        if os.path.exists(model_path):
            srv.load_model(model_path)
        for c in srv.contexts:  # by batch size, from a sweep of settings on this filter
            c.set_batch_tuning(C5_TUNING)
        srv.replay(qk[: int(qo[2048])], qo[:2049], 0.0, 7)
        srv.stats(reset=True)
        ref = None
        for rate in rates:
            counts, lat, wall = srv.replay(qk, qo, rate, 11)
            st = srv.stats(reset=True)
            ref = counts if ref is None else ref
            runs.append({"offered_qps": rate or None, "achieved_qps": round(nq / wall, 1),
                         "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                         "mean_batch": round(st["mean_batch"], 1), "dense_queries": int(st["dense_queries"]),
                         "counts_equal_saturation_run": bool(np.array_equal(counts, ref))})
            log(f"C5 serving, offered {rate or 'all at once'}: {runs[-1]['achieved_qps']:.0f} q/s, "
                f"p50 {runs[-1]['p50_ms']:.2f} ms, p99 {runs[-1]['p99_ms']:.2f} ms")
    finally:
        srv.close()
        ix.close()
        torch.cuda.empty_cache()
    return {"workload": f"C5: {spec.nvalues} values x {spec.nkeys} keys, {info['nnz']} pairs (device-built, "
                        f"value tile 2^10); {nq} C3-mix requests", "saturation_qps": runs[0]["achieved_qps"],
            "ids_per_query_mean": float(ref.mean()), "runs": runs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--scale", type=float, default=1.0, help="filter size factor (1.0 = C2)")
    ap.add_argument("--path", default="auto", choices=["auto", "threshold", "dense", "sparse", "learned"])
    ap.add_argument("--baseline-queries", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tau", type=float, default=None, help="single-query threshold")
    ap.add_argument("--tau-batch", type=float, default=None, help="batch threshold")
    ap.add_argument("--dense-threads", type=int, default=None)
    ap.add_argument("--group", type=int, default=None, help="dense group width")
    ap.add_argument("--no-columns", action="store_true", help="row-only dense groups")
    ap.add_argument("--calib-copy", type=int, default=0, help="MiB copied once by a plain copy kernel (PMC scale)")
    ap.add_argument("--batch-mode", type=int, default=2, choices=[0, 1, 2],
                    help="0 host dispatch, 1 device dispatch + fused launch, 2 = 1 replayed from a hipGraph")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "c2_batch_step.json"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (roofline traffic)")
    ap.add_argument("--model", default=os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json"))
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--tile-shift", type=int, default=0, help="value tile of the built index (0: automatic)")
    ap.add_argument("--tile-shifts", default="9,10,12",
                    help="value-tile sizes (log2) the autotuner also tries, one device-built index each")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3-mix batch leg")
    ap.add_argument("--no-c5", action="store_true", help="skip the C5 serving leg")
    ap.add_argument("--streams", type=int, default=2,
                    help="contexts (one stream, graph and output buffers each) that consecutive timed steps "
                         "alternate over, so one step's compaction runs beside the next one's union; 1: every "
                         "step waits for the previous one")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (across GPUs the path would shard by value window: each GPU "
              "holds the records of its value range and writes its slice of every result). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import adabool_amd as ab
    from adabool_amd import configs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    spec = configs.C2 if args.scale == 1.0 else configs.scaled(configs.C2, args.scale)
    t = time.time()
    pk, pv = configs.make_filter(spec)
    log(f"filter {spec.name}: {pk.size} pairs in {time.time() - t:.1f}s")
    t = time.time()
    ix = ab.Index(pk, pv, spec.nkeys, spec.nvalues, args.tile_shift)
    info = ix.info
    log(f"index: R={info['nrec_total']} list_bytes={info['list_bytes']/1e6:.1f}MB key_stream="
        f"{info['key_stream_bytes']/1e6:.1f}MB tiles={info['ntiles']} device={info['device_bytes']/1e6:.0f}MB "
        f"in {time.time() - t:.1f}s")
    qk, qo = configs.make_queries(spec, configs.C2_Q, args.queries)
    nq = qo.size - 1
    path = {"auto": ab.AB_PATH_AUTO, "threshold": ab.AB_PATH_THRESHOLD, "dense": ab.AB_PATH_DENSE,
            "sparse": ab.AB_PATH_SPARSE, "learned": ab.AB_PATH_LEARNED}[args.path]
    dispatch = "threshold"
    model_meta = None

    def setup_ctx(index):
        """A context with the cost model, calibrated thresholds and the requested knobs."""
        nonlocal dispatch, model_meta
        c = ab.Context(index)
        if os.path.exists(args.model):
            d = c.load_model(args.model)
            model_meta = {k: d[k].get("meta") for k in ("single", "batch") if k in d}
            # thresholds calibrated on the same timings as the model
            if "single" in d:
                c.set_threshold(d["single"]["meta"]["tau"])
            if "batch" in d:
                c.set_batch_threshold(d["batch"]["meta"]["tau"])
            if path in (ab.AB_PATH_AUTO, ab.AB_PATH_LEARNED):
                dispatch = "learned"
        if args.tau is not None:
            c.set_threshold(args.tau)
        if args.tau_batch is not None:
            c.set_batch_threshold(args.tau_batch)
        if args.dense_threads:
            c.set_dense_threads(args.dense_threads)
        if args.group:
            c.set_dense_group(args.group)
        if args.no_columns:
            c.set_dense_columns(False)
        c.set_batch_mode(args.batch_mode)
        return c

    ctx = setup_ctx(ix)
    # the per-query baseline runs on this first index (automatic value tile),
    # whatever the autotuner picks for the batches
    base_ctx = ctx
    if args.calib_copy:
        n = args.calib_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy} MiB read + {args.calib_copy} MiB written")
        del src, dst
    sw = ix.stride_words
    nv = spec.nvalues
    d_keys = torch.from_numpy(qk.view(np.int32)).cuda()
    d_bm = torch.empty(nq * sw, dtype=torch.int32, device="cuda")
    d_ids = torch.empty(nq * nv, dtype=torch.int32, device="cuda")
    d_cnt = torch.empty(nq, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    tuned = None
    explicit = args.dense_threads or args.group or args.no_columns
    if not args.no_autotune and not explicit:
        # launch parameters on the built index, then the same for other value
        # tiles (each index built on the device from the same pairs)
        t = time.time()
        res = base_ctx.autotune(d_keys, qo, path, repeats=2)
        dflt = next((r["ms"] for r in res if r["group"] == 64 and r["dense_threads"] == 512
                     and r["sparse_chunk"] == 4096 and r["dense_columns"]), None)
        tuned = dict(res[0], default_ms=dflt, candidates=len(res), tile_shift=info["tile_shift"],
                     mismatched_candidates=sum(r["mismatched_words"] > 0 for r in res))
        by_tile = {info["tile_shift"]: res[0]["ms"]}
        extra = [int(x) for x in args.tile_shifts.split(",") if x and int(x) != info["tile_shift"]]
        if extra:
            d_pk = torch.from_numpy(pk.view(np.int32)).cuda()
            d_pv = torch.from_numpy(pv.view(np.int32)).cuda()
            for ts in extra:
                ix2 = ab.Index.build_device(d_pk, d_pv, spec.nkeys, spec.nvalues, ts)
                ctx2 = setup_ctx(ix2)
                r2 = ctx2.autotune(d_keys, qo, path, repeats=2)
                tuned["mismatched_candidates"] += sum(r["mismatched_words"] > 0 for r in r2)
                by_tile[ts] = r2[0]["ms"]
                if r2[0]["ms"] < tuned["ms"]:
                    if ctx is not base_ctx:
                        ctx.close()
                        ix.close()
                    ix, ctx = ix2, ctx2
                    tuned.update(r2[0], tile_shift=ts, candidates=tuned["candidates"] + len(r2))
                else:
                    tuned["candidates"] += len(r2)
                    ctx2.close()
                    ix2.close()
            del d_pk, d_pv
            torch.cuda.empty_cache()
            info = ix.info
        tuned["ms_by_tile_shift"] = by_tile
        log(f"autotune: {tuned} in {time.time() - t:.1f}s")

    def run(p, steps, warmup):
        """warmup + timed steps of the batch on dispatch p (no per-kernel events)."""
        for _ in range(warmup):
            ctx.query_batch_device(d_keys, qo, p, d_bm, d_ids, nv, d_cnt)
        ctx.sync()
        step_ms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            pth = ctx.query_batch_device(d_keys, qo, p, d_bm, d_ids, nv, d_cnt)
            ctx.sync()
            step_ms.append((time.perf_counter() - s0) * 1e3)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, step_ms, pth

    def kernel_times(p, steps):
        """Per-kernel-class device time with HIP events on the context stream
        (a graph-mode context runs these steps as individual launches)."""
        ctx.timing(reset=True)
        ctx.set_timing(True)
        for _ in range(steps):
            ctx.query_batch_device(d_keys, qo, p, d_bm, d_ids, nv, d_cnt)
        ctx.sync()
        kt = ctx.timing(reset=True)
        ctx.set_timing(False)
        return kt

    # ---- more contexts for pipelined steps: same index and settings, own
    # stream, graph and output buffers
    ctxs, bufs = [ctx], [(d_bm, d_ids, d_cnt)]
    for _ in range(max(1, args.streams) - 1):
        c2 = setup_ctx(ix)
        if tuned:
            c2.apply_tuning(tuned)
        ctxs.append(c2)
        bufs.append((torch.empty_like(d_bm), torch.empty_like(d_ids), torch.empty_like(d_cnt)))
    torch.cuda.synchronize()

    def run_streams(p, steps, warmup):
        """warmup, then `steps` steps enqueued round-robin over the contexts
        with no wait between them; bracketed by device synchronizations."""
        for c, (b, i, n) in zip(ctxs, bufs):
            for _ in range(warmup):
                c.query_batch_device(d_keys, qo, p, b, i, nv, n, want_paths=False)
            c.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            b, i, n = bufs[k % len(ctxs)]
            ctxs[k % len(ctxs)].query_batch_device(d_keys, qo, p, b, i, nv, n, want_paths=False)
        for c in ctxs:
            c.sync()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    wall_serial, step_ms, paths = run(path, args.steps, args.warmup)
    counts = d_cnt.cpu().numpy().view(np.uint32).astype(np.int64)
    qps_serial = nq * args.steps / wall_serial
    ndense = int((paths == ab.AB_PATH_DENSE).sum())
    log(f"batched ({dispatch}, mode {args.batch_mode}), one step at a time: {qps_serial:.0f} q/s, "
        f"{wall_serial / args.steps * 1e3:.2f} ms/step, dense {ndense}/{nq}")
    wall, qps, streams_equal = wall_serial, qps_serial, None
    if len(ctxs) > 1:
        for _, i, n in bufs[1:]:
            n.fill_(-1)
            i[:nv].fill_(-1)
        wall = run_streams(path, args.steps, args.warmup)
        qps = nq * args.steps / wall
        # every context's outputs equal the serial run's (counts, and the ids of query 0)
        ids0 = d_ids[: int(counts[0])].cpu()
        streams_equal = all(np.array_equal(n.cpu().numpy().view(np.uint32).astype(np.int64), counts)
                            and bool((i[: int(counts[0])].cpu() == ids0).all()) for _, i, n in bufs[1:])
        log(f"batched ({dispatch}, mode {args.batch_mode}), {len(ctxs)} streams: {qps:.0f} q/s, "
            f"{wall / args.steps * 1e3:.2f} ms/step; outputs equal the serial run: {streams_equal}")
    smp = None
    if not args.no_cpu:
        # host copies of the first queries' results of the timed run (the last
        # context's buffers, written by nothing else), for the checker
        ns = min(64, nq)
        sb, si, _ = bufs[-1]
        smp = (sb[: ns * sw].cpu().numpy().view(np.uint32).reshape(ns, sw),
               np.concatenate([si[i * nv: i * nv + int(counts[i])].cpu().numpy().view(np.uint32) for i in range(ns)]),
               counts[:ns].astype(np.uint32))
    ksteps = max(2, min(args.steps, 5))
    kt = kernel_times(path, ksteps)
    log("kernels: " + ", ".join(f"{k}={v[0] / ksteps:.3f}ms/{v[1] // ksteps}" for k, v in kt.items() if v[1]))
    thr_cmp = None
    nsub = max(2, args.steps // 2)
    if dispatch == "learned":
        tw, tms, tp = run(ab.AB_PATH_THRESHOLD, nsub, 1)
        thr_cmp = {"qps_one_step_at_a_time": round(nq * nsub / tw, 1), "ms_per_step": round(tw / nsub * 1e3, 3),
                   "dense_queries": int((tp == ab.AB_PATH_DENSE).sum())}
        if len(ctxs) > 1:
            thr_cmp["qps"] = round(nq * args.steps / run_streams(ab.AB_PATH_THRESHOLD, args.steps, 1), 1)
        else:
            thr_cmp["qps"] = thr_cmp["qps_one_step_at_a_time"]
        log(f"batched (threshold tau_batch={ctx.batch_threshold:.5f}): {thr_cmp['qps']:.0f} q/s")
    mode_cmp = {}
    for m in (0, 1, 2):
        if m == args.batch_mode:
            continue
        ctx.set_batch_mode(m)
        mw, _, _ = run(path, nsub, 1)
        mode_cmp[str(m)] = round(nq * nsub / mw, 1)
        log(f"batched ({dispatch}, mode {m}): {mode_cmp[str(m)]:.0f} q/s")
    ctx.set_batch_mode(args.batch_mode)

    # ---- the device's write ceiling for this output: a plain fill of the ids buffer
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    d_ids.fill_(0)
    torch.cuda.synchronize()
    ev0.record()
    d_ids.fill_(1)
    ev1.record()
    torch.cuda.synchronize()
    fill_gbs = d_ids.numel() * 4 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    log(f"plain fill of the {d_ids.numel() * 4 / 1e9:.1f} GB ids buffer: {fill_gbs:.0f} GB/s")

    # ---- algorithmic bytes (per query: 4|q| + list bytes + bitmap + 4|R| + 4)
    lb = list_bytes_per_query(ix, qk, qo)
    nkeys_q = np.diff(qo).astype(np.float64)
    stride_b = float(ix.stride_bytes)
    alg_q = 4 * nkeys_q + lb + stride_b + 4 * counts + 4
    job_gbs = alg_q.sum() * args.steps / wall / 1e9
    # per kernel class: bytes per step of that class
    dense_mask = paths == ab.AB_PATH_DENSE
    kbytes = {
        "write_ids": float((stride_b + 4 * counts + 4).sum()),
        "dense_tiles": float((lb[dense_mask] + stride_b).sum()),
        "sparse_union": float((4 * nkeys_q[~dense_mask] + lb[~dense_mask]).sum()),
        "popcount": float(stride_b * nq),
        "sparse_zero": float(stride_b * (~dense_mask).sum()),
    }
    kbytes["fused_union"] = kbytes["dense_tiles"] + kbytes["sparse_union"]
    dom = max(kt, key=lambda k: kt[k][0])
    ms_total, launches = kt[dom]
    roof = None
    if dom in kbytes and launches:
        per_launch_bytes = kbytes[dom] * ksteps / launches
        avg_ms = ms_total / launches
        ach = per_launch_bytes / (avg_ms * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.traffic, dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "bytes_per_launch": per_launch_bytes, "avg_launch_ms": avg_ms, "launches": launches}

    # ---- plain per-query threshold dispatch (one query at a time, synchronized)
    nb = min(args.baseline_queries, nq)
    d_bm1 = torch.empty(sw, dtype=torch.int32, device="cuda")
    d_ids1 = torch.empty(nv, dtype=torch.int32, device="cuda")
    d_c1 = torch.empty(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    for i in range(min(8, nb)):
        base_ctx.query_device(qk[qo[i]:qo[i + 1]], ab.AB_PATH_THRESHOLD, d_bm1, d_ids1, d_c1)
        base_ctx.sync()
    lat1 = []
    mism = 0
    t1 = time.perf_counter()
    for i in range(nb):
        s0 = time.perf_counter()
        base_ctx.query_device(qk[qo[i]:qo[i + 1]], ab.AB_PATH_THRESHOLD, d_bm1, d_ids1, d_c1)
        base_ctx.sync()
        lat1.append((time.perf_counter() - s0) * 1e3)
        if i < 16:
            mism += int(int(d_c1.cpu().numpy().view(np.uint32)[0]) != counts[i])
    base_qps = nb / (time.perf_counter() - t1)
    log(f"per-query threshold: {base_qps:.0f} q/s, p50 {np.percentile(lat1, 50):.3f} ms "
        f"p99 {np.percentile(lat1, 99):.3f} ms; count mismatches vs batch: {mism}")
    # the same queries one at a time with the single-query cost model
    learned_pq = None
    if dispatch == "learned":
        latl = []
        tl = time.perf_counter()
        for i in range(nb):
            s0 = time.perf_counter()
            base_ctx.query_device(qk[qo[i]:qo[i + 1]], ab.AB_PATH_LEARNED, d_bm1, d_ids1, d_c1)
            base_ctx.sync()
            latl.append((time.perf_counter() - s0) * 1e3)
        learned_pq = {"qps": round(nb / (time.perf_counter() - tl), 1), "p50_ms": float(np.percentile(latl, 50)),
                      "p99_ms": float(np.percentile(latl, 99))}
        log(f"per-query learned: {learned_pq['qps']:.0f} q/s, p50 {learned_pq['p50_ms']:.3f} ms "
            f"p99 {learned_pq['p99_ms']:.3f} ms")

    c3 = c3_data = None
    if not args.no_c3 and args.scale == 1.0:
        c3, c3_data = c3_leg(ab, torch, ctx, ix, spec, pk, path)

    cpu = None
    if not args.no_cpu:
        more = [("c3_mix",) + c3_data] if c3_data is not None else []
        t = time.time()
        cpu = cpu_baseline(pk, pv, spec, qk, qo, counts, args.cpu_budget, sample=smp, more=more)
        del smp
        log(f"cpu baseline: {cpu['value']:.1f} q/s on {cpu['cores']} threads ({time.time() - t:.1f}s), "
            f"counts match: {cpu['counts_match_gpu']}")

    out = {
        "metric": "queries/sec + p99 latency, learned vs threshold dispatch, bit-exact",
        "value": round(qps, 1),
        "unit": "queries/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded generators, adabool_amd/configs.py)",
        "config": {"workload": f"{spec.name}: {spec.nvalues} values x {spec.nkeys} keys, {pk.size} pairs, "
                               f"Zipf {spec.zipf_s}; {nq} Zipf-weighted queries, LogNormal(ln 200, 1.2) sizes "
                               f"in [8, 20000]; bitmap + ascending ids + counts",
                   "queries_per_step": nq, "dispatch": dispatch, "dense_queries": ndense,
                   "tile_shift": info["tile_shift"]},
        "streams": len(ctxs),
        "one_step_at_a_time": {"qps": round(qps_serial, 1), "ms_per_step": round(wall_serial / args.steps * 1e3, 3)},
        "streams_outputs_equal": streams_equal,
        "latency_ms": {"batch_p50": float(np.percentile(step_ms, 50)), "batch_p99": float(np.percentile(step_ms, 99)),
                       "per_query_threshold_p50": float(np.percentile(lat1, 50)),
                       "per_query_threshold_p99": float(np.percentile(lat1, 99))},
        "per_query_threshold_qps": round(base_qps, 1),
        "per_query_learned": learned_pq,
        "per_query_threshold_tau": base_ctx.threshold,
        "per_query_threshold_tile_shift": base_ctx.index.info["tile_shift"],
        "batched_threshold": thr_cmp,
        "batch_modes_qps": mode_cmp,
        "autotune": tuned,
        "cost_model": model_meta,
        "speedup_vs_per_query_threshold": round(qps / base_qps, 2),
        "job_algorithmic_GBps": round(job_gbs, 1),
        "write_ceiling_fill_GBps": round(fill_gbs, 1),
        "kernel_ms_per_step": {k: round(v[0] / ksteps, 4) for k, v in kt.items() if v[1]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "c3_mix_batch4096": c3,
    }
    if not args.no_c5 and args.scale == 1.0:
        # the C2 steps' outputs (about 33 GB) make room for the result arena
        bufs.clear()
        d_bm = d_ids = d_cnt = sb = si = None
        torch.cuda.empty_cache()
        out["c5_serving"] = c5_leg(ab, torch, args.model)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
