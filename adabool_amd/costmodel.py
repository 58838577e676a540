"""Learned dispatch: timing collection, training and regret of the cost model.

Labels come from the method itself, exactly as the reference harness does it
(test/adsfilter_q_all.cu:414-438): every query is run through BOTH paths and
each is timed.  Two label sets, one per model slot:

* single (MODEL_SINGLE): one synchronized single-query call per path
  (enqueue -> stream synchronize, host wall clock), best of `repeats`;
* batch (MODEL_BATCH): the query's share of a batch, from HIP-event device
  times of device-dispatched batches (one fused union launch) over random
  groups of queries, each group run once on either path: dense -- the
  group's mask and union kernels divided by the group size (one pass over the
  index serves the group; groups of 256, the width the batches are tuned
  to); sparse -- the plan divided by the group size, plus
  the route pass and the slice items in proportion to the query's records
  M(q) (the work of both follows the records).  Timing a sparse query alone would charge it
  a whole launch that a batch shares.

Each model (12 features -> 64 -> 64 -> 2, log milliseconds of dense and
sparse) is trained by the library's own HIP kernels (ab_mlp_train); the
batch network is early-stopped on the regret of a seeded tenth of the
training queries held back for validation.  The
threshold rule (dense iff M(q) >= tau * R, tau fit on the same training
timings) is both the fallback and the yardstick: regret = sum of the chosen
path's time over the sum of the per-query best time, minus one.

Run as a module on a GPU host to produce the committed model:
    python -m adabool_amd.costmodel --out adabool_amd/models/cost_model_c4.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from dataclasses import replace
from typing import List

import numpy as np

from . import AB_PATH_DENSE, AB_PATH_SPARSE, Context, Index
from . import configs

# Seeded indices spanning 1M..16M values and Zipf s 0.8..1.3 (C2-like rows).
TRAIN_INDICES = [
    (1_000_000, 0.8), (2_000_000, 1.3), (4_000_000, 1.0), (8_000_000, 1.2), (16_000_000, 1.1), (2_000_000, 0.9),
]
HELDOUT_INDEX = (4_000_000, 1.1)  # the C2 shape, held out of training


def filter_spec(nvalues: int, s: float, k: int) -> configs.FilterSpec:
    return replace(configs.C2, name=f"C4-{nvalues // 1_000_000}M-s{s}", seed=0xADAB0400 + k, nvalues=nvalues,
                   nkeys=nvalues // 4, zipf_s=s)


def query_mix(spec: configs.FilterSpec, nq: int, seed: int):
    """Half power-law sizes (C3), half lognormal (C2); Zipf-weighted keys
    with the filter's own exponent, plus a uniform-key slice."""
    from . import gen_queries

    n1, n2 = nq // 2, nq // 4
    n3 = nq - n1 - n2
    parts = [
        gen_queries(seed, spec.nkeys, n1, size_dist=2, p1=1.5, lo=1, hi=50_000, key_dist=1, zipf_s=spec.zipf_s,
                    perm_seed=spec.seed),
        gen_queries(seed + 1, spec.nkeys, n2, size_dist=1, p1=200.0, p2=1.2, lo=8, hi=20_000, key_dist=1,
                    zipf_s=spec.zipf_s, perm_seed=spec.seed),
        gen_queries(seed + 2, spec.nkeys, n3, size_dist=1, p1=500.0, p2=1.5, lo=1, hi=50_000, key_dist=0),
    ]
    keys = np.concatenate([p[0] for p in parts])
    off = [np.zeros(1, np.uint64)]
    base = 0
    for k, o in parts:
        off.append(o[1:] + base)
        base += int(o[-1])
    return keys, np.concatenate(off)


def time_both(ctx: Context, ix: Index, keys: np.ndarray, qoff: np.ndarray, repeats: int = 2):
    """Per query: features, M, and best-of-repeats wall ms of dense and sparse."""
    import torch

    nq = qoff.size - 1
    sw = ix.stride_words
    nv = ix.info["nvalues"]
    bm = torch.empty(sw, dtype=torch.int32, device="cuda")
    ids = torch.empty(nv, dtype=torch.int32, device="cuda")
    cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    feats = np.zeros((nq, 12), np.float32)
    M = np.zeros(nq, np.uint64)
    t = np.full((nq, 2), np.inf)
    for i in range(nq):
        q = keys[int(qoff[i]):int(qoff[i + 1])]
        feats[i], M[i] = ix.features_host(q)
        for r in range(repeats):
            for j, p in enumerate((AB_PATH_DENSE, AB_PATH_SPARSE)):
                s0 = time.perf_counter()
                ctx.query_device(q, p, bm, ids, cnt)
                ctx.sync()
                t[i, j] = min(t[i, j], (time.perf_counter() - s0) * 1e3)
    return feats, M, t


def time_batch_shares(ctx: Context, ix: Index, keys: np.ndarray, qoff: np.ndarray, M: np.ndarray, seed: int,
                      repeats: int = 2, group: int = 256):
    """Per query: device ms of its share of a batch on each path (see module doc)."""
    import torch

    nq = qoff.size - 1
    sw = ix.stride_words
    bm = torch.empty(group * sw, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t = np.full((nq, 2), np.inf)
    ctx.set_batch_mode(1)
    ctx.set_dense_group(group)
    ctx.set_timing(True)
    try:
        rng = np.random.default_rng(seed)
        for _ in range(repeats):
            perm = rng.permutation(nq)
            for g0 in range(0, nq, group):
                members = perm[g0:g0 + group]
                gk = np.concatenate([keys[int(qoff[m]):int(qoff[m + 1])] for m in members])
                go = np.concatenate([[0], np.cumsum([int(qoff[m + 1] - qoff[m]) for m in members])]).astype(np.uint64)
                gd = torch.from_numpy(gk.view(np.int32)).cuda() if gk.size else torch.zeros(1, dtype=torch.int32,
                                                                                             device="cuda")
                torch.cuda.synchronize()
                ctx.timing(reset=True)
                ctx.query_batch_device(gd, go, AB_PATH_DENSE, bm)
                k = ctx.timing(reset=True)
                share = (k["fused_union"][0] + k["dense_mask"][0]) / members.size
                t[members, 0] = np.minimum(t[members, 0], share)
                ctx.query_batch_device(gd, go, AB_PATH_SPARSE, bm)
                k = ctx.timing(reset=True)
                fixed = (k["sparse_zero"][0] + k["sparse_plan"][0]) / members.size
                w = M[members].astype(np.float64)
                frac = w / w.sum() if w.sum() > 0 else np.full(members.size, 1.0 / members.size)
                # the route pass and the fused launch's slice items follow the records
                work = k["sparse_union"][0] + k["fused_union"][0]
                t[members, 1] = np.minimum(t[members, 1], fixed + work * frac)
    finally:
        ctx.set_timing(False)
    # a query with no records costs its zero-fill only; keep labels positive
    return np.maximum(t, 1e-6)


def fit_threshold(M: np.ndarray, R: np.ndarray, t: np.ndarray) -> float:
    """tau minimizing total time of the rule dense iff M >= tau * R."""
    frac = M.astype(np.float64) / R.astype(np.float64)
    cands = np.unique(np.concatenate([[0.0, 1.0 + 1e-9], frac]))
    order = np.argsort(frac)
    fs, td, tsp = frac[order], t[order, 0], t[order, 1]
    # choosing tau between fs[k-1] and fs[k]: first k sparse, rest dense
    cs = np.concatenate([[0.0], np.cumsum(tsp)])
    cd = np.concatenate([np.cumsum(td[::-1])[::-1], [0.0]])
    total = cs + cd
    k = int(np.argmin(total))
    if k == 0:
        return float(fs[0]) if fs.size else 0.0
    if k >= fs.size:
        return float(fs[-1] + 1e-9)
    del cands
    return float((fs[k - 1] + fs[k]) / 2)


def regret(choice_dense: np.ndarray, t: np.ndarray) -> float:
    chosen = np.where(choice_dense, t[:, 0], t[:, 1])
    return float(chosen.sum() / np.minimum(t[:, 0], t[:, 1]).sum() - 1.0)


def _train(ctx: Context, slot: int, F, T, tr, epochs: int, va=None, rounds: int = 1):
    """Trains the slot's network on the rows tr.  With a validation mask va,
    training runs in `rounds` rounds of epochs // rounds epochs and keeps the
    parameters of the round with the lowest validation regret (early
    stopping).  Returns the choices on every row, the loss curve and the
    validation regret of the kept parameters."""
    ctx.clear_model(slot)
    Y = np.log(T).astype(np.float32)
    if va is None or rounds <= 1:
        curve = ctx.train_model(F[tr], Y[tr], epochs=epochs, batch=256, lr=2e-3, seed=7 + slot, slot=slot)
        _, sel = ctx.predict(F, slot=slot)
        return sel == AB_PATH_DENSE, curve, None
    per = max(1, epochs // rounds)
    best, best_r, curve = None, np.inf, []
    for r in range(rounds):
        curve.extend(ctx.train_model(F[tr], Y[tr], epochs=per, batch=256, lr=2e-3, seed=7 + slot + 1000 * r,
                                     slot=slot).tolist())
        _, sel = ctx.predict(F[va], slot=slot)
        rv = regret(sel == AB_PATH_DENSE, T[va])
        if rv < best_r:
            best_r, best = rv, ctx.get_model(slot)
    ctx.set_model(*best, slot=slot)
    _, sel = ctx.predict(F, slot=slot)
    return sel == AB_PATH_DENSE, np.asarray(curve, np.float32), best_r


def main(argv: List[str] = None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="cost_model_c4.json")
    ap.add_argument("--data-out", default="cost_timings_c4.npz")
    ap.add_argument("--queries", type=int, default=2000, help="timed queries per index")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--rounds", type=int, default=12, help="early-stopping rounds of the batch network")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0, help="index size factor (for quick runs)")
    ap.add_argument("--reuse-data", default="", help="timings file of an earlier run: keep its features, M and "
                                                     "single-query labels, time the batch shares again")
    ap.add_argument("--reuse-model", default="", help="with --reuse-data: model file whose single slot is kept")
    a = ap.parse_args(argv)

    import torch

    assert torch.cuda.is_available()
    rows = []
    specs = [(filter_spec(int(nv * a.scale), s, k), False) for k, (nv, s) in enumerate(TRAIN_INDICES)]
    specs.append((filter_spec(int(HELDOUT_INDEX[0] * a.scale), HELDOUT_INDEX[1], 99), True))
    old = dict(np.load(a.reuse_data)) if a.reuse_data else None
    for n_spec, (spec, held) in enumerate(specs):
        t0 = time.time()
        pk, pv = configs.make_filter(spec)
        ix = Index(pk, pv, spec.nkeys, spec.nvalues)
        del pk, pv
        ctx = Context(ix)
        keys, qoff = query_mix(spec, a.queries, spec.seed + 17)
        if old is not None:
            sl = slice(n_spec * a.queries, (n_spec + 1) * a.queries)
            f, M, t1 = old["features"][sl], old["M"][sl], old["t_single_ms"][sl]
            assert M.size == qoff.size - 1 and int(old["R"][sl][0]) == ix.info["nrec_total"], "timings file mismatch"
        else:
            f, M, t1 = time_both(ctx, ix, keys, qoff, a.repeats)
        tb = time_batch_shares(ctx, ix, keys, qoff, M, spec.seed, a.repeats)
        R = np.full(M.size, ix.info["nrec_total"], np.uint64)
        rows.append((f, M, R, t1, tb, np.full(M.size, held)))
        print(f"[costmodel] {spec.name}: R={ix.info['nrec_total']} {a.queries} queries timed in "
              f"{time.time() - t0:.1f}s; dense wins single {(t1[:, 0] < t1[:, 1]).mean():.2f} "
              f"batch {(tb[:, 0] < tb[:, 1]).mean():.2f}", file=sys.stderr, flush=True)
        ctx.close()
        ix.close()
    F = np.concatenate([r[0] for r in rows])
    M = np.concatenate([r[1] for r in rows])
    R = np.concatenate([r[2] for r in rows])
    T1 = np.concatenate([r[3] for r in rows])
    TB = np.concatenate([r[4] for r in rows])
    H = np.concatenate([r[5] for r in rows])
    os.makedirs(os.path.dirname(os.path.abspath(a.data_out)), exist_ok=True)
    np.savez_compressed(a.data_out, features=F, M=M, R=R, t_single_ms=T1, t_batch_ms=TB, heldout=H)
    tr, te = ~H, H
    spec = filter_spec(int(TRAIN_INDICES[0][0] * a.scale), TRAIN_INDICES[0][1], 0)
    pk, pv = configs.make_filter(spec)
    ix = Index(pk, pv, spec.nkeys, spec.nvalues)
    ctx = Context(ix)
    out = {"meta": {"trained_on": [list(x) for x in TRAIN_INDICES], "heldout": list(HELDOUT_INDEX),
                    "train_queries": int(tr.sum()), "heldout_queries": int(te.sum())}}
    kept = json.load(open(a.reuse_model)) if a.reuse_model else None
    # a seeded tenth of the training queries validates the batch network
    # (early stopping on validation regret); the held-out index stays unseen
    rng = np.random.default_rng(0xADAB04)
    va = tr & (rng.random(tr.size) < 0.1)
    fit = tr & ~va
    for slot, name, T in ((0, "single", T1), (1, "batch", TB)):
        if kept is not None and name == "single":
            out[name] = kept[name]
            print(name, "kept from", a.reuse_model, flush=True)
            continue
        if slot == 1:
            tau = fit_threshold(M[fit], R[fit], T[fit])
            learned, curve, val_r = _train(ctx, slot, F, T, fit, a.epochs, va=va, rounds=a.rounds)
        else:
            tau = fit_threshold(M[tr], R[tr], T[tr])
            learned, curve, val_r = _train(ctx, slot, F, T, tr, a.epochs)
        thr = M.astype(np.float64) >= tau * R.astype(np.float64)
        res = {
            "tau": tau,
            "train_regret_threshold": regret(thr[tr], T[tr]), "train_regret_learned": regret(learned[tr], T[tr]),
            "heldout_regret_threshold": regret(thr[te], T[te]), "heldout_regret_learned": regret(learned[te], T[te]),
            "first_loss": float(curve[0]), "final_loss": float(curve[-1]),
            "labels": "log ms; " + ("single-query synchronized wall time" if slot == 0 else
                                    "device time of the query's share of a batch of 256"),
        }
        if val_r is not None:
            res["validation_regret_learned"] = val_r
            res["validation_regret_threshold"] = regret(thr[va], T[va])
            res["validation_queries"] = int(va.sum())
        print(name, json.dumps(res), flush=True)
        out[name] = dict(ctx.model_dict(slot), meta=res)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    json.dump(out, open(a.out, "w"))


if __name__ == "__main__":
    main()
