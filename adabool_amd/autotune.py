"""Autotuning of the batched path and its persisted table.

For one index and a sample batch, `ab_ctx_autotune` times the launch grid
(dense group width x work-group size x sparse chunk x column masks) and keeps
the fastest; this module adds the index layout knob (value tile, one index
build per candidate) and batch sizes, and persists the winners per index
shape and batch-size bucket:

    python -m adabool_amd.autotune --out adabool_amd/models/tuning_c2.json

`lookup(table, index, nq)` returns the entry for the nearest tuned batch size
of an index with the same shape.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import AB_PATH_AUTO, Context, Index, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def signature(index: Index) -> str:
    i = index.info
    return f"nv{i['nvalues']}-nk{i['nkeys']}-nnz{i['nnz']}"


def lookup(table: dict, index: Index, nq: int):
    """Tuned entry for this index shape and the nearest batch size (log scale)."""
    ent = table.get("entries", {}).get(signature(index))
    if not ent:
        return None
    sizes = sorted(int(b) for b in ent)
    best = min(sizes, key=lambda b: abs(math.log2(b) - math.log2(max(nq, 1))))
    return ent[str(best)]


def tune(index: Index, keys: np.ndarray, qoff: np.ndarray, batch_sizes, model: str = None, repeats: int = 3,
         log=print):
    import torch

    ctx = Context(index)
    if model and os.path.exists(model):
        d = ctx.load_model(model)
        ctx.set_threshold(d["single"]["meta"]["tau"])
        ctx.set_batch_threshold(d["batch"]["meta"]["tau"])
    out = {}
    nq_all = qoff.size - 1
    for b in batch_sizes:
        b = min(b, nq_all)
        sub = qoff[: b + 1]
        dk = torch.from_numpy(np.ascontiguousarray(keys[: int(sub[-1])]).view(np.int32)).cuda() if sub[-1] else \
            torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        res = ctx.autotune(dk, sub, AB_PATH_AUTO, repeats)
        valid = [r for r in res if not r["mismatched_words"]]
        best, worst = valid[0], valid[-1]
        out[str(b)] = dict(best, candidates=len(res), mismatched_candidates=len(res) - len(valid), worst_ms=worst["ms"],
                           default_ms=next((r["ms"] for r in res if r["group"] == 64 and r["dense_threads"] == 512
                                            and r["sparse_chunk"] == 4096 and r["dense_columns"]), None))
        log(f"[autotune] batch {b}: best {best} ({len(res)} candidates, worst {worst['ms']:.3f} ms)")
    ctx.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="tuning_c2.json")
    ap.add_argument("--batches", default="1,4,16,64,256,1024,4096")
    ap.add_argument("--tile-shifts", default="9,10,11,12")
    ap.add_argument("--model", default=os.path.join(ROOT, "adabool_amd", "models", "cost_model_c4.json"))
    a = ap.parse_args(argv)
    spec = configs.C2
    pk, pv = configs.make_filter(spec)
    qk, qo = configs.make_queries(spec, configs.C3_Q, 4096)
    batches = [int(x) for x in a.batches.split(",")]
    table = {"workload": "C2 filter, C3 queries (power law 1.5 on [1, 50000], Zipf-weighted keys)",
             "entries": {}, "tile_shift_ms": {}}
    best_ts, best_ms, best_entry = None, math.inf, None
    for ts in [int(x) for x in a.tile_shifts.split(",")]:
        t = time.time()
        ix = Index(pk, pv, spec.nkeys, spec.nvalues, ts)
        print(f"[autotune] tile_shift {ts}: index in {time.time() - t:.1f}s", file=sys.stderr, flush=True)
        ent = tune(ix, qk, qo, batches, a.model, log=lambda m: print(m, file=sys.stderr, flush=True))
        ref = ent[str(max(b for b in batches if b <= qo.size - 1))]["ms"]
        table["tile_shift_ms"][str(ts)] = ref
        if ref < best_ms:
            best_ts, best_ms, best_entry, sig = ts, ref, ent, signature(ix)
        ix.close()
    for e in best_entry.values():
        e["tile_shift"] = best_ts
    table["entries"][sig] = best_entry
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    json.dump(table, open(a.out, "w"), indent=1)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
