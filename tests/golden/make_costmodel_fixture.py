"""Writes tests/golden/cost_heldout.npz: a seeded sample of the held-out timed
queries (C2-shaped index, 4M values, Zipf 1.1) that the committed cost model
was never trained on: features, M, R and both label sets.

Input: the timing file written by `python -m adabool_amd.costmodel`
(--data-out, default cost_timings_c4.npz)."""
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else "cost_timings_c4.npz"
out = sys.argv[2] if len(sys.argv) > 2 else "tests/golden/cost_heldout.npz"
d = np.load(src)
h = np.nonzero(d["heldout"])[0]
rng = np.random.default_rng(0xADAB)
pick = np.sort(rng.choice(h, size=min(4000, h.size), replace=False))
np.savez_compressed(out, features=d["features"][pick].astype(np.float32), M=d["M"][pick], R=d["R"][pick],
                    t_single_ms=d["t_single_ms"][pick].astype(np.float32),
                    t_batch_ms=d["t_batch_ms"][pick].astype(np.float32))
print(out, pick.size)
