"""Writes tests/golden/c1_fixture.npz: the C1 workload pinned as data.

* the seeded C1 filter (100k values x 10k keys): pair count and SHA-256 of the
  key and value arrays the generator produces;
* the C1 query (64 distinct uniform keys) and 24 mixed queries (Zipf-weighted
  keys, LogNormal sizes, plus empty / out-of-range / duplicate-key / all-key
  queries), with the oracle's id count and SHA-256 of the ascending ids of
  each (CSR union, oracle/oracle.c), and the ids themselves for the queries
  with at most 5000 of them;
* the index shape the build must produce (records per type, list bytes,
  key-stream bytes at the automatic value tile).

Run from the repository root: python tests/golden/make_c1_fixture.py"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import adabool_amd as ab  # noqa: E402
import oracle  # noqa: E402
from adabool_amd import configs  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def queries():
    q0k, q0o = configs.make_queries(configs.C1, configs.C1_Q)
    qk, qo = ab.gen_queries(0xADAB0701, configs.C1.nkeys, 20, size_dist=1, p1=60.0, p2=1.3, lo=1, hi=3000,
                            key_dist=1, zipf_s=1.0, perm_seed=configs.C1.seed)
    qs = [q0k[q0o[0]:q0o[1]]] + [qk[qo[i]:qo[i + 1]] for i in range(20)]
    nk = configs.C1.nkeys
    qs.append(np.zeros(0, np.uint32))
    qs.append(np.array([nk, nk + 3, 2**32 - 1], np.uint32))
    qs.append(np.array([7, 7, 7, 19, 7], np.uint32))
    qs.append(np.arange(nk, dtype=np.uint32))
    keys = np.concatenate(qs).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64)
    return keys, off


def main():
    out = os.path.join(ROOT, "tests", "golden", "c1_fixture.npz")
    k, v = configs.make_filter(configs.C1)
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    keys, off = queries()
    ids = [csr.union(keys[off[i]:off[i + 1]])[1].astype(np.uint32) for i in range(off.size - 1)]
    small = [a if a.size <= 5000 else np.zeros(0, np.uint32) for a in ids]
    id_off = np.concatenate([[0], np.cumsum([a.size for a in small])]).astype(np.uint64)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues, host_only=True)
    info = ix.info
    np.savez_compressed(
        out, nnz_input=np.uint64(k.size), keys_sha=np.array(sha(k)), values_sha=np.array(sha(v)),
        qkeys=keys, qoff=off, counts=np.array([a.size for a in ids], np.uint64),
        ids_sha=np.array([sha(a) for a in ids]), small_ids=np.concatenate(small).astype(np.uint32), small_off=id_off,
        nrec=np.array(info["nrec"], np.uint64), nnz=np.uint64(info["nnz"]), list_bytes=np.uint64(info["list_bytes"]),
        key_stream_bytes=np.uint64(info["key_stream_bytes"]), tile_shift=np.uint32(info["tile_shift"]))
    print(out, off.size - 1, "queries,", sum(a.size for a in ids), "ids")


if __name__ == "__main__":
    main()
