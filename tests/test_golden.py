"""The committed C1 fixture (tests/golden/c1_fixture.npz, written by
tests/golden/make_c1_fixture.py) pins the seeded generators, the index shape
and the oracle's results; the GPU paths must reproduce the same ids."""
import hashlib
import os

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_fixture.npz")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def fx():
    d = dict(np.load(FIX))
    k, v = configs.make_filter(configs.C1)
    return d, k, v


def test_generators_reproduce_the_filter(fx):
    d, k, v = fx
    assert k.size == int(d["nnz_input"])
    assert sha(k) == str(d["keys_sha"]) and sha(v) == str(d["values_sha"])


def test_index_shape(fx):
    d, k, v = fx
    info = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues, host_only=True).info
    assert info["nrec"] == [int(x) for x in d["nrec"]]
    assert info["nnz"] == int(d["nnz"]) and info["list_bytes"] == int(d["list_bytes"])
    assert info["key_stream_bytes"] == int(d["key_stream_bytes"]) and info["tile_shift"] == int(d["tile_shift"])


def test_oracle_results(fx):
    d, k, v = fx
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    qk, qo = d["qkeys"], d["qoff"]
    for i in range(qo.size - 1):
        ids = csr.union(qk[qo[i]:qo[i + 1]])[1].astype(np.uint32)
        assert ids.size == int(d["counts"][i]) and sha(ids) == str(d["ids_sha"][i]), f"q{i}"
        a, b = int(d["small_off"][i]), int(d["small_off"][i + 1])
        if b > a:
            assert np.array_equal(ids, d["small_ids"][a:b])


@pytest.mark.gpu
@pytest.mark.parametrize("path", [ab.AB_PATH_DENSE, ab.AB_PATH_SPARSE, ab.AB_PATH_AUTO])
def test_gpu_batch_matches_fixture(fx, path):
    d, k, v = fx
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    ctx = ab.Context(ix)
    qk, qo = d["qkeys"], d["qoff"]
    print(f"golden path={path}", flush=True)
    bms, ids, counts, paths = ctx.query_batch(qk, qo, path)
    for i in range(qo.size - 1):
        assert counts[i] == int(d["counts"][i]) and sha(ids[i].astype(np.uint32)) == str(d["ids_sha"][i]), f"q{i}"
    ctx.close()
