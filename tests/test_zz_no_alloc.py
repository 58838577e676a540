"""The no-allocation sparse entry point (ab_list_union_sparse_noalloc) against
ab_list_union_sparse, the entry point it is a variant of, through the C ABI:
bit-exact bitmaps (padding included) on two calls in a row, the second with
a smaller input, plus the oracle and the size limit."""
import ctypes as C

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 18


def _lib():
    ab.load_library()
    return ab._lib


def _run_noalloc(ctx, d_keys, n, d_res, nbytes):
    st = _lib().ab_list_union_sparse_noalloc(ctx._h, d_keys.data_ptr() if n else None, n, d_res.data_ptr(), nbytes)
    if st != 0:
        raise ab.AdaBoolError(st, _lib().ab_last_error().decode())


def _run_existing(ctx, keys, d_res, nbytes):
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    st = _lib().ab_list_union_sparse(ctx._h, k.ctypes.data if k.size else None, k.size, d_res.data_ptr(), nbytes)
    if st != 0:
        raise ab.AdaBoolError(st, _lib().ab_last_error().decode())


def test_noalloc_matches_existing_entry_point_twice():
    import torch

    k, v = configs.make_filter(configs.C1)
    nk, nv = configs.C1.nkeys, configs.C1.nvalues
    ix = ab.Index(k, v, nk, nv)
    ctx = ab.Context(ix)
    csr = oracle.CSR(k, v, nk, nv)
    sw = ix.stride_words
    nbytes = sw * 4
    rng = np.random.default_rng(0xADAB0701)
    # a large query (every key, some twice, some out of range), then a smaller one
    big = np.concatenate([np.arange(nk, dtype=np.uint32), rng.integers(0, nk, 3000).astype(np.uint32),
                          np.array([nk, nk + 5, 2**32 - 1], np.uint32)])
    small = rng.integers(0, nk, 37).astype(np.uint32)
    for name, q in (("big", big), ("small", small)):
        print(f"[noalloc] {name}: {q.size} keys", flush=True)
        d_keys = torch.from_numpy(q.view(np.int32)).cuda()
        got = torch.full((sw,), -1, dtype=torch.int32, device="cuda")
        want = torch.full((sw,), -1, dtype=torch.int32, device="cuda")
        _run_noalloc(ctx, d_keys, q.size, got, nbytes)
        _run_existing(ctx, q, want, nbytes)
        ctx.sync()
        g = got.cpu().numpy().view(np.uint32)
        w = want.cpu().numpy().view(np.uint32)
        assert np.array_equal(g, w), f"{name}: no-alloc bitmap differs from ab_list_union_sparse"
        assert np.array_equal(g, csr.union(q)[0]), f"{name}: no-alloc bitmap differs from the oracle"
    # an empty query writes an all-zero bitmap
    got = torch.full((sw,), -1, dtype=torch.int32, device="cuda")
    _run_noalloc(ctx, torch.zeros(1, dtype=torch.int32, device="cuda"), 0, got, nbytes)
    ctx.sync()
    assert not got.cpu().numpy().any()
    ctx.close()


def test_noalloc_limits_are_errors():
    import torch

    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    ctx = ab.Context(ix)
    sw = ix.stride_words
    d_res = torch.zeros(sw, dtype=torch.int32, device="cuda")
    d_keys = torch.zeros(MAX_KEYS + 1, dtype=torch.int32, device="cuda")
    with pytest.raises(ab.AdaBoolError) as e:
        _run_noalloc(ctx, d_keys, MAX_KEYS + 1, d_res, sw * 4)
    assert e.value.status == 1 and "exceed" in str(e.value)
    with pytest.raises(ab.AdaBoolError) as e:
        _run_noalloc(ctx, d_keys, 4, d_res, sw * 4 - 1)
    assert e.value.status == 1
    # the largest input passes and matches the oracle
    rng = np.random.default_rng(3)
    q = rng.integers(0, configs.C1.nkeys, MAX_KEYS).astype(np.uint32)
    d_keys = torch.from_numpy(q.view(np.int32)).cuda()
    _run_noalloc(ctx, d_keys, MAX_KEYS, d_res, sw * 4)
    ctx.sync()
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    assert np.array_equal(d_res.cpu().numpy().view(np.uint32), csr.union(q)[0])
    ctx.close()
