"""Pins the CPU oracle against known answers, hand-built cases, independent
algorithms and closed forms (none of these retype the oracle's formula)."""
import os
import tempfile

import numpy as np
import pytest

import oracle


def py_union(keys, vals, q, nvalues):
    """Set-based restatement for tiny inputs (test/adsfilter_q_all.cu:196-208)."""
    qs = set(int(x) for x in q)
    out = sorted({int(v) for k, v in zip(keys, vals) if int(k) in qs})
    bm = np.zeros(oracle.stride_words(nvalues), np.uint32)
    for v in out:
        bm[v >> 5] |= np.uint32(1 << (v & 31))
    return bm, np.array(out, np.uint32)


# ---------------------------------------------------------------------------
# KeyBlockMeta known answer (src/AdaBool/inverted_list.h:73-81)
# ---------------------------------------------------------------------------
def test_key_block_meta_known_answer():
    keys = [100, 101, 255, 260, 500, 510]
    starts, pos, buf = oracle.key_block_meta(4, keys)
    assert starts.tolist() == [100, 500, 510]
    assert pos.tolist() == [0, 8, 10]
    assert buf.size == 12
    # block 0: 2-byte offsets 0,1,155,160; block 1: 1-byte 0,10 + 2 pad bytes
    assert buf.tolist() == [0, 0, 1, 0, 155, 0, 160, 0, 0, 10, 0, 0]


def test_key_block_meta_widths():
    # spans 255 -> 1 byte, 256 -> 2 bytes, 65536 -> 4 bytes
    for span, w in ((255, 1), (256, 2), (65535, 2), (65536, 4)):
        keys = [7, 7 + span]
        starts, pos, buf = oracle.key_block_meta(1, keys)
        # the last block's span runs to the final key itself (0) -> 1 byte
        assert pos.tolist() == [0, w, w + 1]
        assert buf.size == w + 1


# ---------------------------------------------------------------------------
# record typing and payloads (src/AdaBool/inverted_list.h:390-470)
# ---------------------------------------------------------------------------
def test_bucket_types_by_hand():
    sizes = [1, 2, 5, 6, 9, 10, 17, 18, 33, 34]
    want = [0, 1, 1, 2, 2, 3, 3, 4, 4, 5]
    keys, vals = [], []
    for i, n in enumerate(sizes):
        base = (i + 1) << 8
        for j in range(n):
            keys.append(0)
            vals.append(base + 3 * j + 1)
    ix = oracle.RefIndex(keys, vals)
    got = []
    for t in range(6):
        hdr, pay, k, segs = ix.arrays(t)
        for h in hdr:
            got.append((int(h) >> 8, t))
        assert segs.tolist() == [0, len(hdr)]  # key 0 only: length max_key + 2
    got.sort()
    assert [t for _, t in got] == want


def test_payload_bytes_by_hand():
    # one ARRAY4 bucket {3, 7, 9} in window 0x100 and one BITSET bucket
    keys = [2, 2, 2] + [5] * 40
    vals = [0x107, 0x103, 0x109] + [0x300 + x for x in range(0, 80, 2)]
    ix = oracle.RefIndex(keys, vals)
    hdr, pay, k, segs = ix.arrays(1)
    assert hdr.tolist() == [0x103]
    assert pay.tolist() == [7, 9, 0, 0]
    assert k.tolist() == [2]
    assert segs.tolist() == [0, 0, 0, 1, 1, 1, 1]  # keys 0..5 -> length 7
    hdr, pay, k, segs = ix.arrays(5)
    assert hdr.tolist() == [0x300]
    bits = np.unpackbits(pay, bitorder="little")
    assert np.nonzero(bits)[0].tolist() == list(range(0, 80, 2))
    assert segs[6] - segs[5] == 1


def test_duplicates_dedupe_and_reference_behaviour():
    # a (k, v) pair repeated: with dedupe the bucket has one value (ARRAY1);
    # without, the reference counts the copies (size 3 -> ARRAY4 with payload
    # holding the duplicates).
    keys, vals = [1, 1, 1], [0x205, 0x205, 0x205]
    d = oracle.RefIndex(keys, vals, dedupe=True)
    assert d.nrec(0) == 1 and d.nrec(1) == 0
    r = oracle.RefIndex(keys, vals, dedupe=False)
    assert r.nrec(0) == 0 and r.nrec(1) == 1
    hdr, pay, _, _ = r.arrays(1)
    assert hdr.tolist() == [0x205] and pay.tolist() == [5, 5, 0, 0]


# ---------------------------------------------------------------------------
# emit rules (src/AdaBool/inverted_list.cu.h:61-94)
# ---------------------------------------------------------------------------
def _ids(bm):
    return np.nonzero(np.unpackbits(bm.view(np.uint8), bitorder="little"))[0].tolist()


def test_emit_array_terminator_and_kminnum():
    nv = 4096
    # ARRAY4: header + bytes until the first zero after kMinNum=1
    assert _ids(oracle.emit(1, 0x103, [7, 9, 0, 0], nv)) == [0x103, 0x107, 0x109]
    # ARRAY8: the first kMinNum=5 bytes are emitted even when zero
    got = _ids(oracle.emit(2, 0x205, [6, 7, 8, 9, 0, 10, 0, 0], nv))
    assert got == sorted({0x205, 0x206, 0x207, 0x208, 0x209, 0x200, 0x20A})
    # ARRAY1: header only
    assert _ids(oracle.emit(0, 0x3FF, [], nv)) == [0x3FF]
    # BITSET: 8 words OR-ed at (header >> 5) + j
    p = np.zeros(32, np.uint8)
    p[0] = 0b101
    p[31] = 0x80
    assert _ids(oracle.emit(5, 0x400, p, nv)) == [0x400, 0x402, 0x4FF]


# ---------------------------------------------------------------------------
# predicate: independent algorithms agree
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_union_algorithms_agree(seed):
    rng = np.random.default_rng(seed)
    nk, nv = int(rng.integers(1, 400)), int(rng.integers(1, 3000))
    n = int(rng.integers(0, 4000))
    keys = rng.integers(0, nk, n).astype(np.uint32)
    vals = rng.integers(0, nv, n).astype(np.uint32)
    if n:
        keys[: n // 10] = keys[0]  # a heavy key -> bitsets
        vals[: n // 10] = rng.integers(0, min(nv, 300), n // 10)
    csr = oracle.CSR(keys, vals, nk, nv)
    ix = oracle.RefIndex(keys, vals)
    for _ in range(5):
        q = rng.integers(0, nk + 20, int(rng.integers(0, 50))).astype(np.uint32)
        bm, ids = csr.union(q)
        assert np.array_equal(bm, oracle.scan_coo(keys, vals, q, nv))
        qb = np.zeros((nk + 31) // 32 + 1, np.uint32)
        for x in q:
            if x < nk:
                qb[x >> 5] |= np.uint32(1 << (int(x) & 31))
        for step in (4, 256):
            assert np.array_equal(bm, ix.dense_probe(qb, nk, nv, step))
        for step in (3, 256):
            assert np.array_equal(bm, ix.sparse_union(q, nv, step)[0])
        assert np.array_equal(ids, oracle.collect_ids(bm, bm.size * 32, nv))
        assert np.array_equal(ids, csr.union_ids(q))
        if n <= 400:
            pbm, pids = py_union(keys, vals, q, nv)
            assert np.array_equal(bm, pbm) and np.array_equal(ids, pids)


def test_union_tiny_by_hand():
    keys = [0, 0, 1, 2, 2, 5]
    vals = [3, 40, 3, 7, 31, 64]
    csr = oracle.CSR(keys, vals, 6, 100)
    assert csr.union([0])[1].tolist() == [3, 40]
    assert csr.union([2, 1])[1].tolist() == [3, 7, 31]
    assert csr.union([5, 5, 99])[1].tolist() == [64]
    assert csr.union([])[1].tolist() == []
    bm, _ = csr.union([0, 2])
    assert bm.size == 4  # align_up(ceil(100/8), 16) = 16 bytes
    assert bm[0] == (1 << 3) | (1 << 7) | (1 << 31) and bm[1] == 1 << 8


def test_union_ids_by_hand_and_at_the_top_of_u32():
    """The bitmap-free union: hand-checked answers, repeated keys and pairs,
    and values next to 2^32 - 1, where a bitmap would take 512 MB."""
    keys = [0, 0, 1, 2, 2, 5, 0]
    vals = [3, 40, 3, 7, 31, 64, 3]  # (0, 3) twice
    csr = oracle.CSR(keys, vals, 6, 100)
    assert csr.union_ids([0]).tolist() == [3, 40]
    assert csr.union_ids([2, 1, 2]).tolist() == [3, 7, 31]
    assert csr.union_ids([5, 5, 99]).tolist() == [64]
    assert csr.union_ids([]).tolist() == []
    top = 2 ** 32 - 2
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 20, 500).astype(np.uint32)
    vals = (top - rng.integers(0, 3000, 500)).astype(np.uint32)
    vals[:3] = [top, 0, 255]
    csr = oracle.CSR(keys, vals, 20, top + 1)
    for _ in range(10):
        q = rng.integers(0, 25, int(rng.integers(0, 12))).astype(np.uint32)
        qs = set(q.tolist())
        want = sorted({int(v) for k, v in zip(keys.tolist(), vals.tolist()) if k in qs})
        assert csr.union_ids(q).tolist() == want


def test_decode_equals_pairs():
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 50, 5000).astype(np.uint32)
    vals = rng.integers(0, 20000, 5000).astype(np.uint32)
    ix = oracle.RefIndex(keys, vals)
    k, v = ix.decode()
    a = np.unique((keys.astype(np.uint64) << 32) | vals)
    b = np.sort((k.astype(np.uint64) << 32) | v)
    assert np.array_equal(a, b)


def test_sparse_units_count():
    # LoadBalanceTransformer::Prepare cuts every list into ceil(len/step) units
    keys = [0] * 10 + [1] * 3
    vals = [x * 256 for x in range(10)] + [x * 256 for x in range(3)]
    ix = oracle.RefIndex(keys, vals)
    _, units = ix.sparse_union([0, 1], 4096, step=4)
    assert units == 3 + 1  # key 0: 10 ARRAY1 records -> 3 units; key 1: 3 -> 1


def test_plant_mistake_is_caught():
    """A mistake planted in the oracle side (a dropped list) is visible to the
    cross-check used by the parity tests."""
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 30, 3000).astype(np.uint32)
    vals = rng.integers(0, 5000, 3000).astype(np.uint32)
    csr = oracle.CSR(keys, vals, 30, 5000)
    q = np.array([1, 2, 3], np.uint32)
    good = csr.union(q)[0]
    planted = csr.union(q[:-1])[0]  # "forgot" the last key
    assert not np.array_equal(good, planted)
    assert np.array_equal(good, oracle.scan_coo(keys, vals, q, 5000))


# ---------------------------------------------------------------------------
# statistics (test/adsfilter_q_all.cu:258-307)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 5, 10, 101])
def test_stats_closed_forms(n):
    rng = np.random.default_rng(n)
    d = rng.exponential(size=n)
    s = oracle.stats(d)
    assert s["mean"] == pytest.approx(d.mean())
    assert s["median"] == pytest.approx(np.median(d))
    assert s["stddev"] == pytest.approx(d.std(ddof=0))
    assert s["min"] == d.min() and s["max"] == d.max()
    for p in (90, 95, 99):
        assert s[f"p{p}"] == pytest.approx(np.percentile(d, p, method="linear"))
    assert oracle.stats([])["mean"] == 0


# ---------------------------------------------------------------------------
# text formats (test/adsfilter_q_all.cu:127-193)
# ---------------------------------------------------------------------------
COO_TEXT = "% matrix market style comment\n# another\n\n 1 2 \n3\t4\r\n5 x\n 7   8 extra\n-1 2\n99999999999 1\n9 10"
QUERY_TEXT = "# q\n4\n\n  8 \r\nx\n15 16\n"


def test_text_parsers():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.coo")
        open(p, "w").write(COO_TEXT)
        k, v, mr, mc = oracle.read_coo(p)
        assert k.tolist() == [1, 3, 7, 9] and v.tolist() == [2, 4, 8, 10]
        assert (mr, mc) == (9, 10)
        q = os.path.join(d, "query_vec_1.coo")
        open(q, "w").write(QUERY_TEXT)
        assert oracle.read_query(q).tolist() == [4, 8, 15]


# ---------------------------------------------------------------------------
# MLP reference: finite differences and closed forms
# ---------------------------------------------------------------------------
def test_mlp_gradient_finite_differences():
    dims = (3, 5, 4, 2)
    rng = np.random.default_rng(0)
    npar = oracle.mlp_nparams(*dims)
    p = rng.normal(scale=0.5, size=npar)
    x = rng.normal(size=(7, 3))
    y = rng.normal(size=(7, 2))
    loss, g = oracle.mlp_loss_grad(p, x, y, dims)
    eps = 1e-6
    for i in rng.choice(npar, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (oracle.mlp_loss_grad(pp, x, y, dims)[0] - oracle.mlp_loss_grad(pm, x, y, dims)[0]) / (2 * eps)
        assert g[i] == pytest.approx(num, rel=1e-4, abs=1e-7)


def test_mlp_closed_forms():
    dims = (2, 3, 3, 2)
    npar = oracle.mlp_nparams(*dims)
    p = np.zeros(npar)
    p[-2:] = [1.5, -2.0]  # b3; all weights zero -> y = b3
    x = np.array([[1.0, 2.0], [3.0, -1.0]])
    y = np.array([[0.5, 0.0], [1.5, -1.0]])
    out = oracle.mlp_forward(p, x, dims)
    assert np.allclose(out, [[1.5, -2.0], [1.5, -2.0]])
    loss, g = oracle.mlp_loss_grad(p, x, y, dims)
    err = out - y
    assert loss == pytest.approx(0.5 * (err ** 2).mean())
    assert np.allclose(g[-2:], err.mean(axis=0) / 2)
    # first Adam step moves every coordinate by ~lr against its gradient sign
    q = np.arange(1.0, 5.0)
    m, v = np.zeros(4), np.zeros(4)
    grad = np.array([0.3, -2.0, 1e-3, -5.0])
    oracle.adam_step(q, m, v, grad, 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(q, np.arange(1.0, 5.0) - 0.01 * np.sign(grad), atol=1e-8)


def _py_bitmap_digest(bm):
    g = 0x9E3779B97F4A7C15
    return sum(int(x) * ((((w + 1) * g) & (2**64 - 1)) | 1) for w, x in enumerate(bm)) % 2**64


def _py_ids_digest(ids):
    return sum((int(x) * 0x9E3779B1) ^ ((int(x) ^ 0x5BD1E995) * 0x27D4EB2F) for x in ids) % 2**64


def test_digests_pinned():
    """The per-query digests: the batch form equals the single-result forms on
    or_union_csr's own outputs, both equal a pure-Python evaluation of their
    definitions, and one changed bit or id changes them."""
    rng = np.random.default_rng(21)
    nk, nv = 300, 5000
    k = rng.integers(0, nk, 20000).astype(np.uint32)
    v = rng.integers(0, nv, 20000).astype(np.uint32)
    csr = oracle.CSR(k, v, nk, nv)
    qs = [rng.choice(nk + 20, size=int(rng.integers(0, 40)), replace=False).astype(np.uint32) for _ in range(25)]
    qk = np.concatenate(qs)
    qo = np.concatenate([[0], np.cumsum([q.size for q in qs])]).astype(np.uint64)
    c, db, di = csr.union_batch_digest(qk, qo, nthreads=3)
    bms, idl = zip(*[csr.union(q) for q in qs])
    assert np.array_equal(c, [x.size for x in idl])
    assert np.array_equal(db, oracle.bitmap_digest(np.stack(bms)))
    flat = np.concatenate(idl)
    offs = np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.uint64)
    d2, asc = oracle.ids_digest(flat, offs, c, nthreads=2)
    assert np.array_equal(di, d2) and asc.all()
    for i in (0, 7, 24):
        assert int(db[i]) == _py_bitmap_digest(bms[i]) and int(di[i]) == _py_ids_digest(idl[i])
    j = int(np.argmax(c > 3))
    b = bms[j].copy()
    b[5] ^= 1 << 9
    assert oracle.bitmap_digest(b)[0] != db[j]
    x = idl[j].copy()
    x[1] += 1 if x[1] + 1 < x[2] else 0
    x[2] ^= 1
    assert oracle.ids_digest(x, [0], [x.size])[0][0] != di[j]
    y = idl[j][::-1].copy()
    assert not oracle.ids_digest(y, [0], [y.size])[1][0]
