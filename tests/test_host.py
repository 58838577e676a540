"""Product host logic (no GPU): the built layouts against the oracle, the C ABI
exports, generators, loaders and statistics."""
import ctypes as C
import os
import re
import tempfile

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "adabool.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(ab_[a-z0-9_]+)\s*\(", hdr))
    lib = C.CDLL(ab.lib_path())
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared, "no declarations parsed"
    # and the Python binding binds what the header declares
    unbound = [s for s in sorted(declared) if s not in ab._SIGS and s not in ab._OPTIONAL_SIGS]
    assert not unbound, unbound


def _pairs_random(seed, nk, nv, n, dups=True):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, nk, n).astype(np.uint32)
    v = rng.integers(0, nv, n).astype(np.uint32)
    if dups and n > 10:
        k[-n // 10:] = k[: n // 10]
        v[-n // 10:] = v[: n // 10]
    # a few dense windows -> every record type
    hk = np.repeat(np.arange(3, dtype=np.uint32), 200)
    hv = (np.tile(np.arange(200, dtype=np.uint32), 3) * 1) % max(nv, 1)
    return np.concatenate([k, hk % nk]), np.concatenate([v, hv])


@pytest.mark.parametrize("case", ["c1", "rand_a", "rand_b", "tiny"])
def test_key_major_layout_matches_oracle(case):
    if case == "c1":
        k, v = configs.make_filter(configs.C1)
    elif case == "rand_a":
        k, v = _pairs_random(1, 500, 70000, 50000)
    elif case == "rand_b":
        k, v = _pairs_random(2, 7, 1000, 20000)
    else:
        k, v = np.array([3, 3, 0], np.uint32), np.array([9, 1, 255], np.uint32)
    ix = ab.Index(k, v, host_only=True)
    ref = oracle.RefIndex(k, v, dedupe=True)
    nk = ix.info["nkeys"]
    for t in range(ab.NTYPES):
        h, p, s = ix.export_type(t)
        rh, rp, _, rs = ref.arrays(t)
        assert np.array_equal(h, rh), f"type {t} headers"
        assert np.array_equal(p, rp), f"type {t} payload"
        # the reference's key_segments_ has max_key+2 entries; ours nkeys+1
        assert rs.size == nk + 1 and np.array_equal(s, rs), f"type {t} segs"


@pytest.mark.parametrize("ts", [8, 9, 12, 13, 16])
def test_value_tiled_layout_decodes_to_pairs(ts):
    k, v = _pairs_random(5, 300, 90000, 60000)
    ix = ab.Index(k, v, 0, 0, ts, host_only=True)
    dk, dv = ix.decode_tiles()
    a = np.unique((k.astype(np.uint64) << 32) | v)
    b = np.sort((dk.astype(np.uint64) << 32) | dv)
    assert np.array_equal(a, b)
    # records come out tile by tile in ascending value tiles
    tiles = dv >> ts
    assert (np.diff(tiles.astype(np.int64)) >= 0).all() or ts >= 16


def test_index_info_counts():
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues, host_only=True)
    inf = ix.info
    ref = oracle.RefIndex(k, v)
    assert inf["nrec"] == [ref.nrec(t) for t in range(6)]
    assert inf["nrec_total"] == sum(inf["nrec"])
    assert inf["list_bytes"] == sum(n * (4 + p) for n, p in zip(inf["nrec"], ab.PAYLOAD_BYTES))
    assert inf["nnz"] == np.unique((k.astype(np.uint64) << 32) | v).size
    assert inf["stride_bytes"] == ((configs.C1.nvalues + 7) // 8 + 15) // 16 * 16
    assert inf["tile_shift"] == 8 and inf["ntiles"] == (configs.C1.nvalues + 255) // 256


def test_build_errors():
    k = np.array([5], np.uint32)
    v = np.array([9], np.uint32)
    with pytest.raises(ab.AdaBoolError) as e:
        ab.Index(k, v, nkeys=3, host_only=True)
    assert e.value.status == 1
    with pytest.raises(ab.AdaBoolError):
        ab.Index(k, v, tile_shift=7, host_only=True)
    with pytest.raises(ab.AdaBoolError) as e:
        ab.Index(np.array([0], np.uint32), np.array([0xFFFFFFFF], np.uint32), host_only=True)
    assert e.value.status == 4
    empty = ab.Index(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 10, 100, host_only=True)
    assert empty.info["nrec_total"] == 0 and empty.info["nkeys"] == 10
    with pytest.raises(ab.AdaBoolError):
        ab.Context(empty)  # host-only handles cannot run queries


def test_key_block_meta_product_matches_oracle():
    ks = [100, 101, 255, 260, 500, 510]
    a = ab.key_block_meta(4, ks)
    assert a[0].tolist() == [100, 500, 510] and a[1].tolist() == [0, 8, 10] and a[2].size == 12
    rng = np.random.default_rng(4)
    for step in (4, 64, 256):
        keys = np.sort(rng.integers(0, 2**24, 3000)).astype(np.uint32)
        keys[:100] = np.sort(rng.integers(0, 50, 100))
        keys = np.sort(keys)
        for x, y in zip(ab.key_block_meta(step, keys), oracle.key_block_meta(step, keys)):
            assert np.array_equal(x, y)


def test_host_features():
    keys = np.array([0, 0, 0, 1, 2, 2], np.uint32)
    vals = np.array([0, 1, 2, 300, 600, 5000], np.uint32)
    ix = ab.Index(keys, vals, 4, 8192, host_only=True)
    # key 0: one ARRAY4 record (3 values); key 1: ARRAY1; key 2: two ARRAY1
    f, m = ix.features_host([0, 2, 3, 9])
    assert m == 3
    assert f[0] == pytest.approx(np.log1p(4))      # 4 query keys
    assert f[1] == pytest.approx(np.log1p(3))      # 3 in range
    assert f[2] == pytest.approx(np.log1p(3))      # M
    assert f[3] == pytest.approx(np.log1p(2))      # max list
    assert f[5] == pytest.approx(np.log1p(8 + 4 + 4))
    assert f[6] == pytest.approx(3 / 4)             # M / R
    assert f[7] == pytest.approx(5 / 8192)
    assert f[9] == pytest.approx(3 / 4)
    assert f[10] == pytest.approx(2 / 3) and f[11] == 0


def test_generators_deterministic_and_shaped():
    a = ab.gen_pairs(77, 3000, 500, lo=5, hi=15, zipf_s=1.0, nthreads=1)
    b = ab.gen_pairs(77, 3000, 500, lo=5, hi=15, zipf_s=1.0, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    k, v = a
    lens = np.bincount(v, minlength=3000)
    assert lens.min() >= 5 and lens.max() <= 15
    # keys distinct within a row
    assert np.unique((v.astype(np.uint64) << 32) | k).size == k.size
    # Zipf: the most frequent key is far above the mean frequency
    cnt = np.bincount(k, minlength=500)
    assert cnt.max() > 10 * cnt.mean()
    ln = ab.gen_pairs(5, 20000, 5000, len_dist=1, p1=16, p2=0.668, lo=1, hi=256, zipf_s=1.1)
    L = np.bincount(ln[1], minlength=20000)
    assert 1 <= L.min() and L.max() <= 256 and 14 <= np.median(L) <= 18
    qk, qo = ab.gen_queries(9, 1000, 50, size_dist=1, p1=40, p2=1.0, lo=8, hi=300, key_dist=1, zipf_s=1.0)
    sizes = np.diff(qo)
    assert sizes.min() >= 8 and sizes.max() <= 300 and qo[-1] == qk.size
    for i in range(50):
        q = qk[qo[i]:qo[i + 1]]
        assert np.unique(q).size == q.size and q.max() < 1000
    pk, po = ab.gen_queries(9, 100000, 2000, size_dist=2, p1=1.5, lo=1, hi=50000, key_dist=1, zipf_s=1.1)
    ps = np.diff(po)
    assert ps.min() >= 1 and ps.max() <= 50000 and np.median(ps) <= 4


def test_text_loaders_match_oracle():
    from tests.test_oracle import COO_TEXT, QUERY_TEXT

    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "filter.coo")
        open(p, "w").write(COO_TEXT)
        k, v, mk, mv = ab.read_coo(p)
        ok, ov, omk, omv = oracle.read_coo(p)
        assert np.array_equal(k, ok) and np.array_equal(v, ov) and (mk, mv) == (omk, omv)
        q = os.path.join(d, "query_vec_3.coo")
        open(q, "w").write(QUERY_TEXT)
        assert np.array_equal(ab.read_query(q), oracle.read_query(q))
        with pytest.raises(ab.AdaBoolError):
            ab.read_query(os.path.join(d, "missing.coo"))


def test_find_query_files_order_and_filter(capsys):
    """FindQueryFiles (test/adsfilter_q_all.cu:211-231): whole-name match of
    query_vec_<digits>.coo over regular files (symlinks followed), sorted as
    strings, so _10 runs before _2.  Known answer written out by hand."""
    with tempfile.TemporaryDirectory() as d:
        for n in ["query_vec_2.coo", "query_vec_10.coo", "query_vec_1.coo", "query_vec_007.coo",
                  "query_vec_.coo", "query_vec_3.coo.bak", "xquery_vec_4.coo", "query_vec_5.COO",
                  "query_vec_1a.coo", "query_vec_-1.coo", "filter.coo"]:
            open(os.path.join(d, n), "w").write("1\n")
        os.mkdir(os.path.join(d, "query_vec_6.coo"))                       # a directory: skipped
        os.symlink(os.path.join(d, "filter.coo"), os.path.join(d, "query_vec_7.coo"))  # to a file: kept
        os.symlink(os.path.join(d, "nowhere"), os.path.join(d, "query_vec_8.coo"))     # dangling: skipped
        want = [os.path.join(d, n) for n in
                ["query_vec_007.coo", "query_vec_1.coo", "query_vec_10.coo", "query_vec_2.coo", "query_vec_7.coo"]]
        assert ab.find_query_files(d) == want
        assert ab.find_query_files(d + "/") == want  # no doubled separator
        assert ab.find_query_files(os.path.join(d, "missing")) == []
        assert "Error reading directory" in capsys.readouterr().err
        empty = os.path.join(d, "empty")
        os.mkdir(empty)
        assert ab.find_query_files(empty) == []
        # the C ABI's sizing contract: nothing written into a short buffer
        lib = C.CDLL(ab.lib_path())
        lib.ab_find_query_files.restype = C.c_int64
        lib.ab_find_query_files.argtypes = [C.c_char_p, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]
        need = C.c_uint64()
        buf = C.create_string_buffer(b"\x7f" * 16, 16)
        assert lib.ab_find_query_files(d.encode(), buf, 16, C.byref(need)) == 5
        assert need.value == sum(len(p) + 1 for p in want) and buf.raw == b"\x7f" * 16


def test_large_coo_parallel_parse_matches_oracle():
    """A text file above the parallel-parse size (chunks cut at line starts)
    with comments, blank lines, CRLF, tabs and bad lines spread through it."""
    from tests.test_oracle import COO_TEXT

    rng = np.random.default_rng(8)
    k = rng.integers(0, 2**32, size=400_000, dtype=np.uint64)
    v = rng.integers(0, 5_000_000, size=400_000, dtype=np.uint64)
    lines = [f"{a} {b}" for a, b in zip(k.tolist(), v.tolist())]
    for i in range(0, len(lines), 997):
        lines[i] = lines[i] + "\r"
    for i in range(5, len(lines), 1009):
        lines[i] = "# comment " + lines[i]
    for i in range(7, len(lines), 1013):
        lines[i] = "\t" + lines[i].replace(" ", "\t")
    text = "%%MatrixMarket\n" + "\n\n".join(lines[:50]) + "\n" + "\n".join(lines[50:]) + "\n" + COO_TEXT
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "big.coo")
        open(p, "w").write(text)
        assert os.path.getsize(p) > (1 << 22)
        got = ab.read_coo(p)
        want = oracle.read_coo(p)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert got[2:] == want[2:]
        assert got[0].size > 390_000


def test_stats_match_oracle():
    rng = np.random.default_rng(2)
    for n in (1, 2, 9, 1000):
        d = rng.lognormal(size=n)
        a, b = ab.stats(d), oracle.stats(d)
        for key in a:
            assert a[key] == pytest.approx(b[key], rel=1e-12)


def test_id_bound_and_argument_errors_on_host_index():
    """The id upper bound from the host copy of the key prefixes: the sum of
    the in-range keys' list lengths (repeats counted), capped at nvalues; and
    the id-only entry points refuse bad arguments before touching a device."""
    rng = np.random.default_rng(21)
    keys = rng.integers(0, 40, 3000).astype(np.uint32)
    vals = rng.integers(0, 5000, 3000).astype(np.uint32)
    ix = ab.Index(keys, vals, 40, 5000, host_only=True)
    pairs = np.unique((keys.astype(np.uint64) << 32) | vals)
    per_key = np.bincount((pairs >> 32).astype(np.int64), minlength=40)
    for q in ([], [3], [3, 3, 7], [39, 40, 10**6], list(range(40))):
        want = min(int(sum(per_key[k] for k in q if k < 40)), 5000)
        assert ix.id_bound(np.array(q, np.uint32)) == want, q
    with pytest.raises(ab.AdaBoolError):
        ab._check(ab._lib.ab_query_batch_ids(None, None, None, 0, None, None, None))
    with pytest.raises(ab.AdaBoolError):
        ab._check(ab._lib.ab_ctx_set_id_budget(None, 1 << 20))
    with pytest.raises(ab.AdaBoolError):
        ab._check(ab._lib.ab_query_id_bound(None, None, 0, None))
