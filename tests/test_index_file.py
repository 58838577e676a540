"""Binary index file: save -> load reproduces every array of both layouts
(checked through the exports and the full tile decode), and damaged files
are refused.  Host-only handles: runs without a GPU."""

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs


def _filter():
    k, v = ab.gen_pairs(11, 20000, 3000, len_dist=1, p1=8.0, p2=0.7, lo=1, hi=200, zipf_s=1.1)
    return k, v


def _arrays(ix):
    out = []
    for t in range(ab.NTYPES):
        out.extend(ix.export_type(t))
    out.extend(ix.decode_tiles())
    return out


def test_save_load_roundtrip(tmp_path):
    k, v = _filter()
    ix = ab.Index(k, v, 3000, 20000, host_only=True)
    p = str(tmp_path / "c.abidx")
    ix.save(p)
    iy = ab.Index.load(p, host_only=True)
    assert iy.info == ix.info
    for a, b in zip(_arrays(ix), _arrays(iy)):
        assert np.array_equal(a, b)
    # and the decoded pairs are the input set
    dk, dv = iy.decode_tiles()
    got = np.unique(dk.astype(np.uint64) << 32 | dv)
    want = np.unique(k.astype(np.uint64) << 32 | v)
    assert np.array_equal(got, want)
    f, m = iy.features_host(np.array([0, 5, 17, 2999, 4000], np.uint32))
    f0, m0 = ix.features_host(np.array([0, 5, 17, 2999, 4000], np.uint32))
    assert m == m0 and np.array_equal(f, f0)


@pytest.mark.parametrize("damage", ["truncate", "flip", "magic", "empty"])
def test_damaged_files_are_refused(tmp_path, damage):
    k, v = _filter()
    ix = ab.Index(k, v, 3000, 20000, host_only=True)
    p = str(tmp_path / "c.abidx")
    ix.save(p)
    data = bytearray(open(p, "rb").read())
    if damage == "truncate":
        data = data[: len(data) // 2]
    elif damage == "flip":
        data[len(data) // 3] ^= 0x10
    elif damage == "magic":
        data[0] ^= 0xFF
    else:
        data = bytearray()
    open(p, "wb").write(bytes(data))
    with pytest.raises(ab.AdaBoolError) as e:
        ab.Index.load(p, host_only=True)
    assert e.value.status == 5  # AB_IO_ERROR


# ---- malformed files with a valid checksum --------------------------------
# File layout (ab_index_save): a 40-byte header, then per section a u64 byte
# count, the bytes and zero padding to 8, then the 64-bit FNV-1a checksum of
# everything before it.  Sections: seg8, then per type (hdr, pay, vhdr, vpay),
# then vblk, vkeys, tile_blk.
_NSEC = 1 + 4 * ab.NTYPES + 3


def _fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _sections(data: bytes):
    out, pos = [], 40
    for _ in range(_NSEC):
        n = int(np.frombuffer(data, np.uint64, 1, pos)[0])
        out.append((pos + 8, n))
        pos += 8 + n + (8 - n % 8) % 8
    return out


def _resave(path, data: bytearray):
    body = bytes(data[:-8])
    open(path, "wb").write(body + np.uint64(_fnv1a(body)).tobytes())


def _sec_u32(data, secs, i):
    off, n = secs[i]
    return np.frombuffer(data, np.uint32, n // 4, off), off


@pytest.mark.parametrize("damage", ["km_header_value", "km_payload_value", "seg8_order", "vblk_records",
                                    "vblk_key_bytes", "vt_header_tile", "vt_key_range", "tile_blk_order",
                                    "bitset_align"])
def test_malformed_file_with_valid_checksum_is_refused(tmp_path, damage):
    """Checksums detect accidents only: every invariant the kernels rely on is
    checked at load, so a file that breaks one is refused even when its
    checksum was recomputed (values, keys, record and key-stream offsets, tile
    membership, prefix order)."""
    nvalues, nkeys = 20000, 3000  # 20000 = 78 * 256 + 32: the last window is partial
    k, v = _filter()
    ix = ab.Index(k, v, nkeys, nvalues, host_only=True)
    p = str(tmp_path / "m.abidx")
    ix.save(p)
    data = bytearray(open(p, "rb").read())
    secs = _sections(bytes(data))
    # pristine re-save loads (the checksum writer here matches the library's)
    _resave(p, data)
    ab.Index.load(p, host_only=True)

    def put_u32(off, val):
        data[off:off + 4] = np.uint32(val).tobytes()

    hdr = lambda t: 1 + 4 * t      # noqa: E731
    pay = lambda t: 2 + 4 * t      # noqa: E731
    vhdr = lambda t: 3 + 4 * t     # noqa: E731
    vblk, vkeys, tblk = 1 + 4 * ab.NTYPES, 2 + 4 * ab.NTYPES, 3 + 4 * ab.NTYPES
    if damage == "km_header_value":
        _, off = _sec_u32(bytes(data), secs, hdr(0))
        put_u32(off, nvalues + 5)
    elif damage == "km_payload_value":
        # an ARRAY4 record moved into the partial last window, its first
        # payload byte past the end of the value space
        _, off = _sec_u32(bytes(data), secs, hdr(1))
        put_u32(off, 78 * 256 + 1)
        poff, _ = secs[pay(1)]
        data[poff] = 250
    elif damage == "seg8_order":
        a, off = _sec_u32(bytes(data), secs, 0)
        row = int(np.argmax(a.reshape(-1, 8)[:, 7] > 0))
        put_u32(off + (row * 8 + 7) * 4, 0xFFFFFFF0)  # the next row's value prefix is then smaller
    elif damage == "vblk_records":
        off, _ = secs[vblk]
        put_u32(off + 12, 0x7FFFFFFF)          # rec_base of block 0
    elif damage == "vblk_key_bytes":
        off, _ = secs[vblk]
        put_u32(off + 4, 0x00FFFFFF)           # off_kb of block 0
    elif damage == "vt_header_tile":
        t = int(np.argmax([len(_sec_u32(bytes(data), secs, vhdr(t))[0]) > 0 for t in range(ab.NTYPES)]))
        a, off = _sec_u32(bytes(data), secs, vhdr(t))
        # a record of tile 0 given a header in the last tile (still < nvalues)
        put_u32(off, (int(a[0]) & 0xFF) | (nvalues - 256) & 0xFFFFFF00 if t != 5 else (nvalues - 256) & 0xFFFFFF00)
    elif damage == "vt_key_range":
        off, _ = secs[vblk]
        put_u32(off, nkeys)                    # start_key of block 0: every key >= nkeys
    elif damage == "tile_blk_order":
        a, off = _sec_u32(bytes(data), secs, tblk)
        put_u32(off + 4, int(a[-1]) + 1)
    else:
        a, off = _sec_u32(bytes(data), secs, hdr(5))
        assert a.size > 0
        put_u32(off, int(a[0]) | 3)
    _resave(p, data)
    with pytest.raises(ab.AdaBoolError) as e:
        ab.Index.load(p, host_only=True)
    assert e.value.status == 5, str(e.value)  # AB_IO_ERROR


def test_missing_file():
    with pytest.raises(ab.AdaBoolError):
        ab.Index.load("/nonexistent/index.abidx", host_only=True)


@pytest.mark.gpu
def test_device_index_save_load_queries(tmp_path):
    """A device-resident index saved (arrays copied back from HBM) and loaded
    again answers a mixed batch exactly as the oracle does."""
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    p = str(tmp_path / "c1.abidx")
    ix.save(p)
    iy = ab.Index.load(p)
    assert iy.info["nrec_total"] == ix.info["nrec_total"] and iy.info["key_stream_bytes"] == ix.info["key_stream_bytes"]
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    qk, qo = ab.gen_queries(0xADAB0601, configs.C1.nkeys, 32, size_dist=1, p1=60.0, p2=1.3, lo=0, hi=3000,
                            key_dist=1, zipf_s=1.0, perm_seed=configs.C1.seed)
    ctx = ab.Context(iy)
    ctx.set_batch_threshold(0.01)
    bms, ids, counts, paths = ctx.query_batch(qk, qo, ab.AB_PATH_THRESHOLD)
    for i in range(32):
        wb, wi = csr.union(qk[qo[i]:qo[i + 1]])
        assert np.array_equal(bms[i], wb) and np.array_equal(ids[i], wi), f"q{i}"
    ctx.close()


def _edge_pairs():
    from .test_gpu_parity import _edge_filter

    k, v, nvalues = _edge_filter()
    return k, v, 12, nvalues


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1", "edge", "zipf", "empty", "ts8", "ts13"])
def test_device_build_is_byte_identical(tmp_path, case):
    """ab_index_build_device writes exactly the host build's layout: the two
    index files are byte-identical."""
    ts = 0
    if case in ("c1", "ts8", "ts13"):
        k, v = configs.make_filter(configs.C1)
        nk, nv = configs.C1.nkeys, configs.C1.nvalues
        ts = {"c1": 0, "ts8": 8, "ts13": 13}[case]
    elif case == "edge":
        k, v, nk, nv = _edge_pairs()
    elif case == "zipf":
        # duplicates, heavy keys (BITSET windows) and a sparse tail
        k, v = ab.gen_pairs(17, 300_000, 20_000, len_dist=1, p1=16.0, p2=0.668, lo=1, hi=256, zipf_s=1.1)
        k = np.concatenate([k, k[:5000]])
        v = np.concatenate([v, v[:5000]])
        nk, nv = 20_000, 300_000
    else:
        k = np.zeros(0, np.uint32)
        v = np.zeros(0, np.uint32)
        nk, nv = 10, 1000
    print(f"device build {case}", flush=True)
    host = ab.Index(k, v, nk, nv, ts, host_only=True)
    dev = ab.Index.build_device(k, v, nk, nv, ts)
    hi, di = dict(host.info), dict(dev.info)
    hi.pop("device_bytes")
    di.pop("device_bytes")
    assert di == hi
    ph, pd = str(tmp_path / "h.abidx"), str(tmp_path / "d.abidx")
    host.save(ph)
    dev.save(pd)
    assert open(ph, "rb").read() == open(pd, "rb").read()


@pytest.mark.gpu
@pytest.mark.parametrize("n,where", [(1, "first"), (1000, "last"), (3_000_001, "first"), (3_000_001, "last"),
                                     (3_000_001, "middle")])
def test_device_build_infers_spaces(n, where):
    """With nkeys = nvalues = 0 the device build takes max+1 of the pairs, as
    the host build does. The sizes cover the strided loop's unrolled body and
    its tail; the largest key and value sit at the first, last or a middle
    pair, in different places for keys and values."""
    rng = np.random.default_rng(n + len(where))
    k = rng.integers(0, 5000, n).astype(np.uint32)
    v = rng.integers(0, 90_000, n).astype(np.uint32)
    pos = {"first": 0, "last": n - 1, "middle": n // 2 + 7}[where]
    k[pos] = 77_777
    v[(pos + n // 3) % n] = 1_234_567
    print(f"device build infers spaces n={n} {where}", flush=True)
    host = ab.Index(k, v, 0, 0, host_only=True)
    dev = ab.Index.build_device(k, v, 0, 0)
    assert dev.info["nkeys"] == host.info["nkeys"] == int(k.max()) + 1
    assert dev.info["nvalues"] == host.info["nvalues"] == int(v.max()) + 1
