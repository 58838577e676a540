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
