"""CPU oracle of the AdaBool OR-filtering path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker / reported CPU baseline -- never as the product path.

What pins each oracle function (besides the reference citation in oracle.c):
see the header comment of ``oracle/oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_VP = C.c_void_p

NTYPES = 6
PAYLOAD_BYTES = (0, 4, 8, 16, 32, 32)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    p = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(p):
        build()
    L = C.CDLL(p)
    sig = {
        "or_csr_build": (C.c_int, [_VP, _VP, C.c_uint64, C.c_uint32, _VP, _VP]),
        "or_union_csr": (C.c_uint64, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, C.c_uint32, _VP, _VP]),
        "or_union_ids": (C.c_uint64, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP]),
        "or_scan_coo": (C.c_uint64, [_VP, _VP, C.c_uint64, _VP, C.c_uint32, C.c_uint32, _VP]),
        "or_collect_ids": (C.c_uint64, [_VP, C.c_uint64, C.c_uint32, _VP]),
        "or_ref_index_build": (_VP, [_VP, _VP, C.c_uint64, C.c_int]),
        "or_ref_index_free": (None, [_VP]),
        "or_ref_index_nrec": (C.c_uint64, [_VP, C.c_int]),
        "or_ref_index_headers": (_VP, [_VP, C.c_int]),
        "or_ref_index_payload": (_VP, [_VP, C.c_int]),
        "or_ref_index_keys": (_VP, [_VP, C.c_int]),
        "or_ref_index_segs": (_VP, [_VP, C.c_int, C.POINTER(C.c_uint64)]),
        "or_payload_bytes": (C.c_int, [C.c_int]),
        "or_key_block_meta": (C.c_uint64, [C.c_uint32, C.c_uint64, _VP, _VP, _VP, _VP]),
        "or_ref_index_decode": (C.c_uint64, [_VP, _VP, _VP]),
        "or_emit": (None, [C.c_int, C.c_uint32, _VP, _VP]),
        "or_dense_probe": (None, [_VP, C.c_uint32, _VP, C.c_uint64, _VP]),
        "or_sparse_union": (C.c_uint64, [_VP, _VP, C.c_uint32, C.c_uint32, _VP]),
        "or_union_batch": (C.c_double, [_VP, _VP, C.c_uint32, _VP, _VP, C.c_uint32, C.c_uint32, C.c_int, _VP,
                                        _VP, C.POINTER(C.c_uint64)]),
        "or_union_batch_digest": (None, [_VP, _VP, C.c_uint32, _VP, _VP, C.c_uint32, C.c_uint32, C.c_int, _VP,
                                         _VP, _VP]),
        "or_bitmap_digest": (None, [_VP, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, _VP]),
        "or_ids_digest": (None, [_VP, _VP, _VP, C.c_uint64, C.c_int, _VP, _VP]),
        "or_stats": (None, [_VP, C.c_uint64, _VP]),
        "or_read_coo": (C.c_int64, [C.c_char_p, _VP, _VP, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "or_read_query": (C.c_int64, [C.c_char_p, _VP]),
        "or_mlp_forward": (None, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP, C.c_int, _VP]),
        "or_mlp_loss_grad": (C.c_double, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, _VP, _VP, C.c_int, _VP]),
        "or_adam_step": (None, [_VP, _VP, _VP, _VP, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def stride_words(nvalues: int) -> int:
    b = (nvalues + 7) // 8
    return ((b + 15) // 16 * 16) // 4


class CSR:
    """CSR-by-key copy of the filter, the oracle's fast predicate."""

    def __init__(self, keys, vals, nkeys: int, nvalues: int):
        k, v = _u32(keys), _u32(vals)
        self.nkeys, self.nvalues = int(nkeys), int(nvalues)
        self.off = np.zeros(self.nkeys + 1, np.uint64)
        self.vals = np.zeros(k.size, np.uint32)
        if lib().or_csr_build(_p(k), _p(v), k.size, self.nkeys, _p(self.off), _p(self.vals)) != 0:
            raise ValueError("key >= nkeys")

    def union(self, q) -> Tuple[np.ndarray, np.ndarray]:
        """(bitmap words [stride], ascending ids) of query q."""
        qa = _u32(q)
        bm = np.zeros(stride_words(self.nvalues), np.uint32)
        ids = np.zeros(max(self.nvalues, 1), np.uint32)
        n = lib().or_union_csr(_p(self.off), _p(self.vals), self.nkeys, _p(qa), qa.size, self.nvalues, _p(bm),
                               _p(ids))
        return bm, ids[:n].copy()

    def union_ids(self, q) -> np.ndarray:
        """Ascending ids of query q, by gather + sort + dedupe (no bitmap)."""
        qa = _u32(q)
        room = sum(int(self.off[k + 1] - self.off[k]) for k in qa.tolist() if k < self.nkeys)
        ids = np.zeros(max(room, 1), np.uint32)
        n = lib().or_union_ids(_p(self.off), _p(self.vals), self.nkeys, _p(qa), qa.size, _p(ids))
        return ids[:n].copy()

    def union_batch(self, qkeys, qoff, nthreads: int = 1):
        """Multi-threaded union of many queries: (wall seconds, counts, latency ms, checksum)."""
        k = _u32(qkeys)
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        nq = off.size - 1
        counts = np.zeros(max(nq, 1), np.uint64)
        lat = np.zeros(max(nq, 1), np.float64)
        cs = C.c_uint64()
        t = lib().or_union_batch(_p(self.off), _p(self.vals), self.nkeys, _p(k), _p(off), nq, self.nvalues,
                                 nthreads, _p(counts), _p(lat), C.byref(cs))
        return t, counts[:nq], lat[:nq], cs.value


    def union_batch_digest(self, qkeys, qoff, nthreads: int = 1):
        """Per query: (count, bitmap digest, id digest) of the oracle's union
        (definitions in oracle.h)."""
        k = _u32(qkeys)
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        nq = off.size - 1
        c = np.zeros(max(nq, 1), np.uint64)
        db = np.zeros(max(nq, 1), np.uint64)
        di = np.zeros(max(nq, 1), np.uint64)
        lib().or_union_batch_digest(_p(self.off), _p(self.vals), self.nkeys, _p(k), _p(off), nq, self.nvalues,
                                    nthreads, _p(c), _p(db), _p(di))
        return c[:nq], db[:nq], di[:nq]


def bitmap_digest(bitmaps, nwords: int = 0, nthreads: int = 1) -> np.ndarray:
    """Digest of each row of a [nq, stride] u32 array (its first nwords words,
    default all)."""
    b = np.ascontiguousarray(np.asarray(bitmaps).view(np.uint32))
    if b.ndim == 1:
        b = b.reshape(1, -1)
    out = np.zeros(max(b.shape[0], 1), np.uint64)
    lib().or_bitmap_digest(_p(b), b.shape[1], nwords or b.shape[1], b.shape[0], nthreads, _p(out))
    return out[:b.shape[0]]


def ids_digest(ids, off, counts, nthreads: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """(digest, strictly ascending) of each id list ids[off[q] : off[q] + counts[q]]."""
    a = np.ascontiguousarray(np.asarray(ids).view(np.uint32))
    o = np.ascontiguousarray(np.asarray(off, dtype=np.uint64))
    c = np.ascontiguousarray(np.asarray(counts, dtype=np.uint32))
    if o.size and int((o + c).max()) > a.size:
        raise ValueError("an id range lies outside the array")
    out = np.zeros(max(o.size, 1), np.uint64)
    asc = np.zeros(max(o.size, 1), np.uint8)
    lib().or_ids_digest(_p(a), _p(o), _p(c), o.size, nthreads, _p(out), _p(asc))
    return out[:o.size], asc[:o.size].astype(bool)


def scan_coo(keys, vals, q, nvalues: int) -> np.ndarray:
    k, v, qa = _u32(keys), _u32(vals), _u32(q)
    bm = np.zeros(stride_words(nvalues), np.uint32)
    lib().or_scan_coo(_p(k), _p(v), k.size, _p(qa), qa.size, nvalues, _p(bm))
    return bm


def collect_ids(bitmap, nbits: int, limit: int) -> np.ndarray:
    b = _u32(bitmap)
    ids = np.zeros(max(limit, 1), np.uint32)
    n = lib().or_collect_ids(_p(b), nbits, limit, _p(ids))
    return ids[:n].copy()


class RefIndex:
    """Restatement of ivf_list's build (6 record types, key_segments_)."""

    def __init__(self, keys, vals, dedupe: bool = True):
        k, v = _u32(keys), _u32(vals)
        self._h = lib().or_ref_index_build(_p(k), _p(v), k.size, 1 if dedupe else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_ref_index_free(self._h)
            self._h = None

    def nrec(self, t: int) -> int:
        return int(lib().or_ref_index_nrec(self._h, t))

    def arrays(self, t: int):
        L = lib()
        n = self.nrec(t)
        pb = PAYLOAD_BYTES[t]

        def view(ptr, ctype, cnt, dt):
            if not cnt:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array((ctype * cnt).from_address(ptr)).copy()

        hdr = view(L.or_ref_index_headers(self._h, t), C.c_uint32, n, np.uint32)
        keys = view(L.or_ref_index_keys(self._h, t), C.c_uint32, n, np.uint32)
        pay = view(L.or_ref_index_payload(self._h, t), C.c_uint8, n * pb, np.uint8)
        sl = C.c_uint64()
        sp = L.or_ref_index_segs(self._h, t, C.byref(sl))
        segs = view(sp, C.c_uint32, sl.value, np.uint32)
        return hdr, pay, keys, segs

    def decode(self) -> Tuple[np.ndarray, np.ndarray]:
        L = lib()
        n = L.or_ref_index_decode(self._h, None, None)
        k = np.zeros(n, np.uint32)
        v = np.zeros(n, np.uint32)
        L.or_ref_index_decode(self._h, _p(k), _p(v))
        return k, v

    def dense_probe(self, qbitmap_words, nq_bits: int, nvalues: int, step: int = 256) -> np.ndarray:
        qb = _u32(qbitmap_words)
        bm = np.zeros(stride_words(nvalues), np.uint32)
        lib().or_dense_probe(self._h, step, _p(qb), nq_bits, _p(bm))
        return bm

    def sparse_union(self, q, nvalues: int, step: int = 256) -> Tuple[np.ndarray, int]:
        qa = _u32(q)
        bm = np.zeros(stride_words(nvalues), np.uint32)
        units = lib().or_sparse_union(self._h, _p(qa), qa.size, step, _p(bm))
        return bm, int(units)


def key_block_meta(step: int, keys):
    k = _u32(keys)
    n = k.size
    if n == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint8)
    nb = (n + step - 1) // step
    nbytes = lib().or_key_block_meta(step, n, _p(k), None, None, None)
    starts = np.zeros(nb + 1, np.uint32)
    pos = np.zeros(nb + 1, np.uint32)
    buf = np.zeros(nbytes, np.uint8)
    lib().or_key_block_meta(step, n, _p(k), _p(starts), _p(pos), _p(buf))
    return starts, pos, buf


def emit(t: int, header: int, payload, nvalues: int) -> np.ndarray:
    p = np.zeros(32, np.uint8)
    pl = np.asarray(payload, dtype=np.uint8)
    p[: pl.size] = pl
    bm = np.zeros(stride_words(nvalues), np.uint32)
    lib().or_emit(t, header, _p(p), _p(bm))
    return bm


def stats(samples):
    d = np.ascontiguousarray(np.asarray(samples, dtype=np.float64)).copy()
    o = np.zeros(9, np.float64)
    lib().or_stats(_p(d), d.size, _p(o))
    return dict(zip(("mean", "median", "stddev", "min", "max", "p50", "p90", "p95", "p99"), o.tolist()))


def read_coo(path: str):
    mr, mc = C.c_uint32(), C.c_uint32()
    n = lib().or_read_coo(path.encode(), None, None, C.byref(mr), C.byref(mc))
    if n < 0:
        raise OSError(path)
    k = np.zeros(n, np.uint32)
    v = np.zeros(n, np.uint32)
    lib().or_read_coo(path.encode(), _p(k), _p(v), C.byref(mr), C.byref(mc))
    return k, v, mr.value, mc.value


def read_query(path: str):
    n = lib().or_read_query(path.encode(), None)
    if n < 0:
        raise OSError(path)
    k = np.zeros(n, np.uint32)
    lib().or_read_query(path.encode(), _p(k))
    return k


# ---- MLP (float64 reference) ----------------------------------------------
MLP_DIMS = (12, 64, 64, 2)


def mlp_nparams(F=12, H1=64, H2=64, O=2) -> int:
    return H1 * F + H1 + H2 * H1 + H2 + O * H2 + O


def mlp_forward(params, x, dims=MLP_DIMS):
    F, H1, H2, O = dims
    p = np.ascontiguousarray(params, dtype=np.float64)
    xa = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, F)
    y = np.zeros((xa.shape[0], O), np.float64)
    lib().or_mlp_forward(_p(p), F, H1, H2, O, _p(xa), xa.shape[0], _p(y))
    return y


def mlp_loss_grad(params, x, y, dims=MLP_DIMS):
    F, H1, H2, O = dims
    p = np.ascontiguousarray(params, dtype=np.float64)
    xa = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, F)
    ya = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, O)
    g = np.zeros_like(p)
    loss = lib().or_mlp_loss_grad(_p(p), F, H1, H2, O, _p(xa), _p(ya), xa.shape[0], _p(g))
    return loss, g


def adam_step(params, m, v, grad, lr, b1, b2, eps, step):
    lib().or_adam_step(_p(params), _p(m), _p(v), _p(np.ascontiguousarray(grad, dtype=np.float64)), params.size,
                       lr, b1, b2, eps, step)
