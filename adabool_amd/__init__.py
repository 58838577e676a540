"""adabool_amd -- MI355X-native AdaBool Boolean OR-filtering engine.

Python host mirror of the reference's filter/query API over the C ABI in
``include/adabool.h`` (library ``adabool_amd/libadabool.so``, built by
``__graft_entry__.build()`` or ``make -C adabool_amd/csrc``).

The reference is a header-only C++ template library; its entry points map to
this module as follows (reference paths relative to its repository root):

====================================================  ==================================
reference                                             here
====================================================  ==================================
``ivf_list::BuildFrom`` (src/AdaBool/inverted_list.h:348)      ``IvfList.BuildFrom``
``ivf_list::Prepare(keys)`` (inverted_list.h:481)              ``IvfList.Prepare`` -> ``SparseWorker``
``ivf_list::Prepare(len, bitmap)`` (inverted_list.h:530)       ``IvfList.PrepareBitmap`` -> ``DenseWorker``
``ivf_list::PrepareBatch`` (inverted_list.h:564)               ``IvfList.PrepareBatch``
``ivf_list::PrepareBatchDevice`` (inverted_list.h:591)         ``IvfList.PrepareBatchDevice``
``ListUnion_sparse`` (test/adsfilter_q_all.cu:64)             ``SparseWorker.run``
``ListUnion_dense`` (test/adsfilter_q_all.cu:26)              ``DenseWorker.run``
``RunArrays`` (src/AdaBool/inverted_list.cu.h:224)             ``BatchWorker.RunArrays``
``BatchListUnion_sparse`` (test/adsfilter_batch.cu:39)        ``IvfList.BatchListUnionSparse``
``ReadCOO`` / ``ReadQuery`` (test/adsfilter_q_all.cu:127,164)  ``read_coo`` / ``read_query``
``CalculateStats`` (test/adsfilter_q_all.cu:258)              ``stats``
====================================================  ==================================

Differences from the reference, by design: results are always written in
full (the caller needs no memset), duplicate input pairs are removed at build
time, batched dense accepts any batch size, and keys >= nkeys are ignored on
every path.  There is no CPU fallback: without the compiled library every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "AB_PATH_AUTO", "AB_PATH_DENSE", "AB_PATH_SPARSE", "AB_PATH_THRESHOLD", "AB_PATH_LEARNED", "AB_PATH_IDS",
    "NTYPES", "PAYLOAD_BYTES", "NFEATURES", "AdaBoolError", "lib_path", "load_library",
    "Index", "Context", "IvfList", "gen_pairs", "gen_queries", "read_coo", "read_query",
    "stats", "key_block_meta", "find_query_files",
]

AB_PATH_AUTO, AB_PATH_DENSE, AB_PATH_SPARSE, AB_PATH_THRESHOLD, AB_PATH_LEARNED, AB_PATH_IDS = range(6)
MODEL_SINGLE, MODEL_BATCH = 0, 1
NTYPES = 6
PAYLOAD_BYTES = (0, 4, 8, 16, 32, 32)
NFEATURES = 12
_STATUS = {0: "AB_OK", 1: "AB_INVALID_ARG", 2: "AB_OOM", 3: "AB_DEVICE_ERROR",
           4: "AB_UNSUPPORTED", 5: "AB_IO_ERROR"}

_HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path() -> str:
    """The in-tree library, or the build named by ADABOOL_LIB (for comparing builds)."""
    return os.environ.get("ADABOOL_LIB") or os.path.join(_HERE, "libadabool.so")


class AdaBoolError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class _IndexInfo(C.Structure):
    _fields_ = [
        ("nkeys", C.c_uint32), ("nvalues", C.c_uint32), ("nnz", C.c_uint64),
        ("nrec", C.c_uint64 * NTYPES), ("nrec_total", C.c_uint64), ("list_bytes", C.c_uint64),
        ("tile_shift", C.c_uint32), ("ntiles", C.c_uint32), ("nblocks", C.c_uint64),
        ("key_stream_bytes", C.c_uint64), ("device_bytes", C.c_uint64), ("stride_bytes", C.c_uint64),
    ]


_lib = None

_VP = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_U64P = C.POINTER(C.c_uint64)
_SIGS = {
    "ab_last_error": (C.c_char_p, []),
    "ab_status_str": (C.c_char_p, [C.c_int]),
    "ab_index_build": (C.c_int, [_VP, _VP, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_VP)]),
    "ab_index_build_host": (C.c_int, [_VP, _VP, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_VP)]),
    "ab_index_load_coo": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_VP)]),
    "ab_index_build_device": (C.c_int, [_VP, _VP, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _VP,
                                        C.POINTER(_VP)]),
    "ab_index_free": (C.c_int, [_VP]),
    "ab_index_save": (C.c_int, [_VP, C.c_char_p]),
    "ab_index_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_VP)]),
    "ab_index_get_info": (C.c_int, [_VP, C.POINTER(_IndexInfo)]),
    "ab_result_stride_bytes": (C.c_uint64, [_VP]),
    "ab_index_export_type": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "ab_index_decode_tiles": (C.c_int64, [_VP, _VP, _VP, C.c_uint64]),
    "ab_key_block_meta": (C.c_uint64, [C.c_uint32, C.c_uint64, _VP, _VP, _VP, _VP]),
    "ab_ctx_create": (C.c_int, [_VP, _VP, C.POINTER(_VP)]),
    "ab_ctx_free": (C.c_int, [_VP]),
    "ab_ctx_sync": (C.c_int, [_VP]),
    "ab_ctx_stream": (_VP, [_VP]),
    "ab_ctx_set_threshold": (C.c_int, [_VP, C.c_double]),
    "ab_ctx_get_threshold": (C.c_double, [_VP]),
    "ab_ctx_set_batch_threshold": (C.c_int, [_VP, C.c_double]),
    "ab_ctx_get_batch_threshold": (C.c_double, [_VP]),
    "ab_ctx_set_dense_group": (C.c_int, [_VP, C.c_uint32]),
    "ab_ctx_set_dense_threads": (C.c_int, [_VP, C.c_uint32]),
    "ab_ctx_set_sparse_chunk": (C.c_int, [_VP, C.c_uint32]),
    "ab_ctx_set_dense_columns": (C.c_int, [_VP, C.c_int]),
    "ab_ctx_set_batch_mode": (C.c_int, [_VP, C.c_int]),
    "ab_ctx_set_timing": (C.c_int, [_VP, C.c_int]),
    "ab_ctx_timing": (C.c_int, [_VP, _VP, _VP, C.c_int]),
    "ab_query": (C.c_int, [_VP, _VP, C.c_uint32, C.c_int, _VP, _VP, _VP, C.POINTER(C.c_int)]),
    "ab_list_union_sparse": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint64]),
    "ab_list_union_dense": (C.c_int, [_VP, C.c_uint64, _VP, _VP, C.c_uint64]),
    "ab_list_union_sparse_noalloc": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint64]),
    "ab_list_union_dense_batch": (C.c_int, [_VP, C.c_uint32, C.c_uint64, _VP, _VP, C.c_uint32]),
    "ab_list_union_sparse_batch": (C.c_int, [_VP, _VP, _VP, C.c_uint32, C.c_uint32, _VP, C.c_uint32]),
    "ab_query_batch": (C.c_int, [_VP, _VP, _VP, C.c_uint32, C.c_int, _VP, _VP, C.c_uint64, _VP, _VP]),
    "ab_compact_ids": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint64, _VP]),
    "ab_query_batch_packed": (C.c_int, [_VP, _VP, _VP, C.c_uint32, C.c_int, _VP, _VP, _VP, _VP, _VP]),
    "ab_query_batch_ids": (C.c_int, [_VP, _VP, _VP, C.c_uint32, _VP, _VP, _VP]),
    "ab_query_id_bound": (C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    "ab_ctx_set_id_budget": (C.c_int, [_VP, C.c_uint64]),
    "ab_ctx_set_id_method": (C.c_int, [_VP, C.c_int]),
    "ab_ctx_set_batch_tuning": (C.c_int, [_VP, _VP, C.c_uint32]),
    "ab_ctx_get_batch_settings": (C.c_int, [_VP, _VP]),
    "ab_ctx_autotune": (C.c_int, [_VP, _VP, _VP, C.c_uint32, C.c_int, C.c_uint32, _VP, C.c_uint32, _VP]),
    "ab_server_config_default": (C.c_int, [_VP]),
    "ab_server_create": (C.c_int, [_VP, _VP, _VP]),
    "ab_server_ctx": (_VP, [_VP, C.c_uint32]),
    "ab_server_submit": (C.c_int, [_VP, _VP, C.c_uint32, _VP]),
    "ab_server_poll": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_int]),
    "ab_server_release": (C.c_int, [_VP, C.c_uint64]),
    "ab_server_stats": (C.c_int, [_VP, _VP, C.c_int]),
    "ab_server_free": (C.c_int, [_VP]),
    "ab_server_replay": (C.c_int, [_VP, _VP, _VP, C.c_uint32, C.c_double, C.c_uint64, _VP, _VP, _VP]),
    "ab_query_features": (C.c_int, [_VP, _VP, _VP, C.c_uint32, _VP, _VP]),
    "ab_query_features_host": (C.c_int, [_VP, _VP, C.c_uint32, _VP, _VP]),
    "ab_gen_pairs": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_double,
                               C.c_double, C.c_uint32, C.c_uint32, C.c_double, C.c_int, C.POINTER(_VP)]),
    "ab_gen_queries": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_double,
                                 C.c_double, C.c_uint32, C.c_uint32, C.c_int, C.c_double, C.POINTER(_VP)]),
    "ab_dataset_len": (C.c_uint64, [_VP]),
    "ab_dataset_count": (C.c_uint64, [_VP]),
    "ab_dataset_a": (_VP, [_VP]),
    "ab_dataset_b": (_VP, [_VP]),
    "ab_dataset_offsets": (_VP, [_VP]),
    "ab_dataset_free": (None, [_VP]),
    "ab_read_query": (C.c_int64, [C.c_char_p, _VP, C.c_uint64]),
    "ab_read_coo": (C.c_int64, [C.c_char_p, _VP, _VP, C.c_uint64, _U32P, _U32P]),
    "ab_find_query_files": (C.c_int64, [C.c_char_p, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "ab_stats": (None, [_VP, C.c_uint64, _VP]),
}

# Symbols of include/adabool.h (checked by tests: every declared entry point
# must be exported).
EXPORTED = tuple(_SIGS.keys())


def load_library(path: Optional[str] = None):
    """Loads libadabool.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or lib_path()
    if not os.path.exists(p):
        raise ImportError(f"adabool_amd: compiled library not found at {p}; run __graft_entry__.build()")
    # PyTorch ships its own HIP runtime, which it loads under a different file
    # name than the libamdhip64.so.7 this library links.  Loading torch first
    # lets this library bind to that same runtime instance; the other order
    # starts two HIP runtimes in one process and torch then sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    # optional entry points added by later modules are bound when present
    for name, (res, args) in _OPTIONAL_SIGS.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


_OPTIONAL_SIGS: dict = {
    "ab_mlp_nparams": (C.c_uint32, []),
    "ab_mlp_set": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "ab_mlp_get": (C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    "ab_mlp_clear": (C.c_int, [_VP, C.c_int]),
    "ab_mlp_predict": (C.c_int, [_VP, C.c_int, _VP, C.c_uint32, _VP, _VP]),
    "ab_mlp_grad": (C.c_int, [_VP, C.c_int, _VP, _VP, C.c_uint32, _VP, _VP]),
    "ab_mlp_train": (C.c_int, [_VP, C.c_int, _VP, _VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                               _VP]),
}


def _check(st: int):
    if st != 0:
        raise AdaBoolError(st, _lib.ab_last_error().decode(errors="replace"))


def _ptr(a) -> int:
    """Address of a numpy array or torch tensor (or None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    if isinstance(a, int):
        return a
    raise TypeError(f"cannot take the address of {type(a)}")


def _buf_bytes(a, n: Optional[int]) -> int:
    """Bytes a device buffer holds: n when given (never more than a tensor's
    own size), else the tensor's size."""
    if hasattr(a, "element_size"):
        have = a.numel() * a.element_size()
        if n is not None and n > have:
            raise ValueError(f"res_len_bytes {n} exceeds the tensor's {have} bytes")
        return have if n is None else n
    if n is None:
        raise ValueError("res_len_bytes is required for a raw device address")
    return n


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


# ---------------------------------------------------------------------------
# synthetic data, text formats, statistics
# ---------------------------------------------------------------------------
def _take_dataset(h) -> Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray], int]:
    L = _lib
    n = L.ab_dataset_len(h)
    cnt = L.ab_dataset_count(h)
    a = np.ctypeslib.as_array((C.c_uint32 * n).from_address(L.ab_dataset_a(h))).copy() if n else np.zeros(0, np.uint32)
    pb = L.ab_dataset_b(h)
    b = np.ctypeslib.as_array((C.c_uint32 * n).from_address(pb)).copy() if (pb and n) else None
    po = L.ab_dataset_offsets(h)
    off = np.ctypeslib.as_array((C.c_uint64 * (cnt + 1)).from_address(po)).copy() if po else None
    L.ab_dataset_free(h)
    return a, b, off, cnt


def gen_pairs(seed: int, nvalues: int, nkeys: int, *, len_dist: int = 0, p1: float = 0.0, p2: float = 0.0,
              lo: int = 5, hi: int = 15, zipf_s: float = 1.0, perm_seed: Optional[int] = None,
              nthreads: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded filter pairs (keys, values): nvalues rows with Zipf(s) keys."""
    load_library()
    h = _VP()
    _check(_lib.ab_gen_pairs(seed, seed if perm_seed is None else perm_seed, nvalues, nkeys, len_dist, p1, p2,
                             lo, hi, zipf_s, nthreads, C.byref(h)))
    a, b, _, _ = _take_dataset(h)
    return a, b


def gen_queries(seed: int, nkeys: int, nq: int, *, size_dist: int = 0, p1: float = 0.0, p2: float = 0.0,
                lo: int = 64, hi: int = 64, key_dist: int = 0, zipf_s: float = 1.0,
                perm_seed: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded queries: (keys concatenated, offsets[nq+1])."""
    load_library()
    h = _VP()
    _check(_lib.ab_gen_queries(seed, seed if perm_seed is None else perm_seed, nkeys, nq, size_dist, p1, p2,
                               lo, hi, key_dist, zipf_s, C.byref(h)))
    a, _, off, _ = _take_dataset(h)
    return a, off


def read_query(path: str) -> np.ndarray:
    """ReadQuery (test/adsfilter_q_all.cu:164-193)."""
    load_library()
    n = _lib.ab_read_query(path.encode(), None, 0)
    if n < 0:
        raise AdaBoolError(5, _lib.ab_last_error().decode())
    out = np.zeros(n, np.uint32)
    _lib.ab_read_query(path.encode(), _ptr(out), n)
    return out


def find_query_files(dir_path: str) -> list:
    """FindQueryFiles (test/adsfilter_q_all.cu:211-231): paths of the regular
    files named query_vec_<digits>.coo in dir_path, sorted lexicographically.
    As the reference does, an unreadable directory is reported on stderr and
    gives an empty list."""
    load_library()
    need = C.c_uint64()
    n = _lib.ab_find_query_files(dir_path.encode(), None, 0, C.byref(need))
    if n < 0:
        import sys
        print(f"Error reading directory: {_lib.ab_last_error().decode()}", file=sys.stderr)
        return []
    buf = C.create_string_buffer(max(need.value, 1))
    n = _lib.ab_find_query_files(dir_path.encode(), buf, need.value, C.byref(need))
    return [x.decode() for x in buf.raw[:need.value].split(b"\0")[:n]]


def read_coo(path: str) -> Tuple[np.ndarray, np.ndarray, int, int]:
    """ReadCOO (test/adsfilter_q_all.cu:127-161): (keys, values, max_key, max_value)."""
    load_library()
    mk, mv = C.c_uint32(), C.c_uint32()
    n = _lib.ab_read_coo(path.encode(), None, None, 0, C.byref(mk), C.byref(mv))
    if n < 0:
        raise AdaBoolError(5, _lib.ab_last_error().decode())
    k = np.zeros(n, np.uint32)
    v = np.zeros(n, np.uint32)
    _lib.ab_read_coo(path.encode(), _ptr(k), _ptr(v), n, C.byref(mk), C.byref(mv))
    return k, v, mk.value, mv.value


def stats(samples: Sequence[float]) -> dict:
    """CalculateStats (test/adsfilter_q_all.cu:258-307)."""
    load_library()
    d = np.ascontiguousarray(np.asarray(samples, dtype=np.float64)).copy()
    o = np.zeros(9, np.float64)
    _lib.ab_stats(_ptr(d), d.size, _ptr(o))
    return dict(zip(("mean", "median", "stddev", "min", "max", "p50", "p90", "p95", "p99"), o.tolist()))


def key_block_meta(step: int, keys) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """ComputeKeyBlockMeta (src/AdaBool/inverted_list.h:90-148): (starts, pos, buf)."""
    load_library()
    k = _u32(keys)
    n = k.size
    if n == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint8)
    nb = (n + step - 1) // step
    nbytes = _lib.ab_key_block_meta(step, n, _ptr(k), None, None, None)
    starts = np.zeros(nb + 1, np.uint32)
    pos = np.zeros(nb + 1, np.uint32)
    buf = np.zeros(nbytes, np.uint8)
    _lib.ab_key_block_meta(step, n, _ptr(k), _ptr(starts), _ptr(pos), _ptr(buf))
    return starts, pos, buf


# ---------------------------------------------------------------------------
# index and contexts
# ---------------------------------------------------------------------------
class Index:
    """A compressed index resident in device memory (or host-only)."""

    def __init__(self, keys, values, nkeys: int = 0, nvalues: int = 0, tile_shift: int = 0,
                 host_only: bool = False):
        load_library()
        k, v = _u32(keys), _u32(values)
        if k.shape != v.shape:
            raise ValueError("keys and values must have the same length")
        h = _VP()
        fn = _lib.ab_index_build_host if host_only else _lib.ab_index_build
        _check(fn(_ptr(k), _ptr(v), k.size, nkeys, nvalues, tile_shift, C.byref(h)))
        self._h = h
        self.host_only = host_only

    @classmethod
    def from_coo(cls, path: str, tile_shift: int = 0) -> "Index":
        load_library()
        obj = cls.__new__(cls)
        h = _VP()
        _check(_lib.ab_index_load_coo(path.encode(), tile_shift, C.byref(h)))
        obj._h = h
        obj.host_only = False
        return obj

    @classmethod
    def build_device(cls, keys, values, nkeys: int = 0, nvalues: int = 0, tile_shift: int = 0) -> "Index":
        """The same index built on the GPU (pairs as host arrays or device
        tensors)."""
        torch = _torch()
        load_library()
        dk = keys if hasattr(keys, "data_ptr") else torch.from_numpy(_u32(keys).view(np.int32)).cuda()
        dv = values if hasattr(values, "data_ptr") else torch.from_numpy(_u32(values).view(np.int32)).cuda()
        if dk.numel() != dv.numel():
            raise ValueError("keys and values must have the same length")
        torch.cuda.synchronize()
        obj = cls.__new__(cls)
        h = _VP()
        _check(_lib.ab_index_build_device(_ptr(dk), _ptr(dv), dk.numel(), nkeys, nvalues, tile_shift, None,
                                          C.byref(h)))
        obj._h = h
        obj.host_only = False
        return obj

    @classmethod
    def load(cls, path: str, host_only: bool = False) -> "Index":
        """Index from a binary index file written by save()."""
        load_library()
        obj = cls.__new__(cls)
        h = _VP()
        _check(_lib.ab_index_load(os.fsencode(path), 1 if host_only else 0, C.byref(h)))
        obj._h = h
        obj.host_only = host_only
        return obj

    def save(self, path: str):
        _check(_lib.ab_index_save(self._h, os.fsencode(path)))

    def close(self):
        if getattr(self, "_h", None):
            _lib.ab_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self) -> dict:
        i = _IndexInfo()
        _check(_lib.ab_index_get_info(self._h, C.byref(i)))
        d = {f: getattr(i, f) for f, _ in _IndexInfo._fields_}
        d["nrec"] = list(i.nrec)
        return d

    @property
    def stride_bytes(self) -> int:
        return int(_lib.ab_result_stride_bytes(self._h))

    @property
    def stride_words(self) -> int:
        return self.stride_bytes // 4

    def id_bound(self, keys) -> int:
        """Upper bound of a query's id count (its keys' list lengths, capped
        at nvalues)."""
        k = _u32(keys)
        b = C.c_uint64()
        _check(_lib.ab_query_id_bound(self._h, _ptr(k), k.size, C.byref(b)))
        return b.value

    def key_segments(self, t: int) -> np.ndarray:
        """key_segments_ prefix of record type t (nkeys + 1 u32)."""
        segs = np.zeros(self.info["nkeys"] + 1, np.uint32)
        _check(_lib.ab_index_export_type(self._h, t, None, None, _ptr(segs)))
        return segs

    def export_type(self, t: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(headers, payload bytes, key_segments prefix) of record type t."""
        inf = self.info
        n = inf["nrec"][t]
        hdr = np.zeros(n, np.uint32)
        pay = np.zeros(n * PAYLOAD_BYTES[t], np.uint8)
        segs = np.zeros(inf["nkeys"] + 1, np.uint32)
        _check(_lib.ab_index_export_type(self._h, t, _ptr(hdr), _ptr(pay) if pay.size else None, _ptr(segs)))
        return hdr, pay, segs

    def decode_tiles(self) -> Tuple[np.ndarray, np.ndarray]:
        n = _lib.ab_index_decode_tiles(self._h, None, None, 0)
        if n < 0:
            raise AdaBoolError(3, _lib.ab_last_error().decode())
        k = np.zeros(n, np.uint32)
        v = np.zeros(n, np.uint32)
        _lib.ab_index_decode_tiles(self._h, _ptr(k), _ptr(v), n)
        return k, v

    def features_host(self, keys) -> Tuple[np.ndarray, int]:
        k = _u32(keys)
        f = np.zeros(NFEATURES, np.float32)
        m = C.c_uint64()
        _check(_lib.ab_query_features_host(self._h, _ptr(k), k.size, _ptr(f), C.byref(m)))
        return f, m.value


def _torch():
    import torch  # noqa: WPS433 -- device memory plumbing only
    if not torch.cuda.is_available():
        raise RuntimeError("adabool_amd: no GPU visible")
    return torch


class Context:
    """Per-stream execution context over an Index (ab_ctx)."""

    def __init__(self, index: Index, stream: Optional[int] = None, _handle=None):
        self.index = index
        self._owned = _handle is None
        if _handle is not None:  # a context owned by a Server
            self._h = _VP(_handle)
            return
        h = _VP()
        _check(_lib.ab_ctx_create(index._h, stream, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                _lib.ab_ctx_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # knobs
    def set_threshold(self, tau: float):
        _check(_lib.ab_ctx_set_threshold(self._h, tau))

    @property
    def threshold(self) -> float:
        return _lib.ab_ctx_get_threshold(self._h)

    def set_batch_threshold(self, tau: float):
        _check(_lib.ab_ctx_set_batch_threshold(self._h, tau))

    @property
    def batch_threshold(self) -> float:
        return _lib.ab_ctx_get_batch_threshold(self._h)

    def set_dense_group(self, w: int):
        _check(_lib.ab_ctx_set_dense_group(self._h, w))

    def set_dense_threads(self, t: int):
        _check(_lib.ab_ctx_set_dense_threads(self._h, t))

    def set_dense_columns(self, on: bool):
        _check(_lib.ab_ctx_set_dense_columns(self._h, 1 if on else 0))

    def set_sparse_chunk(self, n: int):
        _check(_lib.ab_ctx_set_sparse_chunk(self._h, n))

    BATCH_HOST, BATCH_FUSED, BATCH_GRAPH = 0, 1, 2

    def set_batch_mode(self, mode: int):
        """0: host-dispatched launches; 1: device dispatch + one persistent
        union launch; 2 (default): 1 replayed from a cached hipGraph."""
        _check(_lib.ab_ctx_set_batch_mode(self._h, mode))

    def sync(self):
        _check(_lib.ab_ctx_sync(self._h))

    KERNEL_CLASSES = ("decide", "sparse_zero", "sparse_plan", "sparse_union", "dense_mask", "dense_tiles",
                      "popcount", "write_ids", "fused_union", "ids")

    def set_timing(self, on: bool = True):
        _check(_lib.ab_ctx_set_timing(self._h, 1 if on else 0))

    def timing(self, reset: bool = True) -> dict:
        """{class: (total ms, launches)} since the last reset."""
        ms = np.zeros(len(self.KERNEL_CLASSES), np.float64)
        n = np.zeros(len(self.KERNEL_CLASSES), np.uint64)
        _check(_lib.ab_ctx_timing(self._h, _ptr(ms), _ptr(n), 1 if reset else 0))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNEL_CLASSES)}

    @property
    def stream(self) -> int:
        return _lib.ab_ctx_stream(self._h)

    # single queries -------------------------------------------------------
    def query_device(self, keys, path: int, d_bitmap=None, d_ids=None, d_count=None) -> int:
        """Enqueues one query of host keys; returns the path chosen."""
        k = _u32(keys)
        chosen = C.c_int()
        _check(_lib.ab_query(self._h, _ptr(k), k.size, path, _ptr(d_bitmap), _ptr(d_ids), _ptr(d_count),
                             C.byref(chosen)))
        return chosen.value

    def query(self, keys, path: int = AB_PATH_THRESHOLD):
        """One query; returns (bitmap u32 words, ascending ids, path)."""
        torch = _torch()
        sw = self.index.stride_words
        nv = self.index.info["nvalues"]
        bm = torch.empty(sw, dtype=torch.int32, device="cuda")
        ids = torch.empty(max(nv, 1), dtype=torch.int32, device="cuda")
        cnt = torch.empty(1, dtype=torch.int32, device="cuda")
        p = self.query_device(keys, path, bm, ids, cnt)
        self.sync()
        n = int(cnt.cpu().numpy().view(np.uint32)[0])
        return (bm.cpu().numpy().view(np.uint32), ids[:n].cpu().numpy().view(np.uint32), p)

    # batches --------------------------------------------------------------
    def query_batch_device(self, d_keys, qoff, path: int, d_bitmaps=None, d_ids=None, ids_stride: int = 0,
                           d_counts=None, want_paths: bool = True):
        """Enqueues one batch; returns each query's path (which waits for the
        batch), or None with want_paths=False (nothing waits)."""
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        nq = off.size - 1
        paths = np.zeros(max(nq, 1), np.uint8) if want_paths else None
        _check(_lib.ab_query_batch(self._h, _ptr(d_keys), _ptr(off), nq, path, _ptr(d_bitmaps), _ptr(d_ids),
                                   ids_stride, _ptr(d_counts), _ptr(paths)))
        return paths[:nq] if want_paths else None

    def autotune(self, d_keys, qoff, path: int = AB_PATH_AUTO, repeats: int = 3) -> list:
        """Times the launch-parameter grid on this batch and keeps the fastest
        setting whose bitmaps equal the entry settings'; returns every
        candidate, valid ones (mismatched_words == 0) first, fastest first."""
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        cap = 128  # the grid holds 72 candidates
        buf = (_TuneConfig * cap)()
        n = C.c_uint32()
        _check(_lib.ab_ctx_autotune(self._h, _ptr(d_keys), _ptr(off), off.size - 1, path, repeats, buf, cap,
                                    C.byref(n)))
        return [{f: getattr(b, f) for f, _ in _TuneConfig._fields_} for b in buf[: min(cap, n.value)]]

    def apply_tuning(self, cfg: dict):
        self.set_dense_group(int(cfg["group"]))
        self.set_dense_threads(int(cfg["dense_threads"]))
        self.set_sparse_chunk(int(cfg["sparse_chunk"]))
        self.set_dense_columns(bool(cfg["dense_columns"]))

    def set_batch_tuning(self, table):
        """Launch settings by batch size: a list of dicts with min_batch and
        the apply_tuning keys, in ascending min_batch; each batch applies the
        entry with the largest min_batch <= its query count ([] removes it)."""
        arr = (_Tuning * max(len(table), 1))()
        for i, e in enumerate(table):
            arr[i] = _Tuning(int(e["min_batch"]), int(e["group"]), int(e["dense_threads"]), int(e["sparse_chunk"]),
                             1 if e["dense_columns"] else 0)
        _check(_lib.ab_ctx_set_batch_tuning(self._h, C.byref(arr), len(table)))

    def batch_settings(self) -> dict:
        """The launch settings in force (after the last batch's table entry)."""
        t = _Tuning()
        _check(_lib.ab_ctx_get_batch_settings(self._h, C.byref(t)))
        return {"group": t.group, "dense_threads": t.dense_threads, "sparse_chunk": t.sparse_chunk,
                "dense_columns": bool(t.dense_columns)}

    def query_batch_packed(self, d_keys, qoff, path: int, d_ids, id_off, d_bitmaps=None, d_counts=None) -> np.ndarray:
        """Batch with query i's ids written at d_ids + id_off[i] (u32 units)."""
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        ido = np.ascontiguousarray(np.asarray(id_off, dtype=np.uint64))
        nq = off.size - 1
        paths = np.zeros(max(nq, 1), np.uint8)
        _check(_lib.ab_query_batch_packed(self._h, _ptr(d_keys), _ptr(off), nq, path, _ptr(d_bitmaps), _ptr(d_ids),
                                          _ptr(ido), _ptr(d_counts), _ptr(paths)))
        return paths[:nq]

    def set_id_budget(self, values: int):
        _check(_lib.ab_ctx_set_id_budget(self._h, int(values)))

    def set_id_method(self, method: int):
        """Id-only batches: 0 automatic, 1 always by sort, 2 by value slice wherever it applies."""
        _check(_lib.ab_ctx_set_id_method(self._h, int(method)))

    def query_batch_ids_device(self, d_keys, qoff, d_ids, id_off, d_counts=None):
        """Id-only batch (no bitmaps): query i's ascending ids at
        d_ids + id_off[i] (u32 units), its count at d_counts[i]."""
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        ido = np.ascontiguousarray(np.asarray(id_off, dtype=np.uint64))
        _check(_lib.ab_query_batch_ids(self._h, _ptr(d_keys), _ptr(off), off.size - 1, _ptr(d_ids), _ptr(ido),
                                       _ptr(d_counts)))

    def query_batch_ids(self, keys, qoff):
        """Id-only batch of host queries: (list of id arrays, counts).  Each
        query gets room for its id upper bound (Index.id_bound)."""
        torch = _torch()
        k = _u32(keys)
        off = np.asarray(qoff, dtype=np.uint64)
        nq = off.size - 1
        ub = np.array([self.index.id_bound(k[off[q]:off[q + 1]]) for q in range(nq)], np.uint64)
        room = (ub + 3) // 4 * 4
        ido = np.zeros(max(nq, 1), np.uint64)
        if nq:
            ido[1:nq] = np.cumsum(room)[:-1]
        total = int(room.sum()) if nq else 0
        dk = torch.from_numpy(k.view(np.int32)).cuda() if k.size else torch.zeros(1, dtype=torch.int32, device="cuda")
        ids = torch.empty(max(total, 4), dtype=torch.int32, device="cuda")
        cnt = torch.empty(max(nq, 1), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        self.query_batch_ids_device(dk, off, ids, ido[:nq], cnt)
        self.sync()
        counts = cnt[:nq].cpu().numpy().view(np.uint32).astype(np.int64)
        ih = ids.cpu().numpy().view(np.uint32)
        return [ih[int(ido[q]): int(ido[q]) + int(counts[q])].copy() for q in range(nq)], counts

    def query_batch(self, keys, qoff, path: int = AB_PATH_THRESHOLD, want_ids: bool = True):
        """Batch of queries (host arrays in, host arrays out):
        (bitmaps [nq, stride_words] u32, list of id arrays, counts, paths)."""
        torch = _torch()
        k = _u32(keys)
        off = np.asarray(qoff, dtype=np.uint64)
        nq = off.size - 1
        sw = self.index.stride_words
        nv = self.index.info["nvalues"]
        dk = torch.from_numpy(k.view(np.int32)).cuda() if k.size else torch.zeros(1, dtype=torch.int32, device="cuda")
        bms = torch.empty(nq * sw, dtype=torch.int32, device="cuda")
        ids = torch.empty(nq * max(nv, 1), dtype=torch.int32, device="cuda") if want_ids else None
        cnt = torch.empty(max(nq, 1), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        paths = self.query_batch_device(dk, off, path, bms, ids, max(nv, 1), cnt)
        self.sync()
        counts = cnt[:nq].cpu().numpy().view(np.uint32).astype(np.int64)
        bm_h = bms.cpu().numpy().view(np.uint32).reshape(nq, sw)
        out_ids = []
        if want_ids:
            ih = ids.view(nq, max(nv, 1))
            for q in range(nq):
                out_ids.append(ih[q, : counts[q]].cpu().numpy().view(np.uint32))
        return bm_h, out_ids, counts, paths

    # cost model -----------------------------------------------------------
    # slot 0 (MODEL_SINGLE) prices single queries, slot 1 (MODEL_BATCH) a
    # query's share of a batch
    def set_model(self, params, mean, std, slot: int = 0):
        p = np.ascontiguousarray(params, dtype=np.float32)
        mu = np.ascontiguousarray(mean, dtype=np.float32)
        sd = np.ascontiguousarray(std, dtype=np.float32)
        if p.size != _lib.ab_mlp_nparams() or mu.size != NFEATURES or sd.size != NFEATURES:
            raise ValueError("cost model shape mismatch")
        _check(_lib.ab_mlp_set(self._h, slot, _ptr(p), _ptr(mu), _ptr(sd)))

    def get_model(self, slot: int = 0):
        p = np.zeros(_lib.ab_mlp_nparams(), np.float32)
        mu = np.zeros(NFEATURES, np.float32)
        sd = np.zeros(NFEATURES, np.float32)
        _check(_lib.ab_mlp_get(self._h, slot, _ptr(p), _ptr(mu), _ptr(sd)))
        return p, mu, sd

    def clear_model(self, slot: Optional[int] = None):
        for s_ in ((0, 1) if slot is None else (slot,)):
            _check(_lib.ab_mlp_clear(self._h, s_))

    def load_model(self, path: str) -> dict:
        """Loads a model file: {"single": {...}, "batch": {...}} (either may be
        absent) with params/mean/std each."""
        import json
        d = json.load(open(path))
        for slot, name in ((MODEL_SINGLE, "single"), (MODEL_BATCH, "batch")):
            if name in d:
                self.set_model(d[name]["params"], d[name]["mean"], d[name]["std"], slot)
        return d

    def model_dict(self, slot: int) -> dict:
        p, mu, sd = self.get_model(slot)
        return {"dims": [NFEATURES, 64, 64, 2], "params": p.tolist(), "mean": mu.tolist(), "std": sd.tolist()}

    def predict(self, x, slot: int = 0):
        xa = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEATURES)
        y = np.zeros((xa.shape[0], 2), np.float32)
        sel = np.zeros(max(xa.shape[0], 1), np.uint8)
        _check(_lib.ab_mlp_predict(self._h, slot, _ptr(xa), xa.shape[0], _ptr(y), _ptr(sel)))
        return y, sel[: xa.shape[0]]

    def model_grad(self, x, y, slot: int = 0):
        xa = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEATURES)
        ya = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 2)
        g = np.zeros(_lib.ab_mlp_nparams(), np.float32)
        loss = C.c_float()
        _check(_lib.ab_mlp_grad(self._h, slot, _ptr(xa), _ptr(ya), xa.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g

    def train_model(self, x, y, epochs: int = 30, batch: int = 256, lr: float = 1e-3, seed: int = 1,
                    slot: int = 0):
        xa = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NFEATURES)
        ya = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 2)
        curve = np.zeros(max(epochs, 1), np.float32)
        _check(_lib.ab_mlp_train(self._h, slot, _ptr(xa), _ptr(ya), xa.shape[0], epochs, batch, lr, seed,
                                 _ptr(curve)))
        return curve[:epochs]

    def features(self, d_keys, qoff) -> Tuple[np.ndarray, np.ndarray]:
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        nq = off.size - 1
        f = np.zeros((max(nq, 1), NFEATURES), np.float32)
        m = np.zeros(max(nq, 1), np.uint64)
        _check(_lib.ab_query_features(self._h, _ptr(d_keys), _ptr(off), nq, _ptr(f), _ptr(m)))
        return f[:nq], m[:nq]


# ---------------------------------------------------------------------------
# the reference's own interface (ivf_list and its workers)
# ---------------------------------------------------------------------------
class _Tuning(C.Structure):
    _fields_ = [("min_batch", C.c_uint32), ("group", C.c_uint32), ("dense_threads", C.c_uint32),
                ("sparse_chunk", C.c_uint32), ("dense_columns", C.c_int32)]


class _TuneConfig(C.Structure):
    _fields_ = [("group", C.c_uint32), ("dense_threads", C.c_uint32), ("sparse_chunk", C.c_uint32),
                ("dense_columns", C.c_int), ("ms", C.c_double), ("mismatched_words", C.c_uint64)]


class _ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("max_wait_us", C.c_uint32), ("inflight", C.c_uint32),
                ("arena_bytes", C.c_uint64), ("path", C.c_int), ("deliver", C.c_int), ("batch_mode", C.c_int),
                ("ids_only", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int), ("path", C.c_int), ("count", C.c_uint32),
                ("ids", C.c_void_p), ("latency_ms", C.c_double)]


class _ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("batches", C.c_uint64), ("dense_queries", C.c_uint64),
                ("mean_batch", C.c_double), ("qps", C.c_double), ("lat_mean_ms", C.c_double),
                ("lat_p50_ms", C.c_double), ("lat_p90_ms", C.c_double), ("lat_p99_ms", C.c_double),
                ("lat_max_ms", C.c_double), ("arena_used_bytes", C.c_uint64)]


class Server:
    """Serving loop over a resident index (ab_server): thread-safe submit,
    memory-based admission, continuous batching on `inflight` streams.

    Settings of the contexts (thresholds, cost models, group width...) are
    made through `contexts` before the first submission."""

    DELIVER_DEVICE, DELIVER_HOST = 0, 1

    def __init__(self, index: Index, **cfg):
        self.index = index
        c = _ServerConfig()
        _check(_lib.ab_server_config_default(C.byref(c)))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown server setting {k}")
            setattr(c, k, v)
        self.config = {f: getattr(c, f) for f, _ in _ServerConfig._fields_}
        h = _VP()
        _check(_lib.ab_server_create(index._h, C.byref(c), C.byref(h)))
        self._h = h
        self.contexts = [Context(index, _handle=_lib.ab_server_ctx(h, i)) for i in range(c.inflight)]

    def load_model(self, path: str):
        for ctx in self.contexts:
            d = ctx.load_model(path)
        return d

    def submit(self, keys) -> int:
        k = _u32(keys)
        t = C.c_uint64()
        _check(_lib.ab_server_submit(self._h, _ptr(k), k.size, C.byref(t)))
        return t.value

    def poll(self, cap: int = 256, wait_ms: int = 0) -> list:
        """Completed requests: dicts with ticket, status, path, count,
        ids (numpy copy when results are delivered to the host, else the
        device address) and latency_ms."""
        buf = (_Result * cap)()
        n = C.c_uint32()
        _check(_lib.ab_server_poll(self._h, buf, cap, C.byref(n), wait_ms))
        out = []
        for r in buf[: n.value]:
            ids = r.ids
            if self.config["deliver"] == 1 and r.ids and r.status == 0:
                ids = np.ctypeslib.as_array((C.c_uint32 * r.count).from_address(r.ids)).copy() if r.count else \
                    np.zeros(0, np.uint32)
            out.append({"ticket": r.ticket, "status": r.status, "path": r.path, "count": r.count, "ids": ids,
                        "latency_ms": r.latency_ms})
        return out

    def release(self, ticket: int):
        _check(_lib.ab_server_release(self._h, ticket))

    def stats(self, reset: bool = False) -> dict:
        st = _ServerStats()
        _check(_lib.ab_server_stats(self._h, C.byref(st), 1 if reset else 0))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def replay(self, keys, qoff, rate_qps: float = 0.0, seed: int = 1):
        """Open-loop Poisson replay; returns (counts, latency ms from scheduled
        arrival, wall seconds)."""
        k = _u32(keys)
        off = np.ascontiguousarray(np.asarray(qoff, dtype=np.uint64))
        nq = off.size - 1
        counts = np.zeros(max(nq, 1), np.uint32)
        lat = np.zeros(max(nq, 1), np.float64)
        wall = C.c_double()
        _check(_lib.ab_server_replay(self._h, _ptr(k), _ptr(off), nq, rate_qps, seed, _ptr(counts), _ptr(lat),
                                     C.byref(wall)))
        return counts[:nq], lat[:nq], wall.value

    def close(self):
        if getattr(self, "_h", None):
            for ctx in self.contexts:
                ctx.close()
            _lib.ab_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SparseWorker:
    """Prepared sparse query (ivf_list::Prepare(keys), inverted_list.h:481)."""

    def __init__(self, lst: "IvfList", keys):
        self._l = lst
        self.keys = _u32(keys)

    def run(self, d_res, res_len_bytes: Optional[int] = None):
        """ListUnion_sparse (test/adsfilter_q_all.cu:64-105): writes the result
        bitmap into device buffer d_res, min(its size, stride_bytes) bytes; a
        torch tensor's size is its own, a raw address needs res_len_bytes
        (at least ceil(nvalues / 8))."""
        n = _buf_bytes(d_res, res_len_bytes)
        _check(_lib.ab_list_union_sparse(self._l.ctx._h, _ptr(self.keys), self.keys.size, _ptr(d_res), n))


class DenseWorker:
    """Prepared dense query (ivf_list::Prepare(len, bitmap), inverted_list.h:530)."""

    def __init__(self, lst: "IvfList", nbytes: int, bitmap):
        self._l = lst
        self.bitmap = np.ascontiguousarray(np.frombuffer(bytes(bitmap), dtype=np.uint8)[:nbytes]) \
            if not isinstance(bitmap, np.ndarray) else np.ascontiguousarray(bitmap.view(np.uint8)[:nbytes])

    def run(self, d_res, res_len_bytes: Optional[int] = None):
        """ListUnion_dense (test/adsfilter_q_all.cu:26-61); d_res as in
        SparseWorker.run."""
        n = _buf_bytes(d_res, res_len_bytes)
        _check(_lib.ab_list_union_dense(self._l.ctx._h, self.bitmap.size, _ptr(self.bitmap), _ptr(d_res), n))


class BatchWorker:
    """Batched dense worker (PrepareBatch / PrepareBatchDevice,
    inverted_list.h:564-605)."""

    def __init__(self, lst: "IvfList", batch_size: int, nbytes: int, d_bitmaps, keep=None):
        self._l = lst
        self.batch_size = batch_size
        self.len = nbytes
        self.d_bitmaps = d_bitmaps
        self._keep = keep

    def RunArrays(self, d_res, res_len_u32: int):  # noqa: N802 -- reference name
        """RunArrays (src/AdaBool/inverted_list.cu.h:224-228), any batch size."""
        _check(_lib.ab_list_union_dense_batch(self._l.ctx._h, self.batch_size, self.len, _ptr(self.d_bitmaps),
                                              _ptr(d_res), res_len_u32))


class IvfList:
    """ivf_list<256,64> (src/AdaBool/inverted_list.h:315-829) over the engine."""

    def __init__(self):
        self.index: Optional[Index] = None
        self.ctx: Optional[Context] = None

    def BuildFrom(self, kvs, nkeys: int = 0, nvalues: int = 0, tile_shift: int = 0):  # noqa: N802
        """Builds from (key, value) pairs: an [n, 2] array or two arrays."""
        if isinstance(kvs, tuple) and len(kvs) == 2:
            k, v = kvs
        else:
            a = np.asarray(kvs, dtype=np.uint32).reshape(-1, 2)
            k, v = a[:, 0], a[:, 1]
        self.index = Index(k, v, nkeys, nvalues, tile_shift)
        self.ctx = Context(self.index)
        return self

    def Prepare(self, keys) -> SparseWorker:  # noqa: N802
        return SparseWorker(self, keys)

    def PrepareBitmap(self, nbytes: int, bitmap) -> DenseWorker:  # noqa: N802
        return DenseWorker(self, nbytes, bitmap)

    def PrepareBatch(self, batch_size: int, nbytes: int, host_bitmaps) -> BatchWorker:  # noqa: N802
        torch = _torch()
        hb = np.ascontiguousarray(np.asarray(host_bitmaps, dtype=np.uint8).reshape(-1))
        d = torch.from_numpy(hb).cuda()
        torch.cuda.synchronize()
        return BatchWorker(self, batch_size, nbytes, d, keep=d)

    def PrepareBatchDevice(self, batch_size: int, nbytes: int, d_bitmaps) -> BatchWorker:  # noqa: N802
        return BatchWorker(self, batch_size, nbytes, d_bitmaps)

    def BatchListUnionSparse(self, flat_keys, qidx, nq: int, d_res, res_len_bytes: int):  # noqa: N802
        """BatchListUnion_sparse (test/adsfilter_batch.cu:39-79)."""
        k, q = _u32(flat_keys), _u32(qidx)
        _check(_lib.ab_list_union_sparse_batch(self.ctx._h, _ptr(k), _ptr(q), k.size, nq, _ptr(d_res),
                                               res_len_bytes))

    def sync(self):
        self.ctx.sync()
