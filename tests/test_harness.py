"""The compiled C++ side of the boundary: the reference-shaped shim
(include/adabool_ivf.hpp) and the two command-line harnesses built on it
(adabool_amd/harness/ab_q_all, ab_batch), the counterparts of
test/adsfilter_q_all.cu and test/adsfilter_batch.cu.

CPU: the shim compiles as a reference caller would use it; the harnesses'
argument errors.  GPU: both harnesses run on the C1 filter written as the
reference's text files, every method matches their ground truth, and query
files are taken in FindQueryFiles order (query_vec_10 before query_vec_2)."""
import os
import re
import subprocess

import numpy as np
import pytest

import adabool_amd as ab
from adabool_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "adabool_amd", "harness")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

CALLER = r"""
// a caller written against the reference's API, switched by its includes
#include <set>
#include <utility>
#include <vector>
#include "adabool_ivf.hpp"

int run(hipStream_t stream, uint32_t *d_res, uint32_t *d_batch, uint32_t *d_qidx) {
    std::vector<std::pair<uint32_t, uint32_t>> kvs = {{0, 1}, {1, 7}, {2, 300}};
    adabool::ivf_list list;
    list.BuildFrom(kvs);
    std::vector<uint32_t> keys = {0, 2};
    auto sw = list.Prepare(keys);
    std::set<uint32_t> ks = {1};
    auto sw2 = list.Prepare(ks);
    (void)adabool::ListUnion_sparse(*sw, d_res, stream);
    (void)adabool::ListUnion_sparse(*sw2, d_res);
    std::vector<uint8_t> bm(1, 0x5);
    auto dw = list.Prepare(bm.size(), bm.data());
    (void)adabool::ListUnion_dense(*dw, d_res, stream);
    std::vector<uint8_t> bms(32, 0);
    auto bw = list.PrepareBatch(2, 16, bms.data());
    bw->RunArrays(stream, d_batch, 16);
    auto bw2 = list.PrepareBatchDevice(2, 16, (const uint8_t *)d_batch);
    (void)bw2;
    (void)adabool::BatchListUnion_sparse(*sw, d_batch, d_qidx, 64, stream);
    uint32_t *ids = nullptr, *cnt = nullptr;
    const int path = adabool::Query(list, keys, d_res, ids, cnt, stream);
    return path + (int)adabool::FindQueryFiles(".").size();
}
"""


def test_shim_compiles_for_a_reference_caller(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _bin(name):
    p = os.path.join(HARNESS, name)
    if not os.path.exists(p):
        pytest.fail(f"{p} is not built (make -C adabool_amd/csrc)")
    return p


def test_harness_argument_errors(tmp_path):
    r = subprocess.run([_bin("ab_batch"), "f.coo", str(tmp_path), "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "batch_size must be greater than 0" in r.stderr and "Usage" in r.stdout
    for prog in ("ab_q_all", "ab_batch"):
        r = subprocess.run([_bin(prog), "f.coo", str(tmp_path)], capture_output=True, text=True)
        assert r.returncode == 1 and "No query files found" in r.stderr, (prog, r.stderr)


def _write_c1_files(d):
    """The C1 filter as "key value" lines (with a comment and a blank line,
    which ReadCOO skips) and five query files plus two that do not match."""
    k, v = configs.make_filter(configs.C1)
    lines = "% C1 filter\n\n" + "\n".join(f"{a} {b}" for a, b in zip(k.tolist(), v.tolist())) + "\n"
    (d / "filter.coo").write_text(lines)
    qk, qo = configs.make_queries(configs.C1, configs.C1_Q, 3)
    qs = [qk[qo[i]:qo[i + 1]] for i in range(3)]
    nk = int(k.max()) + 1
    qs.append(np.array([5, 5, 17, nk, nk + 100, 3], np.uint32))  # duplicates and keys outside the key space
    qs.append(np.array([], np.uint32))
    names = ["query_vec_1.coo", "query_vec_2.coo", "query_vec_10.coo", "query_vec_3.coo", "query_vec_11.coo"]
    for n, q in zip(names, qs):
        (d / n).write_text("# query\n" + "".join(f"{x}\n" for x in q.tolist()))
    (d / "query_vec_x.coo").write_text("1\n")
    (d / "notes.txt").write_text("1\n")
    return sorted(names)


@pytest.mark.gpu
def test_q_all_harness_on_c1_files(tmp_path):
    order = _write_c1_files(tmp_path)
    r = subprocess.run([_bin("ab_q_all"), str(tmp_path / "filter.coo"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "Final result: ALL TESTS PASSED" in r.stdout
    for m in ("Dense method:  5/5 PASS", "Sparse method: 5/5 PASS", "Auto method:   5/5 PASS"):
        assert m in r.stdout
    assert re.findall(r"\] Processing (\S+)", r.stdout) == order  # query_vec_1, _10, _11, _2, _3
    assert "Sparse Method - Execution Time:" in r.stdout and "P99:" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [2, 5])
def test_batch_harness_on_c1_files(tmp_path, batch):
    _write_c1_files(tmp_path)
    r = subprocess.run([_bin("ab_batch"), str(tmp_path / "filter.coo"), str(tmp_path), str(batch)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    nb = 5 // batch
    assert ("WARNING: Truncating to 4 queries" in r.stdout) == (batch == 2)
    for m in ("Dense method: ", "Sparse method:", "Auto method:  "):
        assert f"{m} {nb}/{nb} batches PASS" in r.stdout
    assert "Final result: ALL TESTS PASSED" in r.stdout and "queries/sec" in r.stdout
