"""Contexts used concurrently: several contexts over one shared index, each
on its own stream, with batches in flight on all of them at once (no wait
between calls), and contexts driven from several host threads at once.
Every result equals the oracle's (one context per host thread, the index
shared read-only, is the documented threading contract)."""
import threading

import numpy as np
import pytest

import adabool_amd as ab
import oracle
from adabool_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    import torch

    assert torch.cuda.is_available()
    k, v = configs.make_filter(configs.C1)
    ix = ab.Index(k, v, configs.C1.nkeys, configs.C1.nvalues)
    csr = oracle.CSR(k, v, configs.C1.nkeys, configs.C1.nvalues)
    return ix, csr


def _queries(seed, nq):
    qk, qo = ab.gen_queries(seed, configs.C1.nkeys, nq, size_dist=2, p1=1.5, lo=1, hi=3000, key_dist=1, zipf_s=1.0,
                            perm_seed=configs.C1.seed)
    return qk, qo


@pytest.mark.parametrize("mode", [1, 2])
def test_batches_in_flight_on_several_streams(c1, mode):
    import torch

    ix, csr = c1
    nv = configs.C1.nvalues
    ctxs, outs, sets = [], [], []
    for s in range(3):
        c = ab.Context(ix)
        c.set_batch_mode(mode)
        c.set_batch_threshold(0.01)
        qk, qo = _queries(0xADAB0700 + s, 48)
        nq = qo.size - 1
        dk = torch.from_numpy(qk.view(np.int32)).cuda()
        ids = torch.full((nq * nv,), -1, dtype=torch.int32, device="cuda")
        cnt = torch.full((nq,), -1, dtype=torch.int32, device="cuda")
        ctxs.append(c)
        outs.append((dk, ids, cnt))
        sets.append((qk, qo))
    torch.cuda.synchronize()
    print(f"streams mode={mode}", flush=True)
    for rep in range(4):  # replays with all three contexts in flight together
        for c, (dk, ids, cnt), (qk, qo) in zip(ctxs, outs, sets):
            c.query_batch_device(dk, qo, ab.AB_PATH_THRESHOLD, None, ids, nv, cnt, want_paths=False)
    for c in ctxs:
        c.sync()
    for (dk, ids, cnt), (qk, qo) in zip(outs, sets):
        h = ids.cpu().numpy().view(np.uint32)
        cn = cnt.cpu().numpy().view(np.uint32)
        for i in range(qo.size - 1):
            want = csr.union(qk[qo[i]:qo[i + 1]])[1]
            assert cn[i] == want.size and np.array_equal(h[i * nv:i * nv + want.size], want), f"q{i}"
    for c in ctxs:
        c.close()


def test_contexts_on_host_threads(c1):
    ix, csr = c1
    errors = []

    def work(t):
        try:
            c = ab.Context(ix)
            c.set_batch_threshold(0.01)
            qk, qo = _queries(0xADAB0710 + t, 24)
            for _ in range(3):
                _, ids, counts, _ = c.query_batch(qk, qo, ab.AB_PATH_AUTO)
                for i in range(qo.size - 1):
                    if not np.array_equal(ids[i], csr.union(qk[qo[i]:qo[i + 1]])[1]):
                        errors.append(f"thread {t} q{i}")
                only, _ = c.query_batch_ids(qk, qo)
                for i in range(qo.size - 1):
                    if not np.array_equal(only[i], ids[i]):
                        errors.append(f"thread {t} id-only q{i}")
            c.close()
        except Exception as e:  # reported below
            errors.append(f"thread {t}: {e!r}")

    print("host threads", flush=True)
    th = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    assert not any(x.is_alive() for x in th), "a worker thread did not finish"
    assert not errors, errors[:5]
