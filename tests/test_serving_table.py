"""The serving module's launch-settings table by batch size (tuning_table):
each sample's winner applies from the geometric mean of its sample size and
the previous one, and repeated sample sizes add nothing."""
from adabool_amd.serving import tuning_table


def _w(group, threads, ms):
    return {"group": group, "dense_threads": threads, "sparse_chunk": 4096, "dense_columns": True, "ms": ms,
            "mismatched_words": 0}


def test_three_samples():
    t = tuning_table([(32, _w(32, 256, 1.2)), (256, _w(64, 256, 2.1)), (1024, _w(256, 512, 6.0))])
    assert [e["min_batch"] for e in t] == [1, 91, 512]
    assert [(e["group"], e["dense_threads"]) for e in t] == [(32, 256), (64, 256), (256, 512)]
    assert all(set(e) == {"group", "dense_threads", "sparse_chunk", "dense_columns", "min_batch", "sample", "ms"}
               for e in t)


def test_small_batches_collapse():
    # fewer queries than the larger samples: those samples are the same size
    t = tuning_table([(20, _w(16, 256, 0.5)), (20, _w(16, 256, 0.5)), (20, _w(8, 512, 0.4))])
    assert [e["min_batch"] for e in t] == [1]
    t = tuning_table([(32, _w(32, 256, 1.0)), (100, _w(64, 256, 1.5)), (100, _w(128, 512, 1.4))])
    assert [e["min_batch"] for e in t] == [1, 57]


def test_min_batch_strictly_ascending():
    t = tuning_table([(1, _w(1, 256, 0.1)), (2, _w(2, 256, 0.2)), (3, _w(3, 256, 0.3))])
    mb = [e["min_batch"] for e in t]
    assert mb == sorted(set(mb)) and mb[0] == 1
