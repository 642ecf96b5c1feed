"""The benchmark's workload generators (splitmix64, wwlib) reproduce the committed
fixtures exactly: the 4096-range batched workload (seed 0) and the 100 req/s serving
stream (seed 1). No device needed."""
import json
import os

import wwlib

HERE = os.path.dirname(os.path.abspath(__file__))
W = json.load(open(os.path.join(HERE, "golden", "workloads.json")))


def _digest(items):
    import hashlib
    text = "\n".join(" ".join(repr(f) for f in it) for it in items)
    return hashlib.sha256(text.encode()).hexdigest()


def test_batch_workload_is_pinned():
    b = W["batch"]
    ranges = wwlib.batch_workload(n=b["n"], seed=b["seed"])
    assert len(ranges) == b["count"]
    assert [list(r) for r in ranges[:64]] == b["first"]
    assert _digest(ranges) == b["sha256"]
    for r in b["reference"]:
        assert tuple(ranges[r["index"]]) == (r["lo"], r["hi"])
    # the shape the benchmark promises: odd starts, widths 2^16..2^28, inside 64 bits
    assert all(lo & 1 and (1 << 16) <= hi - lo <= (1 << 28) and hi < (1 << 64)
               for lo, hi in ranges)


def test_serving_workload_is_pinned():
    s = W["serving"]
    reqs = wwlib.serving_workload(rate=s["rate"], duration_s=s["duration_s"], seed=s["seed"])
    assert len(reqs) == s["count"]
    assert [list(r) for r in reqs[:16]] == s["first"]
    assert _digest(reqs) == s["sha256"]
    assert all(0 <= t < s["duration_s"] and lo & 1 and hi > lo for t, lo, hi in reqs)


def test_serving_streams_are_pinned():
    """The 20 and 50 req/s streams and the wide requests of the 100 req/s stream whose
    reference summaries the GPU suite checks."""
    for key in ("20.0", "50.0"):
        s = W["serving_streams"][key]
        reqs = wwlib.serving_workload(rate=s["rate"], duration_s=s["duration_s"], seed=s["seed"])
        assert len(reqs) == s["count"] and _digest(reqs) == s["sha256"]
        for r in s["reference"]:
            assert tuple(reqs[r["index"]][1:]) == (r["lo"], r["hi"])
    wide = W["serving_streams"]["100.0_wide"]
    reqs = wwlib.serving_workload(rate=100.0, duration_s=10.0, seed=1)
    for r in wide["reference"]:
        assert tuple(reqs[r["index"]][1:]) == (r["lo"], r["hi"]) and r["hi"] - r["lo"] > 1 << 28
