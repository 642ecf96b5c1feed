"""Writes tests/golden/workloads.json: the batched and serving workloads of the benchmark,
pinned, with the reference's own CPU path (oracle/_ref, compiled from wwcpu.cpp / prime.h)
on a sample of their ranges.

  * batched: 4096 ranges, splitmix64 seed 0 (wwlib.batch_workload): the SHA-256 of the
    whole list, its first 64 ranges, and reference summaries of 48 sampled ranges of width
    at most 2^22;
  * serving: Poisson arrivals at 100/s for 10 s, seed 1 (wwlib.serving_workload): the
    SHA-256 of the whole request list, its first 16 requests, and reference summaries of
    16 sampled requests of width at most 2^24;
  * serving_streams: the 20/s and 50/s streams (10 s, seed 1): SHA-256, and reference
    summaries of 8 sampled requests of width at most 2^24 each; plus the two narrowest
    requests of the 100/s stream wider than 2^28.

    python tests/golden/make_workloads.py          (needs oracle/_ref built)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "wwocl-boinc_amd"))

import oracle  # noqa: E402  (for the _ref binding only)
import wwlib  # noqa: E402  (the generators; no device needed)


def digest(items) -> str:
    """SHA-256 of the list written one item per line, fields separated by spaces
    (floats in repr form)."""
    text = "\n".join(" ".join(repr(f) for f in it) for it in items)
    return hashlib.sha256(text.encode()).hexdigest()


def summarize(lo: int, hi: int):
    prp, cks, flags = oracle.ref_range(lo, hi, os.cpu_count() or 1)
    return {"lo": lo, "hi": hi, "prp_count": prp, "checksum": cks, "cpu_result_count": flags}


def main():
    if oracle.ref_lib() is None:
        sys.exit("oracle/_ref/libwwref.so is not built (make -C oracle ref)")
    batch = wwlib.batch_workload(n=4096, seed=0)
    small = [i for i, (lo, hi) in enumerate(batch) if hi - lo <= (1 << 22)]
    pick = small[:: max(1, len(small) // 48)][:48]
    serving = wwlib.serving_workload(rate=100.0, duration_s=10.0, seed=1)
    ssmall = [i for i, (_, lo, hi) in enumerate(serving) if hi - lo <= (1 << 24)]
    spick = ssmall[:: max(1, len(ssmall) // 16)][:16]
    fixture = {
        "about": "benchmark workloads (splitmix64 generators of wwlib) and the reference "
                 "CPU path (oracle/_ref) on sampled ranges",
        "batch": {
            "n": 4096, "seed": 0, "count": len(batch), "sha256": digest(batch),
            "first": batch[:64],
            "reference": [dict(index=i, **summarize(*batch[i])) for i in pick],
        },
        "serving": {
            "rate": 100.0, "duration_s": 10.0, "seed": 1, "count": len(serving),
            "sha256": digest(serving), "first": serving[:16],
            "reference": [dict(index=i, **summarize(serving[i][1], serving[i][2]))
                          for i in spick],
        },
    }
    streams = {}
    for rate in (20.0, 50.0):
        reqs = wwlib.serving_workload(rate=rate, duration_s=10.0, seed=1)
        sm = [i for i, (_, lo, hi) in enumerate(reqs) if hi - lo <= (1 << 24)]
        pk = sm[:: max(1, len(sm) // 8)][:8]
        streams[str(rate)] = {
            "rate": rate, "duration_s": 10.0, "seed": 1, "count": len(reqs),
            "sha256": digest(reqs),
            "reference": [dict(index=i, **summarize(reqs[i][1], reqs[i][2])) for i in pk]}
    wide = sorted((hi - lo, i) for i, (_, lo, hi) in enumerate(serving) if hi - lo > (1 << 28))[:2]
    streams["100.0_wide"] = {
        "rate": 100.0, "duration_s": 10.0, "seed": 1,
        "reference": [dict(index=i, **summarize(serving[i][1], serving[i][2])) for _, i in wide]}
    fixture["serving_streams"] = streams
    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(fixture, f, indent=0)
    print(f"batch: {len(batch)} ranges, {len(pick)} summaries; serving: {len(serving)} "
          f"requests, {len(spick)} summaries; streams: "
          + ", ".join(f"{k}: {len(v['reference'])}" for k, v in streams.items()))


if __name__ == "__main__":
    main()
