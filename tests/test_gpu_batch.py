"""The persistent form (ww_batch_*): per-range results equal the plain form's and the
oracle's; graph replays equal eager launches; results do not depend on the mix."""
import json
import os
import random

import pytest

import oracle
import wwlib
from test_gpu_parity import RANGES

pytestmark = pytest.mark.gpu


def test_batch_edge_ranges_vs_oracle(gctx):
    with wwlib.Batch(gctx, RANGES) as b:
        b.launch()
        got = b.read()
    for (lo, hi), (prp, cks, recs) in zip(RANGES, got):
        want = oracle.range_results(lo, hi)
        assert (prp, cks, recs) == want, (lo, hi)


def test_batch_ragged_mix_vs_oracle(gctx):
    rng = random.Random(17)
    ranges = []
    for _ in range(300):
        bits = rng.randint(8, 64)
        lo = max(127, rng.getrandbits(bits))
        hi = min(lo + rng.choice([1, 2, 61, 8191, 8192, 8193, 20000, 70000]), (1 << 64) - 1)
        if hi > lo:
            ranges.append((lo, hi))
    with wwlib.Batch(gctx, ranges) as b:
        b.launch()
        got = b.read()
    for (lo, hi), g in zip(ranges, got):
        assert g == oracle.range_results(lo, hi), (lo, hi)


def test_batch_workload_equals_plain_per_range(gctx):
    """The 4096-range workload (seed 0). Per-range checksums and counts bit-identical to
    plain per-range launches; 64 of the narrow ones also against the oracle."""
    ranges = wwlib.batch_workload()
    with wwlib.Batch(gctx, ranges) as b:
        b.launch()
        got = b.read()
        ms = b.time_ms()
    total = sum(g[0] for g in got)
    print(f"batch: {len(ranges)} ranges, {total} PRPs in {ms:.1f} ms")
    for (lo, hi), g in zip(ranges, got):
        plain = gctx.run_range(lo, hi, 1 << 32)
        assert (g[0], g[1], g[2]) == (plain.prp_count, plain.checksum, plain.results), (lo, hi)
    narrow = sorted(range(len(ranges)), key=lambda i: ranges[i][1] - ranges[i][0])[:64]
    for i in narrow:
        lo, hi = ranges[i]
        assert got[i] == oracle.range_results(lo, hi), (lo, hi)


def test_batch_graph_replays_equal_eager(gctx):
    ranges = wwlib.batch_workload(n=256, seed=5, wmax=24)
    with wwlib.Batch(gctx, ranges) as b:
        b.launch()
        eager = b.read()
        for _ in range(3):
            b.launch(graph=True)
            assert b.read() == eager
            gctx.run_range(10**12, 10**12 + 12345)  # eager work between replays
            b.launch()
            assert b.read() == eager


def test_batch_result_independent_of_order_and_grid(gctx):
    ranges = wwlib.batch_workload(n=512, seed=9, wmax=22)
    with wwlib.Batch(gctx, ranges) as b:
        b.launch()
        ref = b.read()
    perm = list(range(len(ranges)))
    random.Random(1).shuffle(perm)
    with wwlib.Batch(gctx, [ranges[i] for i in perm]) as b:
        for bpc, claim in ((1, 1), (3, 5), (0, 64), (0, 0)):
            b.set_grid(bpc, claim)
            b.launch()
            got = b.read()
            assert [got[perm.index(i)] for i in range(len(ranges))] == ref


def test_batch_single_range_equals_plain_at_full_size(gctx):
    lo, hi = 1 << 63, (1 << 63) + (1 << 33)
    with wwlib.Batch(gctx, [(lo, hi)]) as b:
        b.launch()
        (prp, cks, recs), = b.read()
        ms = b.time_ms()
        pre, search = b.kernel_times_ms()
        b.launch(graph=True)
        assert b.read() == [(prp, cks, recs)]
        gms = b.time_ms()
        gpre, gsearch = b.kernel_times_ms()
    plain = gctx.run_range(lo, hi, 1 << 32)
    print(f"persistent 2^33 at 2^63: {prp} PRPs in {ms:.1f} ms = {prp / ms * 1e3:.3e} /s "
          f"(pre-sieve {pre:.2f} ms, search {search:.2f} ms)")
    assert (prp, cks, recs) == (plain.prp_count, plain.checksum, plain.results)
    # the split times cover the whole launch, eager and graph-launched
    for total, a, s in ((ms, pre, search), (gms, gpre, gsearch)):
        assert a > 0 and s > 0
        assert abs(a + s - total) <= 0.05 + 0.01 * total


def test_autotune_keeps_results(gctx):
    lo, hi = (1 << 62) + 12345, (1 << 62) + 12345 + (1 << 24)
    plain = gctx.run_range(lo, hi)
    try:
        pre0 = gctx.presieve_bound()
        best, trials = gctx.autotune(lo, hi, bounds=[0, 4096], blocks=[0, 4], claims=[1, 16],
                                     presieve=[0, pre0, 1 << 19])
        assert len(trials) == 8 + (2 if pre0 not in (0, 1 << 19) else 1)
        assert all(t["presieve_bound"] == pre0 for t in trials[:8])
        assert {t["presieve_bound"] for t in trials[8:]} == {0, 1 << 19} - {pre0}
        assert best["ms"] == min(t["ms"] for t in trials)
        assert gctx.deep_bound() == best["deep_bound"]
        assert gctx.presieve_bound() == best["presieve_bound"]
        with wwlib.Batch(gctx, [(lo, hi)]) as b:
            b.launch()
            (prp, cks, recs), = b.read()
        assert (prp, cks, recs) == (plain.prp_count, plain.checksum, plain.results)
    finally:
        gctx.set_deep_bound(wwlib.WW_DEFAULT_DEEP_BOUND)
        gctx.set_presieve_bound(wwlib.WW_DEFAULT_PRESIEVE_BOUND)


def test_batch_workload_sample_vs_reference_code(gctx):
    """48 ranges of the 4096-range benchmark workload, in one persistent launch, against
    the reference's own CPU path (tests/golden/workloads.json)."""
    w = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "workloads.json")))["batch"]["reference"]
    with wwlib.Batch(gctx, [(r["lo"], r["hi"]) for r in w]) as b:
        b.launch(graph=True)
        got = b.read()
    for r, (prp, cks, recs) in zip(w, got):
        assert (prp, cks, len(recs)) == (r["prp_count"], r["checksum"],
                                         r["cpu_result_count"]), r["index"]
