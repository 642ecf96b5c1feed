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


def _restore_defaults(ctx):
    ctx.set_deep_bound(wwlib.WW_DEFAULT_DEEP_BOUND)
    ctx.set_presieve_bound(wwlib.WW_DEFAULT_PRESIEVE_BOUND)
    ctx.set_launch_shape(4, 20)


def test_autotune_stages_and_tuned_results_vs_oracle(gctx):
    """The tuner's stages (deep bound x geometry, work-group shape, pre-sieve bound,
    pre-sieve segment size) each measured, the fastest kept as the context's settings, and
    the tuned configuration's results equal to the oracle's in the persistent, fused and
    plain forms."""
    lo, hi = (1 << 62) + 12345, (1 << 62) + 12345 + (1 << 24)
    want = oracle.range_results(lo, hi)
    try:
        pre0 = gctx.presieve_bound()
        best, trials = gctx.autotune(lo, hi, bounds=[0, 4096], blocks=[0, 4], claims=[1, 16],
                                     presieve=[0, pre0, 1 << 19], waves=[2, 4, 8],
                                     segment_bits=[16, 20])
        assert len(trials[:8]) == 8 and all(t["waves_per_group"] == 4 for t in trials[:8])
        assert {t["waves_per_group"] for t in trials} == {2, 4, 8}
        # 8-wave groups: as many as fit, and 4 * 4 / 8 = 2 per CU
        assert {t["blocks_per_cu"] for t in trials if t["waves_per_group"] == 8} == {0, 2}
        assert {t["presieve_bound"] for t in trials} == {0, pre0, 1 << 19}
        if best["presieve_bound"]:
            assert {16, 20} <= {t["segment_bits"] for t in trials}
        assert best["ms"] == min(t["ms"] for t in trials)
        assert gctx.deep_bound() == best["deep_bound"]
        assert gctx.presieve_bound() == best["presieve_bound"]
        assert gctx.launch_shape() == (best["waves_per_group"], best["segment_bits"])
        with wwlib.Batch(gctx, [(lo, hi)]) as b:
            b.launch()
            assert b.read() == [want]
        gctx.clear_results()
        gctx.search_fused(lo, hi)
        c = gctx.counters()
        assert (c.totalprimecount, c.checksum) == want[:2]
        got = gctx.run_range(lo, hi)
        assert (got.prp_count, got.checksum, got.results) == want
    finally:
        _restore_defaults(gctx)


@pytest.mark.parametrize("waves,bits", [(2, 15), (8, 17), (2, 20), (8, 19)])
def test_launch_shapes_vs_oracle(gctx, waves, bits):
    """Every work-group shape and pre-sieve segment size gives the oracle's results, on the
    edge ranges and a batch of them."""
    from test_gpu_parity import RANGES
    try:
        gctx.set_launch_shape(waves, bits)
        assert gctx.launch_shape() == (waves, bits)
        with wwlib.Batch(gctx, RANGES) as b:
            b.launch(graph=True)
            got = b.read()
        for (lo, hi), g in zip(RANGES, got):
            assert g == oracle.range_results(lo, hi), (lo, hi)
        lo, hi = (1 << 63) - (1 << 23), (1 << 63) + (1 << 23)  # pre-sieved chunks of any size
        gctx.clear_results()
        gctx.search_fused(lo, hi)
        c = gctx.counters()
        want = oracle.range_results(lo, hi)
        assert (c.totalprimecount, c.checksum) == want[:2]
    finally:
        _restore_defaults(gctx)


def test_launch_shape_arguments_checked(gctx):
    for w, bits in [(3, 20), (16, 20), (4, 14), (4, 21)]:
        with pytest.raises(wwlib.WWError) as e:
            gctx.set_launch_shape(w, bits)
        assert e.value.code == wwlib.WW_E_ARG
    assert gctx.launch_shape() == (4, 20)


def test_tuning_cache_round_trip(gctx, tmp_path):
    """ww_tuning_save / ww_tuning_load: a saved configuration applies on a fresh context of
    the same device; files from another device or malformed ones are refused."""
    path = tmp_path / "tune.json"
    try:
        gctx.set_deep_bound(4096)
        gctx.set_presieve_bound(1 << 19)
        gctx.set_launch_shape(8, 18)
        gctx.tuning_save(path)
        text = path.read_text()
        assert '"waves_per_group": 8' in text and '"segment_bits": 18' in text
        with wwlib.Context(0, list_capacity=1 << 16) as ctx:
            t = ctx.tuning_load(path)
            assert (t["deep_bound"], t["presieve_bound"], t["waves_per_group"],
                    t["segment_bits"]) == (4096, 1 << 19, 8, 18)
            assert ctx.deep_bound() == 4096 and ctx.presieve_bound() == 1 << 19
            assert ctx.launch_shape() == (8, 18)
            lo, hi = 10**15, 10**15 + (1 << 22)
            with wwlib.Batch(ctx, [(lo, hi)]) as b:
                b.launch()
                assert b.read() == [oracle.range_results(lo, hi)]
            other = tmp_path / "other.json"
            other.write_text(text.replace('"compute_units": ', '"compute_units": 1'))
            with pytest.raises(wwlib.WWError) as e:
                ctx.tuning_load(other)
            assert e.value.code == wwlib.WW_E_ARG and b"not on this" in wwlib.load().ww_last_error()
            bad = tmp_path / "bad.json"
            bad.write_text(text.replace('"segment_bits": 18', '"segment_bits": 99'))
            with pytest.raises(wwlib.WWError):
                ctx.tuning_load(bad)
            with pytest.raises(wwlib.WWError):
                ctx.tuning_load(tmp_path / "missing.json")
    finally:
        _restore_defaults(gctx)


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
