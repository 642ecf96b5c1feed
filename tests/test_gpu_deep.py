"""The deep sieve on the device, at every bound, in both forms: the PRP list, checksums
and records are those of the oracle (the sieve only thins composite candidates)."""
import random

import pytest

import oracle
import wwlib
from test_deep_sieve import KNOWN_SPSP2

pytestmark = pytest.mark.gpu

BOUNDS = [0, 127, 1024, 4096, 8192, 20011, 65535]
# strong pseudoprimes (base 2) with a factor above 16384, which the wide pre-sieve must keep:
# 3215031751 = 151 * 751 * 28351, 2152302898747 = 6763 * 10627 * 29947,
# 3474749660383 = 1303 * 16927 * 157543
WIDE_SPSP2 = (3215031751, 2152302898747, 3474749660383)


@pytest.fixture
def deep_ctx(gctx):
    deep, pre = gctx.deep_bound(), gctx.presieve_bound()
    yield gctx
    gctx.set_deep_bound(deep)
    gctx.set_presieve_bound(pre)


@pytest.mark.parametrize("bound", BOUNDS)
def test_list_and_ranges_at_bound(deep_ctx, bound):
    ctx = deep_ctx
    ctx.set_deep_bound(bound)
    assert ctx.deep_bound() == bound
    # windows of 32 segments around WIDE_SPSP2, so that the plain form pre-sieves them too
    windows = [(n - 5000, n + 5000) for n in KNOWN_SPSP2[7:]]
    windows += [(n - (1 << 17), n + (1 << 17)) for n in WIDE_SPSP2]
    for lo, hi in [(127, 3_000_000)] + windows:
        ctx.clear_results()
        ctx.clear_prp_count()
        ctx.getsegprimes(lo, hi)
        ctx.finish()
        got = sorted(ctx.prps())
        want = oracle.prp_list(lo, hi)
        assert got == want, (bound, lo, hi)
    rng = random.Random(bound)
    for _ in range(12):
        lo = rng.randrange(1 << 20, (1 << 64) - (1 << 20))
        hi = lo + rng.choice([8191, 8192, 100_000, 1 << 19])
        g = ctx.run_range(lo, hi)
        assert (g.prp_count, g.checksum, g.results) == oracle.range_results(lo, hi), (lo, hi)


@pytest.mark.parametrize("bound", [0, 8192, 65535])
def test_batch_at_bound(deep_ctx, bound):
    ctx = deep_ctx
    ctx.set_deep_bound(bound)
    ranges = [(n - 3000, n + 3000) for n in KNOWN_SPSP2] + [(127, 200_000)]
    ranges += wwlib.batch_workload(n=64, seed=3, wmax=20)
    with wwlib.Batch(ctx, ranges) as b:
        b.launch()
        got = b.read()
        b.launch(graph=True)
        assert b.read() == got
    for (lo, hi), g in zip(ranges, got):
        assert g == oracle.range_results(lo, hi), (bound, lo, hi)


@pytest.mark.parametrize("deep,pre", [(0, 1 << 22), (4096, 1 << 22), (16384, 1 << 20),
                                      (65535, 1 << 22), (16384, 20011)])
def test_batch_with_wide_presieve(deep_ctx, deep, pre):
    ctx = deep_ctx
    ctx.set_deep_bound(deep)
    ctx.set_presieve_bound(pre)
    assert (ctx.deep_bound(), ctx.presieve_bound()) == (deep, pre)
    for n in WIDE_SPSP2:
        assert oracle.is_prp(n) and not oracle.is_prime(n)
    ranges = [(n - 3000, n + 3000) for n in KNOWN_SPSP2] + [(127, 300_000), (127, 4_000_000)]
    ranges += [(3215031751 - (1 << 21), 3215031751 + (1 << 21))]  # whole pre-sieve chunks
    # 16 and 15 segments: the shortest piece that is pre-sieved, and one that is not
    ranges += [(n - (1 << 16), n + (1 << 16)) for n in WIDE_SPSP2]
    ranges += [(n - 61440, n + 61440) for n in WIDE_SPSP2]
    ranges += wwlib.batch_workload(n=48, seed=5, wmax=21)
    ranges += [(1 << 63, (1 << 63) + (1 << 22)), ((1 << 64) - (1 << 22), (1 << 64) - 1)]
    with wwlib.Batch(ctx, ranges) as b:
        b.launch(graph=True)
        got = b.read()
        b.launch()
        assert b.read() == got
    for (lo, hi), g in zip(ranges, got):
        assert g == oracle.range_results(lo, hi), (deep, pre, lo, hi)


def test_graphs_and_servers_survive_a_presieve_table_rebuild():
    """Raising the pre-sieve bound past what the context has built replaces its device
    table. A batch's captured graph is captured again, and a running server keeps its own
    copy of the tables, so both still give the oracle's results afterwards."""
    ranges = [(3215031751 - (1 << 21), 3215031751 + (1 << 21)),
              ((1 << 63) + 12345, (1 << 63) + 12345 + (1 << 22))]
    want = [oracle.range_results(lo, hi) for lo, hi in ranges]
    with wwlib.Context(0, list_capacity=1 << 20) as ctx:
        assert ctx.presieve_bound() == wwlib.WW_DEFAULT_PRESIEVE_BOUND < wwlib.WW_MAX_PRESIEVE_BOUND
        with wwlib.Batch(ctx, ranges) as b, wwlib.Server(ctx, tick_segments=1024) as srv:
            b.launch(graph=True)
            assert b.read() == want
            ctx.set_presieve_bound(wwlib.WW_MAX_PRESIEVE_BOUND)  # builds a larger table
            ctx.set_presieve_bound(wwlib.WW_DEFAULT_PRESIEVE_BOUND)
            b.launch(graph=True)
            assert b.read() == want
            done, ids, _ = wwlib.serve(srv, [(0.0, lo, hi) for lo, hi in ranges],
                                       realtime=False)
        for rid, (lo, hi) in ids.items():
            c = done[rid]
            assert (c["prp_count"], c["checksum"], c["records"]) == want[ranges.index((lo, hi))]
