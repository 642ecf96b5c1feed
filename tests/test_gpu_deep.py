"""The deep sieve on the device, at every bound, in both forms: the PRP list, checksums
and records are those of the oracle (the sieve only thins composite candidates)."""
import random

import pytest

import oracle
import wwlib
from test_deep_sieve import KNOWN_SPSP2

pytestmark = pytest.mark.gpu

BOUNDS = [0, 127, 1024, 4096, 8192, 20011, 65535]


@pytest.fixture
def deep_ctx(gctx):
    yield gctx
    gctx.set_deep_bound(wwlib.WW_DEFAULT_DEEP_BOUND)


@pytest.mark.parametrize("bound", BOUNDS)
def test_list_and_ranges_at_bound(deep_ctx, bound):
    ctx = deep_ctx
    ctx.set_deep_bound(bound)
    assert ctx.deep_bound() == bound
    for lo, hi in [(127, 3_000_000)] + [(n - 5000, n + 5000) for n in KNOWN_SPSP2[7:]]:
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
