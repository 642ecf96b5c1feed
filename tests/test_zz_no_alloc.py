"""The allocation-free entry point ww_search_ranges_noalloc, through the C ABI, against the
batched form (ww_batch_create / launch / read) on the same ranges: per-range checksums and
tested counts, record count and records (sorted; pad = range index) bit-identical, on two
calls in a row, the second with fewer ranges. Device buffers come from the HIP runtime the
library itself uses (ctypes), so the test allocates, the entry point does not."""
import ctypes
import random

import pytest

import wwlib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libww.so is linked against."""

    def __init__(self):
        wwlib.load()
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), max(nbytes, 1)) == 0
        self.bufs.append(p)
        return p

    def put(self, dst, src, nbytes):
        assert self.lib.hipMemcpy(dst, ctypes.cast(src, ctypes.c_void_p), nbytes, H2D) == 0

    def get(self, dst, src, nbytes):
        assert self.lib.hipMemcpy(ctypes.cast(dst, ctypes.c_void_p), src, nbytes, D2H) == 0

    def free_all(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def batched(ctx, ranges, cap):
    with wwlib.Batch(ctx, ranges, result_capacity=cap) as b:
        b.launch(graph=False)
        return b.read()


def noalloc(ctx, hip, ranges, cap):
    n = len(ranges)
    arr = (wwlib.Range * n)(*[wwlib.Range(lo, hi) for lo, hi in ranges])
    d_ranges = hip.alloc(ctypes.sizeof(arr))
    d_sums = hip.alloc(n * ctypes.sizeof(wwlib.RangeSum))
    d_recs = hip.alloc(cap * ctypes.sizeof(wwlib.Result))
    d_count = hip.alloc(4)
    hip.put(d_ranges, arr, ctypes.sizeof(arr))
    rc = wwlib.load().ww_search_ranges_noalloc(ctx.h, d_ranges, n, d_sums, d_recs, cap, d_count)
    assert rc == 0, wwlib.load().ww_last_error()
    ctx.finish()
    sums = (wwlib.RangeSum * n)()
    hip.get(sums, d_sums, ctypes.sizeof(sums))
    count = ctypes.c_uint32()
    hip.get(ctypes.byref(count), d_count, 4)
    m = min(count.value, cap)
    recs = (wwlib.Result * max(m, 1))()
    if m:
        hip.get(recs, d_recs, m * ctypes.sizeof(wwlib.Result))
    out = [(s.prp_count, s.checksum, []) for s in sums]
    for r in recs[:m]:
        out[r.pad][2].append((r.p, r.result, r.rem, r.quot))
    for o in out:
        o[2].sort()
    return out, count.value


def workload(seed, n):
    rng = random.Random(seed)
    ranges = [(127, 5000), (4_294_967_000, 4_294_968_500), (1093 - 2, 3511 + 2),
              ((1 << 64) - 1_000_000, (1 << 64) - 1), (201, 202)]
    while len(ranges) < n:
        w = int(2 ** rng.uniform(10, 22))
        lo = rng.randrange(127, (1 << 64) - w)
        ranges.append((lo, lo + w))
    return ranges[:n]


def test_noalloc_matches_batched_twice_second_smaller():
    hip = Hip()
    cap = 4096
    try:
        with wwlib.Context(0, list_capacity=1 << 16) as ctx:
            for n in (200, 37):  # two calls in a row, the second with fewer ranges
                ranges = workload(n, n)
                want = batched(ctx, ranges, cap)
                got, count = noalloc(ctx, hip, ranges, cap)
                assert count == sum(len(w[2]) for w in want)
                assert got == [(w[0], w[1], sorted(w[2])) for w in want]
                assert got[0][2] and got[2][2]  # records of small p are there
    finally:
        hip.free_all()


def test_noalloc_invalid_ranges_are_empty_and_sizes_are_checked():
    hip = Hip()
    try:
        with wwlib.Context(0, list_capacity=1 << 16) as ctx:
            ranges = [(100, 5000), (10**9, 10**9), (10**12, 10**12 + 100_000)]
            got, _ = noalloc(ctx, hip, ranges, 64)
            assert got[0][:2] == (0, 0) and got[1][:2] == (0, 0)
            want = batched(ctx, ranges[2:], 64)
            assert got[2] == (want[0][0], want[0][1], sorted(want[0][2]))
            lib = wwlib.load()
            d = hip.alloc(64)
            assert lib.ww_search_ranges_noalloc(ctx.h, d, wwlib.WW_NOALLOC_MAX_RANGES + 1, d, d, 0,
                                                d) == wwlib.WW_E_CAPACITY
            assert b"4097 ranges" in lib.ww_last_error()
            assert lib.ww_search_ranges_noalloc(ctx.h, d, 0, d, d, 0, d) == wwlib.WW_E_CAPACITY
            assert lib.ww_search_ranges_noalloc(ctx.h, None, 1, d, d, 0, d) == wwlib.WW_E_ARG
    finally:
        hip.free_all()
