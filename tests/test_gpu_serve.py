"""The serving form (ww_server_*): continuous batching of range requests. Each request's
checksum, count and records are the same alone, in a mix, and in the plain form."""
import json
import os
import random

import pytest

import oracle
import wwlib
from test_gpu_parity import RANGES

pytestmark = pytest.mark.gpu


def test_serve_edge_ranges_vs_oracle(gctx):
    with wwlib.Server(gctx, tick_segments=4096) as srv:
        done, ids, _ = wwlib.serve(srv, [(0.0, lo, hi) for lo, hi in RANGES], realtime=False)
    assert len(done) == len(RANGES)
    for rid, (lo, hi) in ids.items():
        c = done[rid]
        assert (c["prp_count"], c["checksum"], c["records"]) == oracle.range_results(lo, hi)


def test_serve_alone_equals_in_the_mix_equals_plain(gctx):
    mix = wwlib.serving_workload(rate=400.0, duration_s=0.5, seed=1, wmax=26)
    probe = mix[len(mix) // 2][1:]
    with wwlib.Server(gctx, tick_segments=1 << 14) as srv:
        alone, ids_a, _ = wwlib.serve(srv, [(0.0, *probe)], realtime=False)
        mixed, ids_m, _ = wwlib.serve(srv, mix, realtime=True)
    (ca,) = alone.values()
    cm = next(c for rid, c in mixed.items() if ids_m[rid] == probe)
    key = lambda c: (c["prp_count"], c["checksum"], c["records"])
    assert key(ca) == key(cm)
    rng = random.Random(2)
    for rid in rng.sample(sorted(ids_m), 40):
        lo, hi = ids_m[rid]
        plain = gctx.run_range(lo, hi, 1 << 32)
        assert key(mixed[rid]) == (plain.prp_count, plain.checksum, plain.results), (lo, hi)
    # requests overlapped on the device: some were admitted before others finished
    spans = sorted((c["t_admit"], c["t_done"]) for c in mixed.values())
    assert any(a2 < d1 for (a1, d1), (a2, d2) in zip(spans, spans[1:]))


def test_serve_admission_respects_slots(gctx):
    reqs = [(0.0, lo, lo + (1 << 22)) for lo in range(10**15, 10**15 + 64 * (1 << 22), 1 << 22)]
    with wwlib.Server(gctx, slots=8, tick_segments=2048) as srv:
        done, ids, _ = wwlib.serve(srv, reqs, realtime=False)
        st = srv.stats()
    assert st.completed == len(reqs) and st.active == 0 and st.pending == 0
    for rid, (lo, hi) in ids.items():
        plain = gctx.run_range(lo, hi)
        assert (done[rid]["prp_count"], done[rid]["checksum"]) == (plain.prp_count, plain.checksum)


def test_serving_sample_vs_reference_code(gctx):
    """16 requests of the 100 req/s serving stream, submitted together, against the
    reference's own CPU path (tests/golden/workloads.json)."""
    w = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "workloads.json")))["serving"]["reference"]
    with wwlib.Server(gctx) as srv:
        done, ids, _ = wwlib.serve(srv, [(0.0, r["lo"], r["hi"]) for r in w], realtime=False)
    by_range = {ids[rid]: c for rid, c in done.items()}
    for r in w:
        c = by_range[(r["lo"], r["hi"])]
        assert (c["prp_count"], c["checksum"], len(c["records"])) == (
            r["prp_count"], r["checksum"], r["cpu_result_count"]), r["index"]


def test_serve_admission_bound_by_the_hbm_budget(gctx):
    """Each admitted request holds its resident sieve state (the pre-sieve words of all its
    segments, 512 bytes per 8192 integers) in the server's HBM arena until it completes.
    With 64 slots but a budget for three such requests at a time, the budget, not the
    slots, holds requests back; every result still equals the oracle's."""
    w = 1 << 22                   # 512 segments: 256 KiB of words each
    need = (w // 8192) * 512
    reqs = [(0.0, lo, lo + w) for lo in range(3 * 10**14 + 1, 3 * 10**14 + 1 + 12 * w, w)]
    with wwlib.Server(gctx, slots=64, tick_segments=256, hbm_budget=3 * need) as srv:
        done, ids, _ = wwlib.serve(srv, reqs, realtime=False)
        st = srv.stats()
        # a request needing more than the whole budget is refused at once
        with pytest.raises(wwlib.WWError) as e:
            srv.submit(10**12, 10**12 + 4 * w)
        assert e.value.code == wwlib.WW_E_CAPACITY
    assert st.completed == len(reqs) and st.active == 0 and st.reserved_bytes == 0
    assert st.peak_active == 3 and st.peak_reserved_bytes == 3 * need
    # the budget made requests wait: some were admitted only after others had finished
    spans = sorted((c["t_admit"], c["t_done"]) for c in done.values())
    assert sum(1 for a, _ in spans if a > min(d for _, d in spans)) >= len(reqs) - 3
    for rid, (lo, hi) in ids.items():
        c = done[rid]
        assert (c["prp_count"], c["checksum"], c["records"]) == oracle.range_results(lo, hi)


def test_serve_budget_fragments_and_coalesces(gctx):
    """Requests of mixed widths against a tight budget: blocks are freed and merged as
    requests complete, so wide requests behind narrow ones still get admitted."""
    ws = [1 << 20, 1 << 23, 1 << 21, 1 << 23, 1 << 20, 1 << 22, 1 << 23, 1 << 19]
    lo = 5 * 10**13 + 1
    reqs = []
    for w in ws:
        reqs.append((0.0, lo, lo + w))
        lo += w
    budget = ((1 << 23) // 8192) * 512 * 2   # two of the widest at once
    with wwlib.Server(gctx, slots=16, tick_segments=512, hbm_budget=budget) as srv:
        done, ids, _ = wwlib.serve(srv, reqs, realtime=False)
        st = srv.stats()
    assert st.completed == len(reqs) and st.reserved_bytes == 0
    assert st.peak_reserved_bytes <= budget
    for rid, (a, b) in ids.items():
        plain = gctx.run_range(a, b)
        assert (done[rid]["prp_count"], done[rid]["checksum"], done[rid]["records"]) == (
            plain.prp_count, plain.checksum, plain.results)


@pytest.mark.parametrize("stream", ["20.0", "50.0", "100.0_wide"])
def test_serving_streams_vs_reference_code(gctx, stream):
    """Requests of the 20 and 50 req/s streams, and the two narrowest requests of the 100
    req/s stream wider than 2^28, against the reference's own CPU path
    (tests/golden/workloads.json, written by tests/golden/make_workloads.py)."""
    w = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "workloads.json")))["serving_streams"][stream]["reference"]
    if stream.endswith("wide"):
        assert all(r["hi"] - r["lo"] > (1 << 28) for r in w)
    with wwlib.Server(gctx) as srv:
        done, ids, _ = wwlib.serve(srv, [(0.0, r["lo"], r["hi"]) for r in w], realtime=False)
    by_range = {ids[rid]: c for rid, c in done.items()}
    for r in w:
        c = by_range[(r["lo"], r["hi"])]
        assert (c["prp_count"], c["checksum"], len(c["records"])) == (
            r["prp_count"], r["checksum"], r["cpu_result_count"]), r["index"]
