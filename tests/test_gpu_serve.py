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
