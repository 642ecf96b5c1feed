"""Throughput of the MI355X search path: tested primes/s at p ~ 2^63.

One step = the whole range [2^63, 2^63 + 2^33) (196,705,536 tested primes) through the HIP
path, inputs generated on the device, the checksum accumulated on the device. The headline
`value` is the persistent form: the sieve stage (the bucket sieve's fill and split, then the
wide pre-sieve k_presieve, which applies the buckets), then one graph-launched k_batch
(sieve by the small primes, strong PRP of the candidates not known to be prime, Wieferich
and Wall-Sun-Sun in one kernel). The plain form (per 2^32-wide chunk: clear the list, sieve
+ PRP of every candidate, then the test kernel) runs beside it and must give the same step
checksum. Also: the 4096-range batched workload, the autotuner, the serving form's latency,
and the reference's own CPU path on a bounded sample. Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-batch]
                    [--form both|persistent|plain] [--serve-seconds 10]
                    [--serve-rates 20,50,100] [--no-tune] [--copy-probe BYTES]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wwocl-boinc_amd"))

import wwlib  # noqa: E402

LO = 1 << 63
HI = (1 << 63) + (1 << 33)
CHUNK = 1 << 32
BITS = 64  # every p in the range has bit length 64
# HBM bytes per list entry moved by k_ww_tests: FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per
# launch (factors from a 1 GiB copy kernel) over the entries of the launch, measured with
# rocprofv3 --pmc on MI355X for this workload: 806.0 MB fetched + 17.5 MB written per 2^32
# chunk (the 8-byte list read, register spills and the result atomics).
TRAFFIC_BYTES_PER_ENTRY = 8.37
# k_batch keeps candidates and PRPs on chip (LDS queues). Its algorithmic HBM traffic is the
# pre-sieve's two sets of 64 words (512 B each) per 8192-integer segment, read once: the
# odd numbers to drop and those that still need the strong PRP test. The rest is register
# spill reloads, the deep-sieve table, span lookups and the cursor/slot atomics. Measured
# the same way at the defaults (deep bound 2048, pre-sieve 2^20, bucket sieve on, 4-wave
# workgroups): see profiles/k_batch_persistent_2p63_w2p33.md.
BATCH_TRAFFIC_BYTES_PER_PRP = 6.540
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, calibrated on a 1 GiB copy "
                  "(profiles/k_batch_persistent_2p63_w2p33.md)")
PRESIEVE_BYTES_PER_SEGMENT = 1024  # drop words + needs-PRP words
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def mul32_path(b: int) -> int:
    """Algorithmic work per tested prime of bit length b, in 32x32->64 multiplies (MUL32),
    as the reference's algorithm spends it (the roofline's per-unit figure): Wieferich
    21(b-3)+14, Wall-Sun-Sun 68(b-1)+14 (two squares and one product per bit), the prime's
    own PRP 10(b-2); A(b) = 99b - 123. Composite candidates and the sieve are not counted."""
    return 99 * b - 123


def mul32_tests(b: int) -> int:
    """The same for the Wieferich + Wall-Sun-Sun kernel alone: 89b - 103."""
    return 89 * b - 103


def mul32_path_executed(b: int) -> int:
    """MUL32 this implementation's algorithm needs per tested prime, counted the same way
    (64x64->128 = 4, a low or high half = 3 or 4, a 64-bit square = 3; mod-p^2 Montgomery
    square 21, product 26): Wieferich 21(b-2)+14; Wall-Sun-Sun on V(3,1), one square and one
    product per bit, 47(b-2)+45; the prime's own PRP at least 10(b-4). = 78b - 117."""
    return 78 * b - 117


# The box's CPU share for one GPU: worker pools are sized to at most this many threads
# whatever the host reports (nproc and os.cpu_count() show the whole machine there).
CPU_SHARE = 16
CPU_SLICE3 = 1 << 22  # the 2^63 leg: 16 slices of 2^22 at offsets k*2^29


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The reference's own CPU path (wwcpu.cpp + prime.h compiled into oracle/_ref, the
    restated loop of main.cpp:648-697 around them; the oracle's C port when that is not
    built) on the three samples of the plan: the whole range [127, 1e8), the slice
    [2^50, 2^50 + 2^24), and 16 slices [2^63 + k 2^29, +2^22) of the workload range, each
    split into contiguous pieces over the threads. `value` is the 2^63 leg's rate (the
    metric's domain); every leg is reported with its own rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    threads = max(1, min(allowed, CPU_SHARE))
    kind = "reference" if oracle.ref_lib() is not None else "port"

    def run(lo, hi):
        if kind == "reference":
            return oracle.ref_range(lo, hi, threads)[0]
        return oracle.range_summary(lo, hi, threads)[0]

    legs = {}
    t0 = time.perf_counter()
    prp = run(127, 10**8)
    dt = time.perf_counter() - t0
    legs["range_127_1e8"] = {"prps": prp, "seconds": dt, "value": prp / dt}
    t0 = time.perf_counter()
    prp = run(1 << 50, (1 << 50) + (1 << 24))
    dt = time.perf_counter() - t0
    legs["slice_2p50_w2p24"] = {"prps": prp, "seconds": dt, "value": prp / dt}
    t0 = time.perf_counter()
    prp = 0
    for k in range(16):
        lo = LO + k * (1 << 29)
        prp += run(lo, lo + CPU_SLICE3)
    dt = time.perf_counter() - t0
    legs["slices_2p63_16x2p22"] = {"prps": prp, "seconds": dt, "value": prp / dt}
    head = legs["slices_2p63_16x2p22"]
    return {"value": head["value"], "unit": "tested primes/s", "cores": threads, "kind": kind,
            "sample": f"16 slices [2^63 + k*2^29, +2^22), k=0..15: {head['prps']} PRPs in "
                      f"{head['seconds']:.2f} s on {threads} threads",
            "nproc": os.cpu_count(), "affinity_cpus": allowed, "cpu_share": CPU_SHARE,
            "cpu_model": cpu_model(), "legs": legs}


def run_plain(ctx, steps, warmup):
    """Plain form: per 2^32 chunk, clear the list, sieve + PRP, then the test kernel."""
    step_checksum = None
    for _ in range(max(warmup, 1)):  # at least one untimed step gives the step checksum
        ctx.clear_results()
        ctx.search_range(LO, HI, CHUNK)
        ctx.finish()
        step_checksum = ctx.counters().checksum
    ctx.kernel_times()  # discard warmup kernel times
    ctx.clear_results()
    ctx.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.search_range(LO, HI, CHUNK)
    ctx.finish()
    elapsed = time.perf_counter() - t0
    ms_sieve, ms_tests, launches = ctx.kernel_times()
    c = ctx.counters()
    tested = c.totalprimecount
    per_launch = tested / launches
    avg_tests_s = ms_tests / launches * 1e-3
    return {"value": tested / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "prps_per_step": tested // steps,
            "step_checksum": None if step_checksum is None else f"{step_checksum:016X}",
            "kernel_ms_per_step": {"k_sieve_prp": ms_sieve / steps, "k_ww_tests": ms_tests / steps},
            "tests_kernel": {"avg_launch_ms": avg_tests_s * 1e3, "launches": launches,
                             "mul32_per_s": per_launch * mul32_tests(BITS) / avg_tests_s,
                             "entries_per_launch": per_launch}}


def run_persistent(ctx, steps, warmup):
    """Persistent form: the whole range as two graphs, the sieve stage (k_bucket_fill,
    k_bucket_split, k_presieve) and k_batch."""
    with wwlib.Batch(ctx, [(LO, HI)]) as b:
        res = None
        for _ in range(max(warmup, 1)):
            b.launch(graph=True)
            res = b.read()[0]
        ctx.finish()
        t0 = time.perf_counter()
        kms = pre_ms = search_ms = 0.0
        for _ in range(steps):
            b.launch(graph=True)
            kms += b.time_ms()
            pre, search = b.kernel_times_ms()
            pre_ms += pre
            search_ms += search
        ctx.finish()
        elapsed = time.perf_counter() - t0
        last = b.read()[0]
        buckets = b.buckets()
    assert last[:2] == res[:2], "persistent form is not deterministic"
    assert buckets["active"] and buckets["overflows"] == 0, buckets
    prps, cks = res[0], res[1]
    return {"value": prps * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "prps_per_step": prps, "step_checksum": f"{cks:016X}",
            "kernel_ms_per_step": kms / steps,
            "kernels_ms_per_step": {"sieve (k_bucket_fill + k_bucket_split + k_presieve)":
                                    pre_ms / steps, "k_batch": search_ms / steps},
            "bucket_sieve": {"bucket_bytes": buckets["bytes"],
                             "overflowed_chunks": buckets["overflows"]}}


def run_batched_workload(ctx):
    """The 4096-range workload (seed 0): one persistent launch vs plain per-range launches
    (clear, chunk loop, read back), both on the same context."""
    ranges = wwlib.batch_workload()
    with wwlib.Batch(ctx, ranges) as b:
        b.launch(graph=True)
        ref = b.read()
        ctx.finish()
        t0 = time.perf_counter()
        b.launch(graph=True)
        kms = b.time_ms()
        got = b.read()
        t_batch = time.perf_counter() - t0
    assert got == ref
    prps = sum(g[0] for g in got)
    t0 = time.perf_counter()
    plain_prps = 0
    for (lo, hi), g in zip(ranges, got):
        ctx.clear_results()
        ctx.clear_prp_count()
        ctx.search_range(lo, hi, CHUNK)
        c = ctx.counters()
        plain_prps += c.totalprimecount
        assert (c.totalprimecount, c.checksum) == (g[0], g[1]), (lo, hi)
    t_plain = time.perf_counter() - t0
    return {"ranges": len(ranges), "prps": prps,
            "persistent": {"value": prps / t_batch, "seconds": t_batch, "kernel_ms": kms},
            "plain_per_range": {"value": plain_prps / t_plain, "seconds": t_plain},
            "speedup": t_plain / t_batch, "per_range_checksums_equal": True}


def run_serving(ctx, rates, seconds):
    """Serving form: Poisson request streams (seed 1, widths 2^20..2^30, starts in
    [2^32, 2^64 - 2^30]) submitted in real time against the default 64 GiB HBM budget;
    p50/p99 latency (submit to result) and tested primes/s; 16 requests per rate re-run in
    the plain form must agree. The stated configuration runs 10 minutes per rate
    (--serve-only --serve-seconds 600); the default bench run is a short one."""
    out = {}
    for rate in rates:
        reqs = wwlib.serving_workload(rate=rate, duration_s=seconds, seed=1)
        with wwlib.Server(ctx) as srv:
            done, ids, wall = wwlib.serve(srv, reqs, realtime=True,
                                          progress_s=60.0 if seconds > 60 else 0.0)
            st = srv.stats()
        tested = sum(c["prp_count"] for c in done.values())
        span = max(c["t_done"] for c in done.values()) - min(c["t_submit"] for c in done.values())
        rec = {"requests": len(done), "tested": tested, "tested_per_s": tested / span,
               "offered_req_per_s": rate, "seconds": seconds,
               "run": "stated (10 min)" if seconds >= 600 else f"short ({seconds:g} s)",
               "ticks": st.ticks, "peak_active": st.peak_active,
               "peak_reserved_bytes": st.peak_reserved_bytes}
        rec.update(wwlib.latency_summary(done))
        for rid in sorted(ids)[:: max(1, len(ids) // 16)][:16]:
            lo, hi = ids[rid]
            ctx.clear_results()
            ctx.clear_prp_count()
            ctx.search_range(lo, hi, CHUNK)
            c = ctx.counters()
            assert (c.totalprimecount, c.checksum) == (done[rid]["prp_count"],
                                                       done[rid]["checksum"]), (lo, hi)
        rec["checked_vs_plain"] = 16
        out[str(rate)] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-range workload")
    ap.add_argument("--form", choices=["both", "persistent", "plain"], default="both")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of each serving run (0 = skip serving)")
    ap.add_argument("--serve-rates", default="20,50,100", help="request rates (per second)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (library defaults: deep bound 2048, pre-sieve "
                         "bound 2^20, full occupancy)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also launch one plain copy kernel of BYTES (counter calibration)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving legs (--serve-rates, --serve-seconds) and print "
                         "their JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(ranges shard trivially across GPUs: one sub-range per GPU, checksums summed)", file=sys.stderr)
        sys.exit(2)

    info = wwlib.device_info(0)
    if args.serve_only:
        with wwlib.Context(0, list_capacity=wwlib.list_bound(LO, LO + CHUNK)) as ctx:
            serving = run_serving(ctx, [float(r) for r in args.serve_rates.split(",")],
                                  args.serve_seconds)
        print(json.dumps({"metric": "serving p50/p99 latency (Poisson, seed 1)",
                          "device": info.name.decode(), "serving": serving}), flush=True)
        return
    peak = wwlib.peak_mul32(0, 0)
    cap = wwlib.list_bound(LO, LO + CHUNK)
    copy_ms = wwlib.copy_probe(args.copy_probe) if args.copy_probe else None
    plain = pers = batched = tuning = None
    with wwlib.Context(0, list_capacity=cap) as ctx:
        if not args.no_tune:
            t0 = time.perf_counter()
            best, trials = ctx.autotune(LO, LO + (1 << 31))
            tuning = {"sample": "[2^63, 2^63+2^31)", "best": best, "trials": len(trials),
                      "seconds": time.perf_counter() - t0,
                      "by_deep_bound": {str(b): max(t["prps_per_s"] for t in trials
                                                    if t["deep_bound"] == b)
                                        for b in sorted({t["deep_bound"] for t in trials})},
                      "by_presieve_bound": {
                          str(b): max(t["prps_per_s"] for t in trials
                                      if t["presieve_bound"] == b)
                          for b in sorted({t["presieve_bound"] for t in trials})},
                      "by_waves_per_group": {
                          str(w): max(t["prps_per_s"] for t in trials
                                      if t["waves_per_group"] == w)
                          for w in sorted({t["waves_per_group"] for t in trials})},
                      "by_segment_bits": {
                          str(b): max(t["prps_per_s"] for t in trials
                                      if t["segment_bits"] == b)
                          for b in sorted({t["segment_bits"] for t in trials})}}
        if args.form in ("both", "persistent"):
            pers = run_persistent(ctx, args.steps, args.warmup)
        if args.form in ("both", "plain"):
            plain = run_plain(ctx, args.steps, args.warmup)
        if not args.no_batch:
            batched = run_batched_workload(ctx)
        serving = None
        if args.serve_seconds > 0:
            serving = run_serving(ctx, [float(r) for r in args.serve_rates.split(",")],
                                  args.serve_seconds)

    head = pers if pers is not None else plain
    out = {
        "metric": "tested primes/s at p~2^63 (batched vs plain, 1 GPU)",
        "value": head["value"],
        "unit": "tested primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the number line [2^63, 2^63+2^33), generated on the device",
        "config": {"workload": ("persistent form (sieve stage + one k_batch, graph-launched)"
                                if pers else
                                "plain form (2^32-wide chunks)") +
                               ", range [2^63, 2^63+2^33)",
                   "range": [LO, HI], "prps_per_step": head["prps_per_step"],
                   "step_checksum": head["step_checksum"], "device": info.name.decode(),
                   "arch": info.arch.decode(), "cus": info.compute_units},
    }
    if pers is not None:
        kern_s = pers["kernels_ms_per_step"]["k_batch"] * 1e-3
        ach = pers["prps_per_step"] * mul32_path(BITS) / kern_s
        out["roofline"] = {
            "bound": "valu", "kernel": "k_batch", "achieved": ach / 1e12, "peak": peak / 1e12,
            "unit": "TMUL32/s", "frac": ach / peak,
            "traffic": pers["prps_per_step"] * BATCH_TRAFFIC_BYTES_PER_PRP,
            "traffic_algorithmic": ((HI - LO) // 8192) * PRESIEVE_BYTES_PER_SEGMENT,
            "traffic_algorithmic_note": "the pre-sieve's drop words and needs-PRP words, "
                                        "512 B each per 8192-integer segment",
            "traffic_source": TRAFFIC_SOURCE,
            "per_unit": f"{mul32_path(BITS)} MUL32 per tested prime (A(b) = 99b-123 at b=64: "
                        "PRP + Wieferich + Wall-Sun-Sun in one kernel, counted as the "
                        "reference's algorithm spends them; this build proves most primes "
                        "by sieving to sqrt(hi) and runs the PRP test only where that "
                        "cannot)",
            "executed_per_unit": f"{mul32_path_executed(BITS)} MUL32 per tested prime "
                                 "(78b-117: this build's ladders)",
            "avg_launch_ms": pers["kernels_ms_per_step"]["k_batch"], "launches": args.steps,
            "step_kernels_ms": pers["kernels_ms_per_step"]}
        # the same kernel against HBM: measured bytes per launch over the launch time
        hbm = pers["prps_per_step"] * BATCH_TRAFFIC_BYTES_PER_PRP / kern_s
        out["roofline_hbm"] = {
            "bound": "hbm", "kernel": "k_batch", "achieved": hbm / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": hbm / 1e9 / HBM_PEAK_GBS,
            "traffic": pers["prps_per_step"] * BATCH_TRAFFIC_BYTES_PER_PRP,
            "note": "not the bound: the kernel keeps candidates and PRPs on chip and reads "
                    "only the pre-sieve words (1 KB per 8192 integers) from HBM"}
    if plain is not None:
        tk = plain["tests_kernel"]
        key = "roofline" if pers is None else "roofline_plain_tests_kernel"
        out[key] = {
            "bound": "valu", "kernel": "k_ww_tests", "achieved": tk["mul32_per_s"] / 1e12,
            "peak": peak / 1e12, "unit": "TMUL32/s", "frac": tk["mul32_per_s"] / peak,
            "traffic": tk["entries_per_launch"] * TRAFFIC_BYTES_PER_ENTRY,
            "traffic_algorithmic": tk["entries_per_launch"] * 8,
            "per_unit": f"{mul32_tests(BITS)} MUL32 per tested prime (89b-103, b=64)",
            "avg_launch_ms": tk["avg_launch_ms"], "launches": tk["launches"]}
        out["plain"] = {k: plain[k] for k in ("value", "ms_per_step", "prps_per_step",
                                              "step_checksum", "kernel_ms_per_step")}
    if pers is not None and plain is not None:
        out["persistent_vs_plain"] = pers["value"] / plain["value"]
        assert pers["step_checksum"] == plain["step_checksum"], "forms disagree"
    if batched is not None:
        out["batched_workload"] = batched
    if tuning is not None:
        out["tuning"] = tuning
    if serving is not None:
        out["serving"] = serving
    if copy_ms is not None:
        out["copy_probe"] = {"bytes": args.copy_probe, "ms": copy_ms}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
