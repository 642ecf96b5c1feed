"""Throughput of the MI355X search path: tested primes/s at p ~ 2^63, plain form.

One step = the whole range [2^63, 2^63 + 2^33) through the HIP path (per 2^32-wide
chunk: clear the list, sieve + PRP, Wieferich + Wall-Sun-Sun), inputs generated on the
device, the checksum accumulated on the device. Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
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
# rocprofv3 --pmc on MI355X for this workload. Algorithmic: the 8-byte list read.
TRAFFIC_BYTES_PER_ENTRY = 8.00


def mul32_path(b: int) -> int:
    """Algorithmic 32x32->64 multiplies per tested prime of bit length b over the whole
    path: Wieferich 21(b-3)+14, Wall-Sun-Sun 68(b-1)+14, PRP of the prime 10(b-2)."""
    return 99 * b - 123


def mul32_tests(b: int) -> int:
    """The same for the Wieferich + Wall-Sun-Sun kernel alone."""
    return 89 * b - 103


def cpu_baseline():
    """The reference's own CPU path (compiled into oracle/_ref) on 16 sampled 2^22-wide
    slices of the workload, one slice range per thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = min(16, os.cpu_count() or 1)
    kind = "reference" if oracle.ref_lib() is not None else "port"
    t0 = time.perf_counter()
    total = 0
    for k in range(16):
        lo = LO + k * (1 << 29)
        if kind == "reference":
            prp = oracle.ref_range(lo, lo + (1 << 22), threads)[0]
        else:
            prp = oracle.range_summary(lo, lo + (1 << 22), threads)[0]
        total += prp
    dt = time.perf_counter() - t0
    return {"value": total / dt, "unit": "tested primes/s", "cores": threads, "kind": kind,
            "sample": f"16 slices [2^63 + k*2^29, +2^22), k=0..15: {total} PRPs in {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also launch one plain copy kernel of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(ranges shard trivially across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    info = wwlib.device_info(0)
    peak = wwlib.peak_mul32(0, 0)
    cap = wwlib.list_bound(LO, LO + CHUNK)
    copy_ms = wwlib.copy_probe(args.copy_probe) if args.copy_probe else None
    step_checksum = None
    with wwlib.Context(0, list_capacity=cap) as ctx:
        for _ in range(args.warmup):
            ctx.clear_results()
            ctx.search_range(LO, HI, CHUNK)
            ctx.finish()
            step_checksum = ctx.counters().checksum
        ctx.kernel_times()  # discard warmup kernel times
        ctx.clear_results()
        ctx.finish()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.search_range(LO, HI, CHUNK)
        ctx.finish()
        elapsed = time.perf_counter() - t0
        ms_sieve, ms_tests, launches = ctx.kernel_times()
        c = ctx.counters()

    tested = c.totalprimecount
    per_step = tested // args.steps
    value = tested / elapsed
    # dominant kernel: Wieferich + Wall-Sun-Sun; algorithmic MUL32 per launch over the
    # average launch duration (events on the context's stream)
    avg_tests_s = ms_tests / launches * 1e-3
    per_launch = tested / launches
    ach = per_launch * mul32_tests(BITS) / avg_tests_s
    path_ach = tested * mul32_path(BITS) / elapsed
    out = {
        "metric": "tested primes/s at p~2^63 (plain form, 1 GPU)",
        "value": value,
        "unit": "tested primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the number line [2^63, 2^63+2^33), generated on the device",
        "config": {"workload": "plain form, range [2^63, 2^63+2^33) in 2^32-wide chunks",
                   "range": [LO, HI], "chunk": CHUNK, "prps_per_step": per_step,
                   "step_checksum": None if step_checksum is None else f"{step_checksum:016X}",
                   "device": info.name.decode(),
                   "arch": info.arch.decode(), "cus": info.compute_units},
        "roofline": {"bound": "valu", "kernel": "k_ww_tests",
                     "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TMUL32/s",
                     "frac": ach / peak,
                     "traffic": per_launch * TRAFFIC_BYTES_PER_ENTRY,
                     "traffic_algorithmic": per_launch * 8,
                     "per_unit": f"{mul32_tests(BITS)} MUL32 per tested prime (89b-103, b=64)",
                     "avg_launch_ms": avg_tests_s * 1e3, "launches": launches},
        "roofline_path": {"achieved": path_ach / 1e12, "peak": peak / 1e12, "unit": "TMUL32/s",
                          "frac": path_ach / peak,
                          "per_unit": f"{mul32_path(BITS)} MUL32 per tested prime (99b-123)"},
        "kernel_ms_per_step": {"k_sieve_prp": ms_sieve / args.steps,
                               "k_ww_tests": ms_tests / args.steps},
    }
    if copy_ms is not None:
        out["copy_probe"] = {"bytes": args.copy_probe, "ms": copy_ms}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
