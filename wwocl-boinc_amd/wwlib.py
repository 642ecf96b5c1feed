"""ctypes binding of libww.so, the C ABI declared in include/ww_api.h.

The host side of the search path for Python callers (tests, bench.py). It loads the
in-tree library built by build.py next to this file and fails loudly when it is
missing: there is no CPU fallback for the device entry points.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libww.so")
HOST_EXE = os.path.join(HERE, "bin", "wwocl")

WW_OK = 0
WW_E_ARG = -1
WW_E_HIP = -2
WW_E_NODEV = -3
WW_E_CAPACITY = -4
WW_MIN_START = 127
WW_DEFAULT_RESULTS = 15000
WW_DEFAULT_DEEP_BOUND = 2048
WW_DEFAULT_PRESIEVE_BOUND = 1048576
WW_MAX_PRESIEVE_BOUND = 4194304
WW_NOALLOC_MAX_RANGES = 4096

U64 = ctypes.c_uint64
U32 = ctypes.c_uint32


class Result(ctypes.Structure):
    """ww_result: one result record (main.cpp:1458-1493 as one struct)."""

    _fields_ = [("p", U64), ("result", ctypes.c_int32), ("rem", ctypes.c_int32),
                ("quot", ctypes.c_int32), ("pad", ctypes.c_int32)]

    def key(self):
        return (self.p, self.result, self.rem, self.quot)


class Counters(ctypes.Structure):
    """ww_counters: the device counters read at a checkpoint (main.cpp:793-796)."""

    _fields_ = [("checksum", U64), ("totalprimecount", U64), ("primecount", U32),
                ("primecount_max", U32), ("resultcount", U32), ("reserved", U32)]


class Range(ctypes.Structure):
    _fields_ = [("lo", U64), ("hi", U64)]


class RangeSum(ctypes.Structure):
    _fields_ = [("checksum", U64), ("prp_count", U64)]


class DeepPrime(ctypes.Structure):
    _fields_ = [("ell", U32), ("m", U32), ("rho", U32), ("c32_m", U32), ("bar_ell", U32),
                ("pad", U32), ("bar_m", U64)]


class WidePrime(ctypes.Structure):
    _fields_ = [("ell", U32), ("pad", U32), ("m", U64), ("bar_ell", U64), ("bar_m", U64)]


class Tuning(ctypes.Structure):
    _fields_ = [("deep_bound", U32), ("blocks_per_cu", ctypes.c_int32), ("claim", U32),
                ("presieve_bound", U32), ("ms", ctypes.c_double),
                ("prps_per_s", ctypes.c_double), ("waves_per_group", U32),
                ("segment_bits", U32)]

    def as_dict(self):
        return {"deep_bound": self.deep_bound, "blocks_per_cu": self.blocks_per_cu,
                "claim": self.claim, "presieve_bound": self.presieve_bound, "ms": self.ms,
                "prps_per_s": self.prps_per_s, "waves_per_group": self.waves_per_group,
                "segment_bits": self.segment_bits}


class TuneSpace(ctypes.Structure):
    _fields_ = [("deep_bounds", ctypes.POINTER(U32)), ("n_deep_bounds", U32),
                ("blocks_per_cu", ctypes.POINTER(ctypes.c_int32)), ("n_blocks_per_cu", U32),
                ("claims", ctypes.POINTER(U32)), ("n_claims", U32),
                ("presieve_bounds", ctypes.POINTER(U32)), ("n_presieve_bounds", U32),
                ("waves_per_group", ctypes.POINTER(U32)), ("n_waves_per_group", U32),
                ("segment_bits", ctypes.POINTER(U32)), ("n_segment_bits", U32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", U32), ("tick_segments", U32), ("hbm_budget", U64),
                ("bytes_per_request", U64), ("claim", U32), ("blocks_per_cu", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("id", U64), ("lo", U64), ("hi", U64), ("checksum", U64), ("prp_count", U64),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("nrecords", U32), ("pad", U32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("t_start", ctypes.c_double), ("submitted", U64), ("completed", U64),
                ("tested", U64), ("ticks", U64), ("segments", U64), ("pending", U64),
                ("active", U64), ("reserved_bytes", U64), ("peak_reserved_bytes", U64),
                ("peak_active", U64)]


class BucketInfo(ctypes.Structure):
    _fields_ = [("spans", U32), ("active", U32), ("bucket_cap", U32), ("overflows", U32),
                ("bytes", U64)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int), ("clock_khz", ctypes.c_int),
                ("global_mem", U64), ("warp_size", ctypes.c_int)]


class WWError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libww error {code}: {msg}")
        self.code = code


# (name, restype, argtypes) of every entry point of include/ww_api.h
_SIGNATURES = [
    ("ww_last_error", ctypes.c_char_p, []),
    ("ww_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("ww_device_info", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(DeviceInfo)]),
    ("ww_open", ctypes.c_int, [ctypes.c_int, U64, U32, ctypes.POINTER(ctypes.c_void_p)]),
    ("ww_close", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_list_bound", U64, [U64, U64]),
    ("ww_set_deep_bound", ctypes.c_int, [ctypes.c_void_p, U32]),
    ("ww_get_deep_bound", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U32)]),
    ("ww_set_presieve_bound", ctypes.c_int, [ctypes.c_void_p, U32]),
    ("ww_get_presieve_bound", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U32)]),
    ("ww_clear_results", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_clear_prp_count", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_getsegprimes", ctypes.c_int, [ctypes.c_void_p, U64, U64]),
    ("ww_wieferich_wallsunsun", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_search_range", ctypes.c_int, [ctypes.c_void_p, U64, U64, U64]),
    ("ww_finish", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_read_counters", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Counters)]),
    ("ww_read_results", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Result), U32,
                                       ctypes.POINTER(U32)]),
    ("ww_read_prps", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U64), U64,
                                    ctypes.POINTER(U64)]),
    ("ww_kernel_times", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(U64)]),
    ("ww_residues", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U64), U64,
                                   ctypes.POINTER(U64)]),
    ("ww_prp_flags", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U64), U64,
                                    ctypes.POINTER(ctypes.c_uint8)]),
    ("ww_batch_create", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Range), U32, U32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    ("ww_batch_set_grid", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, U32]),
    ("ww_batch_launch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("ww_batch_launch_graph", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("ww_batch_read", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(RangeSum),
                                     ctypes.POINTER(Result), U32, ctypes.POINTER(U32),
                                     ctypes.POINTER(U32)]),
    ("ww_batch_time", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_double)]),
    ("ww_batch_kernel_times", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]),
    ("ww_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_autotune", ctypes.c_int, [ctypes.c_void_p, U64, U64, ctypes.POINTER(TuneSpace),
                                   ctypes.POINTER(Tuning), ctypes.POINTER(Tuning), U32,
                                   ctypes.POINTER(U32)]),
    ("ww_tuning_save", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("ww_tuning_load", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(Tuning)]),
    ("ww_set_launch_shape", ctypes.c_int, [ctypes.c_void_p, U32, U32]),
    ("ww_get_launch_shape", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
    ("ww_server_create", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerConfig),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    ("ww_server_submit", ctypes.c_int, [ctypes.c_void_p, U64, U64, ctypes.POINTER(U64)]),
    ("ww_server_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RequestInfo), U32,
                                      ctypes.POINTER(U32), ctypes.c_int]),
    ("ww_server_records", ctypes.c_int, [ctypes.c_void_p, U64, ctypes.POINTER(Result), U32,
                                         ctypes.POINTER(U32)]),
    ("ww_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    ("ww_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("ww_peak_mul32", ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]),
    ("ww_copy_probe", ctypes.c_int, [ctypes.c_int, U64, ctypes.POINTER(ctypes.c_double)]),
    ("ww_cpu_is_prime", ctypes.c_int, [U64]),
    ("ww_cpu_is_prp", ctypes.c_int, [U64]),
    ("ww_cpu_residues", ctypes.c_int, [ctypes.POINTER(U64), U64, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(U64)]),
    ("ww_cpu_spsp2", ctypes.c_int, [U64, ctypes.c_int32, ctypes.c_int32]),
    ("ww_search_fused", ctypes.c_int, [ctypes.c_void_p, U64, U64]),
    ("ww_search_ranges_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, U32,
                                                ctypes.c_void_p, ctypes.c_void_p, U32,
                                                ctypes.c_void_p]),
    ("ww_cpu_range", ctypes.c_int, [U64, U64, ctypes.POINTER(U64)]),
    ("ww_set_bucket_sieve", ctypes.c_int, [ctypes.c_void_p, U64]),
    ("ww_get_bucket_sieve", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(U64)]),
    ("ww_set_bucket_capacity", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("ww_prime_table", ctypes.c_int, [ctypes.c_void_p, U64, U64, ctypes.POINTER(U32),
                                      ctypes.POINTER(U64)]),
    ("ww_prime_orders", ctypes.c_int, [ctypes.c_void_p, U64, U64, ctypes.POINTER(U32),
                                       ctypes.POINTER(U64)]),
    ("ww_cpu_bucket_hits", ctypes.c_int, [U64, U64, U32, ctypes.POINTER(U64),
                                          ctypes.POINTER(ctypes.c_uint8), U64,
                                          ctypes.POINTER(U64)]),
    ("ww_cpu_order2", ctypes.c_int, [U32, ctypes.POINTER(U32)]),
    ("ww_batch_buckets", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(BucketInfo)]),
    ("ww_deep_table", U32, [U32, ctypes.POINTER(DeepPrime), U32]),
    ("ww_presieve_table", U32, [U32, U32, ctypes.POINTER(WidePrime), U32]),
]

EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libww.so (once). Raises FileNotFoundError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != WW_OK:
        msg = load().ww_last_error()
        raise WWError(rc, msg.decode() if msg else "")


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(load().ww_device_count(ctypes.byref(n)))
    return n.value


def device_info(device: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    _check(load().ww_device_info(device, ctypes.byref(info)))
    return info


def peak_mul32(device: int = 0, kind: int = 0) -> float:
    """Measured issue rate (ops/s) of v_mad_u64_u32 (kind 0) or v_mul_hi_u32 (kind 1)."""
    v = ctypes.c_double(0)
    _check(load().ww_peak_mul32(device, kind, ctypes.byref(v)))
    return v.value


def copy_probe(bytes_: int, device: int = 0) -> float:
    """Duration (ms) of one plain device-to-device copy kernel of `bytes_` bytes."""
    v = ctypes.c_double(0)
    _check(load().ww_copy_probe(device, bytes_, ctypes.byref(v)))
    return v.value


def list_bound(low: int, high: int) -> int:
    return load().ww_list_bound(low, high)


def cpu_is_prime(n: int) -> bool:
    return bool(load().ww_cpu_is_prime(n))


def cpu_is_prp(n: int) -> bool:
    return bool(load().ww_cpu_is_prp(n))


def deep_table(bound: int):
    """[(ell, m, rho, c32_m, bar_ell, bar_m)] of the deep sieve up to `bound` (host-side,
    no device)."""
    cap = 8192
    buf = (DeepPrime * cap)()
    n = load().ww_deep_table(bound, buf, cap)
    return [(buf[i].ell, buf[i].m, buf[i].rho, buf[i].c32_m, buf[i].bar_ell, buf[i].bar_m)
            for i in range(n)]


def presieve_table(above: int, bound: int):
    """[(ell, m, bar_ell, bar_m)] of the wide pre-sieve, above < ell <= bound (host-side)."""
    cap = 300000
    buf = (WidePrime * cap)()
    n = load().ww_presieve_table(above, bound, buf, cap)
    return [(buf[i].ell, buf[i].m, buf[i].bar_ell, buf[i].bar_m) for i in range(n)]


def cpu_residues(ps, top_w: int = -1, top_f: int = -1, wss_skip: int = -1):
    """Host build of the device ladders: [(s0, s1, c20, c21)] per p, from the given
    leading-bit bounds (-1 = each p's own) and Wall-Sun-Sun table steps (5 or 6, -1 = the
    most each p allows)."""
    n = len(ps)
    pin = (U64 * max(n, 1))(*ps)
    out = (U64 * (4 * max(n, 1)))()
    _check(load().ww_cpu_residues(pin, n, top_w, top_f, wss_skip, out))
    return [tuple(out[4 * i:4 * i + 4]) for i in range(n)]


def cpu_spsp2(n: int, top: int = -1, tmax: int = -1) -> bool:
    """Host build of the device's windowed strong PRP test with the given bounds."""
    r = load().ww_cpu_spsp2(n, top, tmax)
    if r < 0:
        _check(r)
    return bool(r)


def cpu_bucket_hits(g0: int, nbits: int, ell: int, cap: int = 1 << 16):
    """Host build of the bucket sieve's walk of one prime: [(bit, in pseudoprime class)]
    for the multiples ell*k of [g0, g0 + 2 nbits) with k prime to 30030 x 17 x 19 x ... x 53."""
    bits = (U64 * max(cap, 1))()
    resc = (ctypes.c_uint8 * max(cap, 1))()
    n = U64()
    _check(load().ww_cpu_bucket_hits(g0, nbits, ell, bits, resc, cap, ctypes.byref(n)))
    if n.value > cap:
        raise WWError(WW_E_CAPACITY, f"{n.value} hits, buffer {cap}")
    return [(bits[i], bool(resc[i])) for i in range(n.value)]


def cpu_order2(ell: int) -> int:
    """Host build of the order table's arithmetic: ord_ell(2) for a prime 2^20 < ell < 2^32."""
    out = U32()
    _check(load().ww_cpu_order2(ell, ctypes.byref(out)))
    return out.value


def cpu_range(low: int, high: int):
    out = (U64 * 3)()
    _check(load().ww_cpu_range(low, high, out))
    return tuple(out)


@dataclass
class RangeResult:
    prp_count: int
    checksum: int
    results: list  # sorted (p, result, rem, quot), Wieferich before Wall-Sun-Sun


def sort_results(recs):
    """Order of the reference's results file: by p, Wieferich (1093) first for equal p
    (stable bubble sort over records appended W-kernel-first, main.cpp:843-854)."""
    return sorted(recs, key=lambda r: (r[0], 0 if r[1] == 1093 else 1))


class Context:
    """One device context: stream, counters, PRP list and result records."""

    def __init__(self, device: int = 0, list_capacity: int = 1 << 24,
                 result_capacity: int = WW_DEFAULT_RESULTS):
        self.lib = load()
        self.h = ctypes.c_void_p()
        _check(self.lib.ww_open(device, list_capacity, result_capacity, ctypes.byref(self.h)))
        self.list_capacity = list_capacity
        self.result_capacity = result_capacity

    def close(self):
        if self.h:
            self.lib.ww_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bucket_sieve(self, min_width: int):
        """Ranges at least min_width wide get the bucket sieve in batches made afterwards
        (2**64 - 1: none)."""
        _check(self.lib.ww_set_bucket_sieve(self.h, min_width))

    def set_bucket_capacity(self, scale: float):
        """Scale of the bucket capacities of batches made afterwards (below 1: overflows,
        for testing the exact fallback)."""
        _check(self.lib.ww_set_bucket_capacity(self.h, scale))

    def bucket_sieve(self) -> int:
        v = U64()
        _check(self.lib.ww_get_bucket_sieve(self.h, ctypes.byref(v)))
        return v.value

    def prime_table(self, first: int, n: int):
        """(entries [first, first + n) of the device's odd-prime table, its length)."""
        out = (U32 * max(n, 1))()
        total = U64()
        _check(self.lib.ww_prime_table(self.h, first, n, out, ctypes.byref(total)))
        return list(out[:n]), total.value

    def prime_orders(self, first: int, n: int):
        """(entries [first, first + n) of the device's order table, its length): ord_l(2)
        of each table prime l above 2^20, 0 below."""
        out = (U32 * max(n, 1))()
        total = U64()
        _check(self.lib.ww_prime_orders(self.h, first, n, out, ctypes.byref(total)))
        return list(out[:n]), total.value

    def set_deep_bound(self, bound: int):
        _check(self.lib.ww_set_deep_bound(self.h, bound))

    def deep_bound(self) -> int:
        v = U32()
        _check(self.lib.ww_get_deep_bound(self.h, ctypes.byref(v)))
        return v.value

    def set_presieve_bound(self, bound: int):
        _check(self.lib.ww_set_presieve_bound(self.h, bound))

    def presieve_bound(self) -> int:
        v = U32()
        _check(self.lib.ww_get_presieve_bound(self.h, ctypes.byref(v)))
        return v.value

    def autotune(self, lo: int, hi: int, bounds=None, blocks=None, claims=None, presieve=None,
                 waves=None, segment_bits=None):
        """Tune the persistent form on [lo, hi); returns (best, [trials]) as dicts. A None
        list takes the library's built-in search space for that knob."""
        keep = []

        def arr(t, v):
            if not v:
                return None, 0
            a = (t * len(v))(*v)
            keep.append(a)
            return ctypes.cast(a, ctypes.POINTER(t)), len(v)
        sp = TuneSpace()
        sp.deep_bounds, sp.n_deep_bounds = arr(U32, bounds)
        sp.blocks_per_cu, sp.n_blocks_per_cu = arr(ctypes.c_int32, blocks)
        sp.claims, sp.n_claims = arr(U32, claims)
        sp.presieve_bounds, sp.n_presieve_bounds = arr(U32, presieve)
        sp.waves_per_group, sp.n_waves_per_group = arr(U32, waves)
        sp.segment_bits, sp.n_segment_bits = arr(U32, segment_bits)
        best = Tuning()
        cap = 512
        trials = (Tuning * cap)()
        n = U32()
        _check(self.lib.ww_autotune(self.h, lo, hi, ctypes.byref(sp), ctypes.byref(best), trials,
                                    cap, ctypes.byref(n)))
        return best.as_dict(), [trials[i].as_dict() for i in range(min(n.value, cap))]

    def set_launch_shape(self, waves_per_group: int, segment_bits: int):
        _check(self.lib.ww_set_launch_shape(self.h, waves_per_group, segment_bits))

    def launch_shape(self):
        w, b = U32(), U32()
        _check(self.lib.ww_get_launch_shape(self.h, ctypes.byref(w), ctypes.byref(b)))
        return w.value, b.value

    def tuning_save(self, path: str):
        _check(self.lib.ww_tuning_save(self.h, str(path).encode()))

    def tuning_load(self, path: str) -> dict:
        t = Tuning()
        _check(self.lib.ww_tuning_load(self.h, str(path).encode(), ctypes.byref(t)))
        return t.as_dict()

    # reference kernel-level entry points
    def clear_results(self):
        _check(self.lib.ww_clear_results(self.h))

    def clear_prp_count(self):
        _check(self.lib.ww_clear_prp_count(self.h))

    def search_fused(self, low: int, high: int):
        _check(self.lib.ww_search_fused(self.h, low, high))

    def getsegprimes(self, low: int, high: int):
        _check(self.lib.ww_getsegprimes(self.h, low, high))

    def wieferich_wallsunsun(self):
        _check(self.lib.ww_wieferich_wallsunsun(self.h))

    def search_range(self, low: int, high: int, chunk: int = 0):
        _check(self.lib.ww_search_range(self.h, low, high, chunk))

    def finish(self):
        _check(self.lib.ww_finish(self.h))

    def counters(self) -> Counters:
        c = Counters()
        _check(self.lib.ww_read_counters(self.h, ctypes.byref(c)))
        return c

    def results(self):
        cap = self.result_capacity
        buf = (Result * cap)()
        n = U32(0)
        _check(self.lib.ww_read_results(self.h, buf, cap, ctypes.byref(n)))
        return [buf[i].key() for i in range(n.value)]

    def prps(self):
        n = U64(0)
        _check(self.lib.ww_read_prps(self.h, None, 0, ctypes.byref(n)))
        if n.value == 0:
            return []
        buf = (U64 * n.value)()
        _check(self.lib.ww_read_prps(self.h, buf, n.value, ctypes.byref(n)))
        return list(buf)

    def kernel_times(self):
        a, b, n = ctypes.c_double(), ctypes.c_double(), U64()
        _check(self.lib.ww_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def residues(self, ps):
        n = len(ps)
        if n == 0:
            return []
        pin = (U64 * n)(*ps)
        out = (U64 * (4 * n))()
        _check(self.lib.ww_residues(self.h, pin, n, out))
        return [tuple(out[4 * i:4 * i + 4]) for i in range(n)]

    def prp_flags(self, ns):
        n = len(ns)
        if n == 0:
            return []
        nin = (U64 * n)(*ns)
        out = (ctypes.c_uint8 * n)()
        _check(self.lib.ww_prp_flags(self.h, nin, n, out))
        return [bool(v) for v in out]

    # convenience: one whole range, counters cleared first
    def run_range_fused(self, low: int, high: int, chunk: int = 1 << 32) -> RangeResult:
        """run_range with each chunk through ww_search_fused (no PRP list)."""
        self.clear_results()
        self.clear_prp_count()
        s = low
        while s < high:
            e = min(high, s + chunk)
            self.search_fused(s, e)
            s = e
        self.finish()
        c = self.counters()
        if c.resultcount > self.result_capacity:
            raise WWError(WW_E_CAPACITY, f"{c.resultcount} results overflowed the buffer")
        return RangeResult(c.totalprimecount, c.checksum, sort_results(self.results()))

    def run_range(self, low: int, high: int, chunk: int = 0) -> RangeResult:
        self.clear_results()
        self.clear_prp_count()
        self.search_range(low, high, chunk)
        self.finish()
        c = self.counters()
        if c.resultcount > self.result_capacity:
            raise WWError(WW_E_CAPACITY, f"{c.resultcount} results overflowed the buffer")
        return RangeResult(c.totalprimecount, c.checksum, sort_results(self.results()))


class Batch:
    """Many ranges searched by one persistent launch (ww_batch_*)."""

    def __init__(self, ctx: Context, ranges, result_capacity: int = 1 << 16):
        self.ctx = ctx
        self.lib = ctx.lib
        self.ranges = [(int(lo), int(hi)) for lo, hi in ranges]
        self.n = len(self.ranges)
        arr = (Range * self.n)(*[Range(lo, hi) for lo, hi in self.ranges])
        self.h = ctypes.c_void_p()
        self.result_capacity = result_capacity
        _check(self.lib.ww_batch_create(ctx.h, arr, self.n, result_capacity, ctypes.byref(self.h)))

    def set_grid(self, blocks_per_cu: int, claim: int = 0):
        _check(self.lib.ww_batch_set_grid(self.h, blocks_per_cu, claim))

    def launch(self, graph: bool = False):
        fn = self.lib.ww_batch_launch_graph if graph else self.lib.ww_batch_launch
        _check(fn(self.ctx.h, self.h))

    def time_ms(self) -> float:
        v = ctypes.c_double()
        _check(self.lib.ww_batch_time(self.ctx.h, self.h, ctypes.byref(v)))
        return v.value

    def kernel_times_ms(self):
        """(pre-sieve kernel, search kernel) milliseconds of the last launch."""
        pre, search = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.ww_batch_kernel_times(self.ctx.h, self.h, ctypes.byref(pre),
                                              ctypes.byref(search)))
        return pre.value, search.value

    def buckets(self) -> dict:
        """The batch's bucket sieve: ranges covered, whether launches use it, bucket
        capacity, device bytes and overflowed chunks so far."""
        bi = BucketInfo()
        _check(self.lib.ww_batch_buckets(self.ctx.h, self.h, ctypes.byref(bi)))
        return {"spans": bi.spans, "active": bool(bi.active), "bucket_cap": bi.bucket_cap,
                "overflows": bi.overflows, "bytes": bi.bytes}

    def read(self):
        """[(prp_count, checksum, sorted records)] per range, in table order."""
        sums = (RangeSum * self.n)()
        recs = (Result * self.result_capacity)()
        nrec, total = U32(), U32()
        _check(self.lib.ww_batch_read(self.ctx.h, self.h, sums, recs, self.result_capacity,
                                      ctypes.byref(nrec), ctypes.byref(total)))
        if total.value > self.result_capacity:
            raise WWError(WW_E_CAPACITY, f"{total.value} records overflowed the buffer")
        per = [[] for _ in range(self.n)]
        for i in range(nrec.value):
            r = recs[i]
            per[r.pad].append(r.key())
        return [(sums[i].prp_count, sums[i].checksum, sort_results(per[i]))
                for i in range(self.n)]

    def close(self):
        if self.h:
            self.lib.ww_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------
# Synthetic workloads. Neither the reference nor the configuration fixes a random
# generator, so these use splitmix64 (Steele, Lea & Flood 2014) with the stated seeds.
# ---------------------------------------------------------------------------------

class SplitMix64:
    def __init__(self, seed: int):
        self.x = seed & (2**64 - 1)

    def next(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)

    def unit(self) -> float:
        return (self.next() >> 11) / float(1 << 53)


def batch_workload(n: int = 4096, seed: int = 0, wmin: int = 16, wmax: int = 28,
                   smin: int = 1 << 40):
    """n ranges: widths floor(2^(wmin + (wmax-wmin) u)) (log-uniform), starts uniform in
    [smin, 2^64 - 2^wmax] and forced odd."""
    rng = SplitMix64(seed)
    smax = (1 << 64) - (1 << wmax)
    out = []
    for _ in range(n):
        w = int(2.0 ** (wmin + (wmax - wmin) * rng.unit()))
        s = (smin + rng.next() % (smax - smin + 1)) | 1
        out.append((s, s + w))
    return out


def serving_workload(rate: float, duration_s: float, seed: int = 1, wmin: int = 20,
                     wmax: int = 30, smin: int = 1 << 32):
    """Poisson arrivals at `rate` per second for `duration_s`: [(t_arrival, lo, hi)],
    widths log-uniform in [2^wmin, 2^wmax], starts uniform in [smin, 2^64 - 2^wmax], odd."""
    import math
    rng = SplitMix64(seed)
    smax = (1 << 64) - (1 << wmax)
    t, out = 0.0, []
    while True:
        t += -math.log(1.0 - rng.unit()) / rate
        if t >= duration_s:
            return out
        w = int(2.0 ** (wmin + (wmax - wmin) * rng.unit()))
        s = (smin + rng.next() % (smax - smin + 1)) | 1
        out.append((t, s, s + w))


class Server:
    """Continuous-batching server (ww_server_*): submit ranges, poll completions."""

    def __init__(self, ctx: Context, slots: int = 0, tick_segments: int = 0,
                 hbm_budget: int = 0, bytes_per_request: int = 0, claim: int = 0,
                 blocks_per_cu: int = -1):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = ServerConfig(slots, tick_segments, hbm_budget, bytes_per_request, claim,
                           blocks_per_cu)
        self.h = ctypes.c_void_p()
        _check(self.lib.ww_server_create(ctx.h, ctypes.byref(cfg), ctypes.byref(self.h)))

    def submit(self, lo: int, hi: int) -> int:
        rid = U64()
        _check(self.lib.ww_server_submit(self.h, lo, hi, ctypes.byref(rid)))
        return rid.value

    def poll(self, wait_ms: int = 0, cap: int = 1024):
        """Completed requests: list of dicts (id, lo, hi, checksum, prp_count, times,
        sorted records)."""
        buf = (RequestInfo * cap)()
        n = U32()
        _check(self.lib.ww_server_poll(self.h, buf, cap, ctypes.byref(n), wait_ms))
        out = []
        for i in range(n.value):
            r = buf[i]
            recs = []
            if r.nrecords:
                rb = (Result * r.nrecords)()
                m = U32()
                _check(self.lib.ww_server_records(self.h, r.id, rb, r.nrecords, ctypes.byref(m)))
                recs = sort_results([rb[k].key() for k in range(m.value)])
            out.append({"id": r.id, "lo": r.lo, "hi": r.hi, "checksum": r.checksum,
                        "prp_count": r.prp_count, "t_submit": r.t_submit,
                        "t_admit": r.t_admit, "t_done": r.t_done, "records": recs})
        return out

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(self.lib.ww_server_get_stats(self.h, ctypes.byref(st)))
        return st

    def close(self):
        if self.h:
            self.lib.ww_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(server: Server, requests, realtime: bool = True, progress_s: float = 0.0):
    """Submit (t_arrival, lo, hi) requests at their arrival times (seconds from now) and
    collect every completion. Returns (completions by id, {id: (lo, hi)}, wall seconds).
    progress_s > 0 prints a progress line to stderr that often."""
    import sys
    import time
    t0 = time.perf_counter()
    ids, done = {}, {}
    i = 0
    next_note = progress_s
    while i < len(requests) or len(done) < len(ids):
        now = time.perf_counter() - t0
        if progress_s > 0 and now >= next_note:
            print(f"serve: {now:.0f} s, {i} of {len(requests)} submitted, {len(done)} done",
                  file=sys.stderr, flush=True)
            next_note += progress_s
        while i < len(requests) and (not realtime or requests[i][0] <= now):
            _, lo, hi = requests[i]
            ids[server.submit(lo, hi)] = (lo, hi)
            i += 1
        wait = 0
        if i < len(requests) and realtime:
            wait = max(0, int((requests[i][0] - (time.perf_counter() - t0)) * 1000))
        for c in server.poll(wait_ms=min(wait, 20) if wait else 2):
            done[c["id"]] = c
    return done, ids, time.perf_counter() - t0


def latency_summary(done):
    lat = sorted(c["t_done"] - c["t_submit"] for c in done.values())
    if not lat:
        return {}
    def pct(q):
        return lat[min(len(lat) - 1, int(q * len(lat)))]
    return {"p50_ms": pct(0.50) * 1e3, "p99_ms": pct(0.99) * 1e3, "max_ms": lat[-1] * 1e3,
            "mean_ms": sum(lat) / len(lat) * 1e3}
