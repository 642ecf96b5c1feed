"""ctypes binding of the CPU oracle (liboracle.so) and of oracle/_ref/libwwref.so.

TEST INFRASTRUCTURE ONLY: the checker. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libwwref.so")
REF_TREE = "/root/reference"

U64 = ctypes.c_uint64


class Summary(ctypes.Structure):
    _fields_ = [("prp_count", U64), ("checksum", U64), ("result_count", U64),
                ("cpu_result_count", U64)]


class Rec(ctypes.Structure):
    _fields_ = [("p", U64), ("result", ctypes.c_int32), ("rem", ctypes.c_int32),
                ("quot", ctypes.c_int32), ("pad", ctypes.c_int32)]


def build(with_ref: bool | None = None):
    """Compile liboracle.so and, when the reference tree is present, _ref/libwwref.so."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(REF_TREE)
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_TREE}"], check=True)


_o = None
_r = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build(with_ref=False)
        o = ctypes.CDLL(ORACLE_SO)
        o.or_invert.restype = U64
        o.or_invert.argtypes = [U64]
        o.or_strong_prp.argtypes = [U64, U64]
        o.or_is_candidate.argtypes = [U64]
        o.or_is_prp.argtypes = [U64]
        o.or_is_prime.argtypes = [U64]
        o.or_m2p_one.argtypes = [U64, ctypes.POINTER(U64)]
        o.or_wieferich.argtypes = [U64, ctypes.POINTER(U64)]
        o.or_wallsunsun.argtypes = [U64, ctypes.POINTER(U64)]
        o.or_wieferich_cpu_flag.argtypes = [U64, ctypes.POINTER(U64)]
        o.or_wallsunsun_cpu_flag.argtypes = [U64, ctypes.POINTER(U64)]
        o.or_find_wheel_offset.argtypes = [ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_int32)]
        o.or_range.restype = U64
        o.or_range.argtypes = [U64, U64, ctypes.POINTER(Summary), ctypes.POINTER(Rec), U64]
        o.or_range_mt.argtypes = [U64, U64, ctypes.c_int, ctypes.POINTER(Summary)]
        o.or_prp_list.restype = U64
        o.or_prp_list.argtypes = [U64, U64, ctypes.POINTER(U64), U64]
        o.or_sieve_primes.restype = U64
        o.or_sieve_primes.argtypes = [U64, U64, ctypes.POINTER(U64), U64]
        _o = o
    return _o


def ref_lib():
    """The reference's own CPU arithmetic (compiled from its sources), or None."""
    global _r
    if _r is None:
        if not os.path.exists(REF_SO):
            return None
        r = ctypes.CDLL(REF_SO)
        r.ref_range.argtypes = [U64, U64, ctypes.c_int, ctypes.POINTER(U64)]
        r.ref_strong_prp2.argtypes = [U64]
        r.ref_wieferich.argtypes = [U64, ctypes.POINTER(U64)]
        r.ref_wallsunsun.argtypes = [U64, ctypes.POINTER(U64)]
        _r = r
    return _r


def strong_prp(base: int, n: int) -> bool:
    return bool(lib().or_strong_prp(base, n))


def is_candidate(n: int) -> bool:
    return bool(lib().or_is_candidate(n))


def is_prp(n: int) -> bool:
    return bool(lib().or_is_prp(n))


def is_prime(n: int) -> bool:
    return bool(lib().or_is_prime(n))


def invert(p: int) -> int:
    return lib().or_invert(p)


def m2p_one(p: int):
    out = (U64 * 2)()
    lib().or_m2p_one(p, out)
    return tuple(out)


def wieferich(p: int):
    out = (U64 * 2)()
    lib().or_wieferich(p, out)
    return tuple(out)


def wallsunsun(p: int):
    out = (U64 * 2)()
    lib().or_wallsunsun(p, out)
    return tuple(out)


def cpu_flags(p: int):
    """(wieferichCPU flag, slice, wallsunsunCPU flag, slice) restated."""
    a, b = U64(0), U64(0)
    fw = lib().or_wieferich_cpu_flag(p, ctypes.byref(a))
    fs = lib().or_wallsunsun_cpu_flag(p, ctypes.byref(b))
    return fw, a.value, fs, b.value


def find_wheel_offset(start: int):
    s = U64(start)
    idx = ctypes.c_int32(0)
    lib().or_find_wheel_offset(ctypes.byref(s), ctypes.byref(idx))
    return s.value, idx.value


def range_summary(lo: int, hi: int, threads: int = 1):
    s = Summary()
    if threads > 1:
        lib().or_range_mt(lo, hi, threads, ctypes.byref(s))
    else:
        lib().or_range(lo, hi, ctypes.byref(s), None, 0)
    return s.prp_count, s.checksum, s.result_count, s.cpu_result_count


def range_results(lo: int, hi: int, cap: int = 1 << 16):
    """(prp_count, checksum, sorted records) over [lo, hi)."""
    s = Summary()
    buf = (Rec * cap)()
    n = lib().or_range(lo, hi, ctypes.byref(s), buf, cap)
    if s.result_count > cap:
        raise RuntimeError("result buffer too small")
    recs = [(buf[i].p, buf[i].result, buf[i].rem, buf[i].quot) for i in range(n)]
    return s.prp_count, s.checksum, recs


def prp_list(lo: int, hi: int):
    n = lib().or_prp_list(lo, hi, None, 0)
    buf = (U64 * max(n, 1))()
    lib().or_prp_list(lo, hi, buf, n)
    return list(buf[:n])


def sieve_primes(lo: int, hi: int):
    n = lib().or_sieve_primes(lo, hi, None, 0)
    if n == 2**64 - 1:
        raise MemoryError("or_sieve_primes allocation failed")
    buf = (U64 * max(n, 1))()
    lib().or_sieve_primes(lo, hi, buf, n)
    return list(buf[:n])


def ref_range(lo: int, hi: int, threads: int = 1):
    r = ref_lib()
    if r is None:
        return None
    out = (U64 * 3)()
    r.ref_range(lo, hi, threads, out)
    return tuple(out)


def ref_slices(p: int):
    r = ref_lib()
    if r is None:
        return None
    a, b = U64(0), U64(0)
    fw = r.ref_wieferich(p, ctypes.byref(a))
    fs = r.ref_wallsunsun(p, ctypes.byref(b))
    return fw, a.value, fs, b.value
