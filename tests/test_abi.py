"""The C ABI library: it loads, it exports every entry point include/ww_api.h
declares, and its host-side checks agree with the oracle. No device calls."""
import os
import random
import re

import pytest

import oracle
import wwlib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ww_api.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ww_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = wwlib.load()
    declared = header_functions()
    assert len(declared) >= 20
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(wwlib.EXPORTED) == declared


C_CALLER = r"""
#include <stdio.h>
#include "ww_api.h"
/* a C caller of the ABI: host-side entry points only (no device needed) */
int main(void)
{
    uint64_t out[3], r[4];
    uint64_t p = 1093;
    if (ww_cpu_range(127, 100000, out) != WW_OK) return 1;
    if (ww_cpu_residues(&p, 1, -1, -1, -1, r) != WW_OK) return 2;
    if (ww_cpu_spsp2(4294967297ull, -1, -1) != 1) return 3;
    if (ww_list_bound(1ull << 40, (1ull << 40) + 1000000) == 0) return 4;
    printf("%llu %llu %llu %llu\n", (unsigned long long)out[0], (unsigned long long)out[1],
           (unsigned long long)r[0], (unsigned long long)r[1]);
    return 0;
}
"""


def test_header_is_plain_c_and_links(tmp_path):
    """include/ww_api.h compiles as C11 (plain C types only) and a C program links and runs
    against libww.so."""
    import subprocess
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    lib_dir = os.path.join(ROOT, "wwocl-boinc_amd")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", str(exe), str(src), "-L", lib_dir, "-lww", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    prp, cks = oracle.range_summary(127, 100000)[:2]
    s0, s1 = oracle.wieferich(1093)
    assert r.stdout.split() == [str(prp), str(cks), str(s0), str(s1)]


def test_no_device_is_reported_cleanly():
    if wwlib.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(wwlib.WWError) as e:
        wwlib.Context(0, list_capacity=1024)
    assert e.value.code == wwlib.WW_E_NODEV


def test_cpu_checks_match_oracle():
    rng = random.Random(1)
    ns = [42799, 49141, 2047, 3277, 49981, 1093, 3511, 18446744073709551557,
          (1 << 64) - 1, 3825123056546413051, 3825123056546413049]
    ns += [rng.getrandbits(rng.randint(8, 64)) | 1 for _ in range(5000)]
    for n in ns:
        if n < 3:
            continue
        assert wwlib.cpu_is_prime(n) == oracle.is_prime(n), n
        assert wwlib.cpu_is_prp(n) == oracle.is_prp(n), n


@pytest.mark.parametrize("lo,hi", [(127, 300_000), ((1 << 32) - 100_000, (1 << 32) + 100_000),
                                   (1 << 63, (1 << 63) + 200_000), (495_700, 495_900),
                                   ((1 << 64) - 200_000, (1 << 64) - 1)])
def test_cpu_range_matches_oracle(lo, hi):
    """ww_cpu_range is the startup self-test's CPU side, written independently of the
    device arithmetic (128-bit products reduced by division, a Fibonacci-pair ladder):
    pinned here against the oracle, including 495791 (the one list member on which the
    reference's single-digit Wall-Sun-Sun phase runs) and the top of the domain."""
    prp, cks, _, flags = oracle.range_summary(lo, hi)
    assert wwlib.cpu_range(lo, hi) == (prp, cks, flags)


def test_list_bound_is_an_upper_bound():
    for lo, hi in [(127, 128), (127, 10**6), (10**12, 10**12 + 10**6),
                   (1 << 63, (1 << 63) + (1 << 20)), ((1 << 64) - 1000, (1 << 64) - 1)]:
        n = len(oracle.prp_list(lo, hi))
        assert wwlib.list_bound(lo, hi) >= n
    assert wwlib.list_bound(5, 5) == 0
