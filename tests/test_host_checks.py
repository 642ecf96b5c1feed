"""The drop-in host's failure detection, driven on the CPU (no GPU): the validation unit
wwocl-boinc_amd/host/wwocl_checks.cpp gets injected counters, records and PRP lists
through a fake device view (tests/host_checks_main.cpp), built with AddressSanitizer and
UndefinedBehaviorSanitizer. Each branch is the reference's:

- composite record that is no base-2 PRP: abort (main.cpp:912-916);
- composite base-2 PRP record: dropped silently (main.cpp:907-911);
- more records than slots: abort (main.cpp:811-814);
- list longer than its buffer: abort (main.cpp:802-805);
- unknown result code: abort (main.cpp:896-899);
- checkPRPS on a corrupted list: abort (main.cpp:477-516);
- self-test mismatches: abort (main.cpp:703-728).

Expected lists come from the oracle; the messages and their streams from main.cpp."""
import os
import shutil
import subprocess

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRCS = [os.path.join(ROOT, "tests", "host_checks_main.cpp"),
        os.path.join(ROOT, "wwocl-boinc_amd", "host", "wwocl_checks.cpp"),
        os.path.join(ROOT, "wwocl-boinc_amd", "csrc", "ww_selftest.cpp")]

# a few records the reference writes (tests.txt and the known Wieferich primes)
T3 = 2276306935816523    # Wieferich special instance (-1 -3 p)
T4 = 338772621946054253  # WallSunSun special instance (+0 +91 p)
SPSP = 42799             # 127 * 337: a base-2 strong pseudoprime with no factor <= 113
COMPOSITE = 1000001      # 101 * 9901: not a base-2 PRP


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("hc") / "host_checks"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I",
           os.path.join(ROOT, "include"), "-o", str(out), *SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(out)


def call(driver, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([driver, *map(str, args)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    return r.returncode, r.stdout, r.stderr


def records_file(tmp_path, recs):
    p = tmp_path / "recs.txt"
    p.write_text("".join(f"{a} {b} {c} {d}\n" for a, b, c, d in recs))
    return p


def list_file(tmp_path, xs):
    p = tmp_path / "list.txt"
    p.write_text("".join(f"{x}\n" for x in xs))
    return p


def test_fixture_numbers():
    assert not oracle.is_prime(SPSP) and oracle.is_prp(SPSP)
    assert not oracle.is_prime(COMPOSITE) and not oracle.is_prp(COMPOSITE)
    assert oracle.is_prime(T3) and oracle.is_prime(T4)


def test_records_sorted_formatted_and_pseudoprimes_dropped(driver, tmp_path):
    recs = [(T4, 2, 0, 91), (3511, 1093, 1, 0), (SPSP, 1093, 1, 5), (T3, 1093, -1, -3),
            (1093, 2, 0, -7), (1093, 1093, 1, 0), (SPSP, 2, 0, 17)]
    f = records_file(tmp_path, recs)
    rc, out, _ = call(driver, "results", f, len(recs), 10, 100, 15000)
    assert rc == 0, out
    lines = out.splitlines()
    assert lines[0] == "OK dropped=2 checksum=0123456789ABCDEF"
    assert lines[1:] == [
        "1093 is a Wieferich prime",
        "1093 is a WallSunSun special instance (+0 -7 p)",
        "3511 is a Wieferich prime",
        "2276306935816523 is a Wieferich special instance (-1 -3 p)",
        "338772621946054253 is a WallSunSun special instance (+0 +91 p)",
    ]


def test_composite_result_aborts(driver, tmp_path):
    f = records_file(tmp_path, [(1093, 1093, 1, 0), (COMPOSITE, 2, 0, 3)])
    rc, out, _ = call(driver, "results", f, 2, 0, 100, 15000)
    assert rc == 3
    assert out == f"FATAL both=0\nComputation error, GPU found a composite result: {COMPOSITE}\n"


def test_result_overflow_aborts(driver, tmp_path):
    f = records_file(tmp_path, [])
    rc, out, _ = call(driver, "results", f, 15001, 0, 100, 15000)
    assert rc == 3 and out == "FATAL both=0\nError: number of results (15001) overflowed array.\n"


def test_prime_array_overflow_aborts(driver, tmp_path):
    f = records_file(tmp_path, [])
    rc, out, _ = call(driver, "results", f, 0, 101, 100, 15000)
    assert rc == 3 and out == "FATAL both=0\nError: prime array overflow!\n"
    rc, out, _ = call(driver, "results", f, 0, 100, 100, 15000)  # at capacity: fine
    assert rc == 0


def test_unknown_result_code_aborts(driver, tmp_path):
    f = records_file(tmp_path, [(1093, 7, 1, 0)])
    rc, out, _ = call(driver, "results", f, 1, 0, 100, 15000)
    assert rc == 3 and out == "FATAL both=0\nError: hostResult returned 7.\n"


def test_failed_device_read_aborts(driver):
    rc, out, _ = call(driver, "read_failure")
    assert rc == 3 and out == "FATAL both=1\nERROR: read counters: injected read failure\n"


LO, HI = 42000, 60000  # holds the pseudoprimes 42799 and 49141


def test_check_prps_accepts_the_true_list(driver, tmp_path):
    good = oracle.prp_list(LO, HI)
    assert SPSP in good and 49141 in good
    rc, out, _ = call(driver, "prps", list_file(tmp_path, good[::-1]), LO, HI)
    assert rc == 0 and out == f"OK prps={len(good)}\n"


def test_check_prps_composite_entry_aborts(driver, tmp_path):
    bad = oracle.prp_list(LO, HI)
    bad[10] = COMPOSITE
    rc, out, _ = call(driver, "prps", list_file(tmp_path, bad), LO, HI)
    assert rc == 3 and out == f"FATAL both=0\nComputation error, GPU found a composite: {COMPOSITE}\n"


def test_check_prps_missing_prime_aborts(driver, tmp_path):
    good = oracle.prp_list(LO, HI)
    primes = [p for p in good if oracle.is_prime(p)]
    bad = [p for p in good if p != primes[3]]
    rc, out, _ = call(driver, "prps", list_file(tmp_path, bad), LO, HI)
    assert rc == 3
    assert out == (f"FATAL both=1\nERROR: gpu prime count: {len(primes) - 1} does not match "
                   f"cpu: {len(primes)}\n")


def test_check_prps_wrong_prime_aborts(driver, tmp_path):
    good = oracle.prp_list(LO, HI)
    primes = sorted(p for p in good if oracle.is_prime(p))
    bad = list(good)
    bad[bad.index(primes[0])] = 1009  # a prime outside the range
    rc, out, _ = call(driver, "prps", list_file(tmp_path, bad), LO, HI)
    assert rc == 3
    assert out == f"FATAL both=1\nERROR: gpu/cpu prime mismatch 1009 vs {primes[0]}\n"


def test_selftest_mismatches_abort(driver):
    rc, out, _ = call(driver, "selftest", 100, 5, 2, 100, 5, 2)
    assert rc == 0 and out == "OK\n"
    rc, out, _ = call(driver, "selftest", 99, 5, 2, 100, 5, 2)
    assert rc == 3 and out == "FATAL both=1\nERROR: gpu / cpu prp count mismatch 99 / 100\n"
    rc, out, _ = call(driver, "selftest", 100, 6, 2, 100, 5, 2)
    assert rc == 3 and out == "FATAL both=1\nERROR: gpu / cpu checksum mismatch 6 / 5\n"
    rc, out, _ = call(driver, "selftest", 100, 5, 3, 100, 5, 2)
    assert rc == 3 and out == "FATAL both=1\nERROR: gpu / cpu result count mismatch 3 / 2\n"
    rc, out, _ = call(driver, "selftest_fused", 100, 5, 2, 100, 5, 2)
    assert rc == 0
    rc, out, _ = call(driver, "selftest_fused", 100, 5, 2, 100, 7, 2)
    assert rc == 3 and "ERROR: persistent kernel / cpu mismatch" in out


def test_selftest_catches_a_corrupted_device_checksum(driver):
    """The self-test's CPU side against a device result with one wrong residue slice: the
    CPU triple comes from ww_cpu_range's own formulation (via the oracle's equal values)."""
    prp, cks, _, flags = oracle.range_summary(10**12, 10**12 + 200000)
    rc, _, _ = call(driver, "selftest", prp, cks, flags, prp, cks, flags)
    assert rc == 0
    rc, out, _ = call(driver, "selftest", prp, (cks + 1) % (1 << 64), flags, prp, cks, flags)
    assert rc == 3 and "checksum mismatch" in out


def test_program_hook_exits_with_failure_on_both_streams(driver):
    rc, out, err = call(driver, "default_hook")
    assert rc == 1
    msg = "Error: injected failure through the program's own hook.\n"
    assert out == msg and err == msg
