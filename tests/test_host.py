"""The drop-in host program (wwocl-boinc_amd/bin/wwocl): command line, checkpoint files
and results file, against the reference's behaviour (main.cpp) and known answers
(tests.txt). CPU tests stop before the device is opened; GPU tests run whole ranges."""
import os
import subprocess

import pytest

import oracle
import wwlib

EXE = wwlib.HOST_EXE


def run(tmp_path, *args, timeout=600, env=None):
    # the tuning cache stays in the test's directory unless a test names one
    e = dict(os.environ, WW_TUNE_CACHE=str(tmp_path / "no-cache.json"))
    if env:
        e.update(env)
    r = subprocess.run([EXE, *map(str, args)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=timeout, env=e)
    return r.returncode, r.stdout, r.stderr


def write_state(tmp_path, name, start, stop, current, cksm=0, psum=0.0, pcnt=0):
    (tmp_path / name).write_text(f"{start} {stop} {current} {cksm} {psum:f} {pcnt}\n")


def results(tmp_path):
    return (tmp_path / "results-WW.txt").read_text()


def no_gpu():
    return wwlib.device_count() == 0


# ----------------------------------------------------------------------------- CPU


def test_usage_without_arguments(tmp_path):
    rc, out, _ = run(tmp_path)
    assert rc == 1 and "Usage:" in out


def test_end_must_exceed_start(tmp_path):
    rc, out, err = run(tmp_path, "1e3", 1000)
    assert rc == 1 and "END must be > START" in out + err


def test_unparsable_exponent_form(tmp_path):
    rc, out, _ = run(tmp_path, "0e5", 1000)
    assert rc == 1 and "unable to parse START" in out
    rc, out, _ = run(tmp_path, "1e30", "1e31")
    assert rc == 1 and "START is out of range" in out


def test_finished_workunit_resumes_complete(tmp_path):
    """A state file whose current == END: 'Workunit complete.', exit 0 (main.cpp:1544-1551)."""
    write_state(tmp_path, "stateA-WW.txt", 2276300000000000, 2276360000000000,
                2276360000000000, 7)
    rc, out, _ = run(tmp_path, "227630e10", "227636e10")
    assert rc == 0
    assert "Resuming search from checkpoint. Current: 2276360000000000" in out
    assert "Workunit complete." in out


def test_garbage_state_a_falls_back_to_b(tmp_path):
    (tmp_path / "stateA-WW.txt").write_text("garbage\n")
    write_state(tmp_path, "stateB-WW.txt", 1000, 5000, 5000)
    rc, out, err = run(tmp_path, 1000, 5000)
    assert rc == 0 and "Cannot parse stateA-WW.txt" in err and "Workunit complete." in out


def test_newest_state_file_wins(tmp_path):
    write_state(tmp_path, "stateA-WW.txt", 1000, 5000, 5000)
    write_state(tmp_path, "stateB-WW.txt", 1000, 5000, 3000)
    rc, out, _ = run(tmp_path, 1000, 5000)
    assert rc == 0 and "Current: 5000" in out


@pytest.mark.skipif(not no_gpu(), reason="would start a search on the visible GPU")
def test_mismatched_state_starts_fresh_and_truncates_results(tmp_path):
    write_state(tmp_path, "stateA-WW.txt", 1, 2, 2)
    (tmp_path / "results-WW.txt").write_text("old line\n")
    rc, out, _ = run(tmp_path, 1000, 5000)
    assert "Beginning a new search" in out
    assert results(tmp_path) == ""
    assert rc == 1  # no device here


@pytest.mark.skipif(not no_gpu(), reason="would start a search on the visible GPU")
def test_start_clamps_to_127_and_exponent_form(tmp_path):
    rc, out, _ = run(tmp_path, 1, "6e10")
    assert "START below minimum. Starting search at 127" in out
    assert "Starting search at: 127\nStopping search at: 60000000000" in out


def test_device_choice_as_boinc_gets_it(tmp_path):
    """boinc_get_opencl_ids (main.cpp:1288-1308): without init_data.xml the program is
    standalone and takes device 0; init_data.xml's <gpu_device_num> or a --device argument
    names another. Devices past the last one are refused."""
    code, out, _ = run(tmp_path, 127, 1000, "--device", 7)
    assert code != 0 and "device 7 of" in out and "init_data.xml not found" not in out
    (tmp_path / "init_data.xml").write_text(
        "<app_init_data>\n<gpu_device_num>5</gpu_device_num>\n</app_init_data>\n")
    code, out, _ = run(tmp_path, 127, 1000)
    assert code != 0 and "device 5 of" in out and "init_data.xml not found" not in out
    (tmp_path / "init_data.xml").unlink()
    (tmp_path / "results-WW.txt").unlink()
    if no_gpu():
        code, out, _ = run(tmp_path, 127, 1000)
        assert "init_data.xml not found, using device 0." in out
        assert code != 0 and "device 0 of 0" in out


# ----------------------------------------------------------------------------- GPU


def expected_lines(lo, hi):
    lines = []
    for p, code, rem, quot in oracle.range_results(lo, hi)[2]:
        if not oracle.is_prime(p):
            continue
        if code == 1093:
            lines.append(f"{p} is a Wieferich prime" if (rem, quot) == (1, 0) else
                         f"{p} is a Wieferich special instance ({rem:+d} {quot:+d} p)")
        elif code == 1:
            lines.append(f"{p} is a WallSunSun prime")
        else:
            lines.append(f"{p} is a WallSunSun special instance (+0 {quot:+d} p)")
    return lines


@pytest.mark.gpu
def test_host_small_range_matches_oracle(tmp_path):
    # a BOINC client's init_data.xml names the device (the box's only one)
    (tmp_path / "init_data.xml").write_text("<app_init_data><gpu_device_num>0</gpu_device_num>"
                                            "</app_init_data>\n")
    rc, out, err = run(tmp_path, 1, "3e6")
    assert rc == 0, out + err
    assert "init_data.xml not found" not in out
    assert "GPU Wieferich/WallSunSun kernel tests passed" in err  # stdout only with -debug
    assert "Kernel profile:" in err
    text = results(tmp_path).splitlines()
    prp, cks, _ = oracle.range_results(127, 3_000_000)
    assert text[:-1] == expected_lines(127, 3_000_000)
    assert text[-1] == f"{cks:016X}"
    assert "1093 is a Wieferich prime" in text and "3511 is a Wieferich prime" in text


@pytest.mark.gpu
@pytest.mark.parametrize("start,end,want", [
    ("18446744013709551615", "18446744073709551615", "509DE6503B2A6E8E\n"),
    ("227630e10", "227636e10",
     "2276306935816523 is a Wieferich special instance (-1 -3 p)\n6522CBC4CA2E4CFF\n"),
    ("338772621946054253", "338772681946054253",
     "338772621946054253 is a WallSunSun special instance (+0 +91 p)\n589A69274A6AB5A8\n"),
], ids=["T1", "T3", "T4"])
def test_host_known_answers(tmp_path, start, end, want):
    rc, out, err = run(tmp_path, start, end)
    assert rc == 0, out + err
    assert results(tmp_path) == want
    assert "Workunit complete." in out


@pytest.mark.gpu
def test_host_known_answer_t2_last_line(tmp_path):
    rc, out, err = run(tmp_path, 1, "6e10")
    assert rc == 0, out + err
    assert results(tmp_path).splitlines()[-1] == "8B77D2DA053842DD"


@pytest.mark.gpu
def test_host_resume_gives_the_same_answer(tmp_path):
    """Kill-and-resume equivalent: a checkpoint in the middle of T3 carries the partial
    checksum; the resumed run must end with the same results file tail."""
    lo, hi = 2276300000000000, 2276360000000000
    mid = lo + 3 * (1 << 32)
    with wwlib.Context(0, list_capacity=wwlib.list_bound(1 << 32, 1 << 33)) as ctx:
        part = ctx.run_range(lo, mid, 1 << 32)
    write_state(tmp_path, "stateA-WW.txt", lo, hi, mid, part.checksum, 1.0, 1)
    (tmp_path / "results-WW.txt").write_text("")
    rc, out, err = run(tmp_path, "227630e10", "227636e10")
    assert rc == 0, out + err
    assert "Resuming search from checkpoint. Current: %d" % mid in out
    assert results(tmp_path).splitlines()[-1] == "6522CBC4CA2E4CFF"
    st = (tmp_path / "stateB-WW.txt").read_text().split()
    assert st[:3] == [str(lo), str(hi), str(hi)]


@pytest.mark.gpu
def test_host_testprimes_checks_every_chunk(tmp_path):
    """-testprimes (main.cpp:1731-1735): the two-kernel form, and every chunk's PRP list
    checked against the CPU's primes of the chunk before the tests run."""
    lo, hi = 10**9, 10**9 + 2_000_000
    rc, out, err = run(tmp_path, lo, hi, "-testprimes")
    assert rc == 0, out + err
    assert "Comparing all GPU generated primes with CPU generated primes." in out
    prp, cks, _ = oracle.range_results(lo, hi)
    assert f"GPU primes match CPU ({prp} PRPs)" in out
    assert results(tmp_path).splitlines() == expected_lines(lo, hi) + [f"{cks:016X}"]


@pytest.mark.gpu
@pytest.mark.parametrize("flags,target", [(["-debug", "-ws", "5"], 5.0), (["-compute"], 100.0),
                                          (["-compute", "-ws", "5"], 100.0)],
                         ids=["debug-ws5", "compute", "compute-ignores-ws"])
def test_host_kernel_time_targets(tmp_path, flags, target):
    """profileGPU (main.cpp:959-1107, 1424-1444): the chunk is sized from a timed probe to
    10 ms, -ws N ms, or 100 ms with -compute (which ignores -ws); -debug prints the kernel
    time split (main.cpp:1632-1661). The answer does not depend on any of it."""
    lo, hi = 10**12, 10**12 + 10_000_000
    rc, out, err = run(tmp_path, lo, hi, *flags)
    assert rc == 0, out + err
    line = next(l for l in err.splitlines() if l.startswith("Kernel profile:"))
    probe_ms = float(line.split()[2])
    est, actual = (int(x) for x in line.split(":")[-1].split("/"))
    assert probe_ms > 0 and actual >= 1 << 20
    expect = est * target / probe_ms  # the printed probe time is rounded to 1 us
    if expect >= 1 << 34:
        assert actual == 1 << 34
    else:
        assert abs(actual - max(1 << 20, expect)) <= 0.01 * expect + 1, (actual, expect)
    if "-ws" in flags and "-compute" not in flags:
        assert "Kernel work size of 5 specified on command line." in err
    if "-compute" in flags:
        assert "Compute mode enabled." in out and "Kernel work size" not in err
    if "-debug" in flags:
        assert "Percent of time spent in each kernel:" in out and "Prime Generator:" in out
        assert "GPU persistent kernel test passed." in out
    _, cks, _ = oracle.range_results(lo, hi)
    assert results(tmp_path).splitlines() == expected_lines(lo, hi) + [f"{cks:016X}"]


@pytest.mark.gpu
def test_host_tune_writes_and_later_runs_load_the_cache(tmp_path):
    """-tune runs the library's autotuner and saves the per-device cache; the next start
    loads it. The answer is the same either way."""
    cache = tmp_path / "tuning.json"
    lo, hi = 10**12, 10**12 + 4_000_000
    _, cks, _ = oracle.range_results(lo, hi)
    want = expected_lines(lo, hi) + [f"{cks:016X}"]
    rc, out, err = run(tmp_path, lo, hi, "-tune", env={"WW_TUNE_CACHE": str(cache)})
    assert rc == 0, out + err
    assert "Tuned (" in out and f"Tuning saved to {cache}" in err
    assert cache.exists() and '"segment_bits"' in cache.read_text()
    assert results(tmp_path).splitlines() == want
    for f in ("stateA-WW.txt", "stateB-WW.txt"):
        if (tmp_path / f).exists():
            (tmp_path / f).unlink()
    rc, out, err = run(tmp_path, lo, hi, env={"WW_TUNE_CACHE": str(cache)})
    assert rc == 0, out + err
    assert f"Tuning loaded from {cache}" in err
    assert results(tmp_path).splitlines() == want
