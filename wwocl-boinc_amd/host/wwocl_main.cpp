/*
 * wwocl_main.cpp -- the drop-in host: the reference's command line, checkpoint files and
 * results file, driving the gfx950 path through the C ABI (include/ww_api.h).
 *
 *   wwocl START END [-testprimes] [-compute] [-debug] [-ws N]
 *
 * Behaviour follows main.cpp of the reference:
 *   START/END parsing with "Me X" notation          main.cpp:1193-1274
 *   A/B state files and their selection             main.cpp:103-221, 1518-1551
 *   startup self-test against the CPU path          main.cpp:590-738 (testKernels)
 *   PRP list check (-testprimes)                    main.cpp:451-523 (checkPRPS)
 *   chunk loop, checkpoint every > 60 s             main.cpp:1683-1795
 *   result validation, sorting, line formats        main.cpp:768-956
 *   final checksum line                             main.cpp:1758-1782
 * BOINC's library is absent: the program keeps the standalone behaviour (messages go to
 * both stdout and stderr where the reference prints both in standalone mode, progress as
 * "Tests done" lines) and takes the device the way boinc_get_opencl_ids does
 * (main.cpp:1288-1308): --device N on the command line, else <gpu_device_num> or
 * <gpu_opencl_dev_index> from init_data.xml, else device 0 with the reference's
 * "init_data.xml not found" message. The device work-size probe (main.cpp:959-1107) is
 * replaced by a fixed chunk size: the checksum does not depend on chunking.
 */
#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <unistd.h>
#include <vector>

#include "ww_api.h"

namespace {

const char *kResults = "results-WW.txt"; /* main.cpp:40 */
const char *kStateA = "stateA-WW.txt";   /* main.cpp:41 */
const char *kStateB = "stateB-WW.txt";   /* main.cpp:42 */
const uint32_t kNumResults = WW_DEFAULT_RESULTS;

/* both streams, as the reference does in standalone mode */
void say(const char *fmt, ...)
{
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    vfprintf(stdout, fmt, a);
    vfprintf(stderr, fmt, b);
    va_end(b);
    va_end(a);
    fflush(stdout);
}

[[noreturn]] void die(const char *fmt, ...)
{
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    vfprintf(stdout, fmt, a);
    vfprintf(stderr, fmt, b);
    va_end(b);
    va_end(a);
    fflush(stdout);
    exit(EXIT_FAILURE);
}

void check(int rc, const char *what)
{
    if (rc != WW_OK) die("ERROR: %s: %s\n", what, ww_last_error());
}

/* START or END: "M" parsed as unsigned decimal, or "MeX" with M parsed by strtoull base 0
   and X by strtoul, then M * 10^X with overflow check (main.cpp:1193-1261). */
uint64_t parse_bound(const char *arg, const char *name)
{
    const char *e = strstr(arg, "e");
    if (!e) {
        uint64_t v = 0;
        if (sscanf(arg, "%" SCNu64, &v) != 1) v = 0;
        return v;
    }
    char *end = nullptr;
    uint64_t num = strtoull(arg, &end, 0);
    int32_t exp = (int32_t)strtoul(end + 1, nullptr, 0);
    if (num == 0 || exp == 0) die("ERROR: unable to parse %s\n", name);
    for (int i = 0; i < exp; ++i) {
        if (num > UINT64_MAX / 10) die("ERROR: %s is out of range\n", name);
        num *= 10;
    }
    return num;
}

struct State {
    uint64_t current = 0, cksm = 0;
    double psum = 0;
    int pcnt = 0;
};

/* One state file; valid when it parses and its start/stop match (main.cpp:146-176). */
bool read_one(const char *path, uint64_t start, uint64_t stop, State &s)
{
    FILE *in = fopen(path, "r");
    if (!in) return false;
    uint64_t a, b;
    int n = fscanf(in, "%" SCNu64 " %" SCNu64 " %" SCNu64 " %" SCNu64 " %lf %d\n", &a, &b,
                   &s.current, &s.cksm, &s.psum, &s.pcnt);
    fclose(in);
    if (n != 6) {
        fprintf(stderr, "Cannot parse %s !!!\n", path);
        return false;
    }
    return a == start && b == stop;
}

/* Newest valid of A and B; *write_a_next names the file the next checkpoint overwrites
   (main.cpp:178-208). */
bool read_state(uint64_t start, uint64_t stop, State &st, bool &write_a_next)
{
    State a, b;
    bool ga = read_one(kStateA, start, stop, a);
    bool gb = read_one(kStateB, start, stop, b);
    if (ga && gb) {
        if (a.current > b.current) gb = false;
        else ga = false;
    }
    if (ga) {
        st = a;
        write_a_next = false;
        return true;
    }
    if (gb) {
        st = b;
        write_a_next = true;
        return true;
    }
    return false;
}

/* The file toggles only after a successful close (main.cpp:103-130). */
void write_state(uint64_t start, uint64_t stop, const State &s, bool &write_a_next)
{
    const char *path = write_a_next ? kStateA : kStateB;
    FILE *out = fopen(path, "w");
    if (!out) {
        fprintf(stderr, "Cannot open %s !!!\n", path);
        return;
    }
    if (fprintf(out, "%" PRIu64 " %" PRIu64 " %" PRIu64 " %" PRIu64 " %lf %d\n", start, stop,
                s.current, s.cksm, s.psum, s.pcnt) < 0) {
        fprintf(stderr, "Cannot write to %s !!! Continuing...\n", path);
        fclose(out);
        return;
    }
    if (fclose(out) == 0) write_a_next = !write_a_next;
}

/* Append and fsync (report_solution, main.cpp:741-765). */
void report_solution(const std::string &text)
{
    FILE *f = fopen(kResults, "a");
    if (!f) die("Cannot open %s !!!\n", kResults);
    if (fprintf(f, "%s", text.c_str()) < 0) die("Cannot write to %s !!!\n", kResults);
    fflush(f);
    fsync(fileno(f));
    fclose(f);
}

/* The results-file line of one record, validated on the CPU: primes are formatted
   (main.cpp:869-905), base-2 PRPs that are not prime are dropped (main.cpp:907-911),
   anything else is a computation error (main.cpp:912-916). Returns false when dropped. */
bool format_record(const ww_result &r, std::string &line, uint32_t &dropped)
{
    char buf[256];
    if (ww_cpu_is_prime(r.p)) {
        if (r.result == 1093) {
            if (r.rem == 1 && r.quot == 0)
                snprintf(buf, sizeof buf, "%" PRIu64 " is a Wieferich prime\n", r.p);
            else
                snprintf(buf, sizeof buf, "%" PRIu64 " is a Wieferich special instance (%+d %+d p)\n",
                         r.p, r.rem, r.quot);
        } else if (r.result == 1) {
            snprintf(buf, sizeof buf, "%" PRIu64 " is a WallSunSun prime\n", r.p);
        } else if (r.result == 2) {
            snprintf(buf, sizeof buf, "%" PRIu64 " is a WallSunSun special instance (+0 %+d p)\n",
                     r.p, r.quot);
        } else {
            die("Error: hostResult returned %d.\n", r.result);
        }
        line = buf;
        return true;
    }
    if (ww_cpu_is_prp(r.p)) {
        ++dropped;
        return false;
    }
    fprintf(stderr, "Computation error, GPU found a composite result: %" PRIu64 "\n", r.p);
    exit(EXIT_FAILURE);
}

struct Run {
    ww_ctx *ctx = nullptr;
    uint64_t list_cap = 0;
    int debug = 0;
    bool fused = true; /* chunks through the persistent kernel (no PRP list) */
};

/* checkPRPS (main.cpp:451-523): every list entry is prime or a base-2 PRP with no factor
   <= 113, and the primes are exactly those of [lo, hi). The reference compares against
   primesieve; here the CPU enumerates the candidates of the range with the host checks. */
uint32_t check_prps(Run &run, uint64_t lo, uint64_t hi)
{
    uint64_t n = 0;
    check(ww_read_prps(run.ctx, nullptr, 0, &n), "ww_read_prps");
    std::vector<uint64_t> gpu(n);
    check(ww_read_prps(run.ctx, gpu.data(), n, &n), "ww_read_prps");
    std::vector<uint64_t> primes;
    for (uint64_t p : gpu) {
        if (ww_cpu_is_prime(p)) primes.push_back(p);
        else if (!ww_cpu_is_prp(p)) die("Computation error, GPU found a composite: %" PRIu64 "\n", p);
    }
    std::sort(primes.begin(), primes.end());
    std::vector<uint64_t> cpu;
    for (uint64_t v = lo | 1; v < hi && v >= lo; v += 2) {
        if (ww_cpu_is_prp(v) && ww_cpu_is_prime(v)) cpu.push_back(v);
        if (v + 2 < v) break;
    }
    if (primes.size() != cpu.size())
        die("ERROR: gpu prime count: %zu does not match cpu: %zu\n", primes.size(), cpu.size());
    for (size_t i = 0; i < cpu.size(); ++i)
        if (primes[i] != cpu[i])
            die("ERROR: gpu/cpu prime mismatch %" PRIu64 " vs %" PRIu64 "\n", primes[i], cpu[i]);
    return (uint32_t)n;
}

/* testKernels (main.cpp:590-738): the device path and the CPU path over
   [start, start + min(2e6, chunk)) must agree on PRP count, checksum and result count. */
void test_kernels(Run &run, uint64_t start, uint64_t top, uint64_t chunk)
{
    uint64_t stop = start + std::min<uint64_t>(2000000, chunk);
    if (stop < start || stop > top) stop = top;
    fprintf(stderr, "Verifying GPU kernels using range %" PRIu64 " to %" PRIu64 "...\n", start, stop);
    if (run.debug) printf("Verifying GPU kernels using range %" PRIu64 " to %" PRIu64 "...\n", start, stop);
    check(ww_clear_results(run.ctx), "clear results");
    check(ww_clear_prp_count(run.ctx), "clear PRP count");
    check(ww_getsegprimes(run.ctx, start, stop), "getsegprimes");
    check(ww_wieferich_wallsunsun(run.ctx), "wieferich/wallsunsun");
    uint64_t cpu[3];
    check(ww_cpu_range(start, stop, cpu), "cpu range");
    uint32_t gpu_prps = check_prps(run, start, stop);
    if (gpu_prps != cpu[0])
        die("ERROR: gpu / cpu prp count mismatch %u / %u\n", gpu_prps, (uint32_t)cpu[0]);
    say("GPU prime generator kernel test passed. (%u PRPs)\n", gpu_prps);
    ww_counters k;
    check(ww_read_counters(run.ctx, &k), "read counters");
    if (k.checksum != cpu[1])
        die("ERROR: gpu / cpu checksum mismatch %" PRIu64 " / %" PRIu64 "\n", k.checksum, cpu[1]);
    if (k.resultcount != cpu[2])
        die("ERROR: gpu / cpu result count mismatch %u / %u\n", k.resultcount, (uint32_t)cpu[2]);
    say("GPU Wieferich/WallSunSun kernel tests passed. (%" PRIu64 " checksum, %u results)\n",
        k.checksum, k.resultcount);
    if (run.fused) {
        /* the chunk body the search uses must agree with the CPU path too */
        check(ww_clear_results(run.ctx), "clear results");
        check(ww_search_fused(run.ctx, start, stop), "persistent kernel");
        ww_counters f;
        check(ww_read_counters(run.ctx, &f), "read counters");
        if (f.totalprimecount != cpu[0] || f.checksum != cpu[1] || f.resultcount != cpu[2])
            die("ERROR: persistent kernel / cpu mismatch: %" PRIu64 " / %" PRIu64 " PRPs, checksum %"
                PRIu64 " / %" PRIu64 ", %u / %u results\n",
                f.totalprimecount, cpu[0], f.checksum, cpu[1], f.resultcount, (uint32_t)cpu[2]);
        if (run.debug) printf("GPU persistent kernel test passed.\n");
    }
}

/* getResults (main.cpp:768-956). */
void get_results(Run &run, int ckpt_time, State &st)
{
    ww_counters k;
    check(ww_read_counters(run.ctx, &k), "read counters");
    st.cksm += k.checksum;
    if (k.primecount_max > run.list_cap) die("Error: prime array overflow!\n");
    uint32_t dropped = 0;
    if (k.resultcount > 0) {
        if (k.resultcount > kNumResults)
            die("Error: number of results (%u) overflowed array.\n", k.resultcount);
        std::vector<ww_result> recs(k.resultcount);
        uint32_t n = 0;
        check(ww_read_results(run.ctx, recs.data(), k.resultcount, &n), "read results");
        recs.resize(n);
        /* by p, Wieferich records before Wall-Sun-Sun records of the same p: the order the
           reference's stable bubble sort yields (main.cpp:843-854) */
        std::stable_sort(recs.begin(), recs.end(), [](const ww_result &a, const ww_result &b) {
            if (a.p != b.p) return a.p < b.p;
            return (a.result == 1093) && (b.result != 1093);
        });
        std::string text;
        for (const ww_result &r : recs) {
            std::string line;
            if (format_record(r, line, dropped)) text += line;
        }
        report_solution(text);
    }
    if (run.debug) {
        printf("PRPs tested since last checkpoint: %" PRIu64 "\n", k.totalprimecount);
        printf("Max number of results: %u\n", kNumResults);
        printf("Number of results: %u\n", k.resultcount);
        printf("PRPs removed from results: %u\n", dropped);
        printf("Largest kernel PRP count: %u\n", k.primecount_max);
        printf("Max kernel PRP count for malloc: %" PRIu64 "\n", run.list_cap);
    }
    if (ckpt_time == 0) ckpt_time = 1;
    double p_sec = ((double)(k.totalprimecount / (uint64_t)ckpt_time)) / 1000000.0;
    st.psum += p_sec;
    ++st.pcnt;
}

/* The device a BOINC client assigned, as boinc_get_opencl_ids finds it (main.cpp:1288-1308):
   "--device N" (older clients pass it on the command line), else the device number in
   init_data.xml, else standalone mode: device 0. */
int pick_device(int argc, char *argv[])
{
    for (int i = 3; i + 1 < argc; ++i)
        if (strcmp(argv[i], "--device") == 0) return atoi(argv[i + 1]);
    FILE *f = fopen("init_data.xml", "r");
    if (!f) {
        printf("init_data.xml not found, using device 0.\n");
        return 0;
    }
    std::string xml;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) xml.append(buf, n);
    fclose(f);
    for (const char *tag : {"<gpu_device_num>", "<gpu_opencl_dev_index>"}) {
        size_t at = xml.find(tag);
        if (at != std::string::npos) return atoi(xml.c_str() + at + strlen(tag));
    }
    return 0;
}

void usage(const char *prog)
{
    printf("Usage: %s START END -testprimes -compute -debug -ws #\nWhere 127 <= START < END < 2^64\n", prog);
    printf("-testprimes is optional and will compare all GPU generated primes with CPU generated primes (very slow)\n\n");
    printf("-compute is optional and uses the largest chunk size.\n\n");
    printf("-debug is optional and will display more information.\n\n");
    printf("-ws # is optional chunk size scale, 1..100 (default 100 = 2^32 integers per chunk).\n\n");
    printf("Example format: %s 10000000 20000000\n", prog);
    printf("            Or: %s 1e7 2e7\n", prog);
    exit(EXIT_FAILURE);
}

} // namespace

int main(int argc, char *argv[])
{
    fprintf(stderr, "\nWW-MI355X: Wieferich and WallSunSun prime search on gfx950 (drop-in for WWocl 1.0.6)\n");
    printf("WW-MI355X: Wieferich and WallSunSun prime search on gfx950 (drop-in for WWocl 1.0.6)\n");
    fprintf(stderr, "Command line: ");
    for (int i = 0; i < argc; i++) fprintf(stderr, "%s ", argv[i]);
    fprintf(stderr, "\n");
    if (argc < 3) usage(argv[0]);

    int test_prps = 0, debug = 0, compute = 0, ws = 0;
    for (int w = 3; w < argc; ++w) {
        if (strcmp(argv[w], "-compute") == 0) {
            compute = 1;
            say("Compute mode enabled.\n");
        } else if (strcmp(argv[w], "-testprimes") == 0) {
            test_prps = 1;
            say("Comparing all GPU generated primes with CPU generated primes.\n");
        } else if (strcmp(argv[w], "-debug") == 0) {
            debug = 1;
            printf("Debug info enabled.\n");
        } else if (strcmp(argv[w], "-ws") == 0 && w + 1 < argc) {
            int v = 0;
            if (sscanf(argv[w + 1], "%d", &v) == 1 && v > 0 && v <= 100) ws = v;
        }
    }

    uint64_t bottom = parse_bound(argv[1], "START");
    uint64_t top = parse_bound(argv[2], "END");
    if (bottom < WW_MIN_START) {
        say("START below minimum. Starting search at 127\n");
        bottom = WW_MIN_START;
    }
    if (top <= bottom) die("ERROR: END must be > START\n");
    say("Starting search at: %" PRIu64 "\nStopping search at: %" PRIu64 "\n", bottom, top);

    State st;
    bool write_a_next = true;
    if (read_state(bottom, top, st, write_a_next)) {
        say("Resuming search from checkpoint. Current: %" PRIu64 "\n", st.current);
    } else {
        printf("Beginning a new search with parameters from the command line\n");
        st = State();
        st.current = bottom;
        write_a_next = true;
        FILE *f = fopen(kResults, "w"); /* a fresh start truncates the results file */
        if (!f) die("Cannot open %s !!!\n", kResults);
        fclose(f);
    }
    if (st.current == top) {
        say("Workunit complete.\n");
        return 0;
    }

    /* chunk size: 2^32 integers (a 3.2 GB list at most), scaled down by -ws */
    uint64_t chunk = 1ull << 32;
    if (!compute && ws) chunk = std::max<uint64_t>(1ull << 20, chunk / 100 * (uint64_t)ws);

    const int device = pick_device(argc, argv);
    int ndev = 0;
    if (ww_device_count(&ndev) != WW_OK || device < 0 || device >= ndev)
        die("Error: boinc_get_opencl_ids() failed: device %d of %d\n", device, ndev);
    ww_device_info_t info;
    check(ww_device_info(device, &info), "device info");
    say("GPU Info:\n  Name: \t\t%s\n  Arch: \t\t%s\n  Compute Units: \t%d\n", info.name, info.arch,
        info.compute_units);

    Run run;
    run.debug = debug;
    run.fused = !test_prps;
    run.list_cap = ww_list_bound(1ull << 32, (1ull << 32) + chunk); /* worst case: low p */
    if (run.list_cap < ww_list_bound(WW_MIN_START, WW_MIN_START + chunk))
        run.list_cap = ww_list_bound(WW_MIN_START, WW_MIN_START + chunk);
    check(ww_open(device, run.list_cap, kNumResults, &run.ctx), "ww_open");
    if (debug) printf("chunk size: %" PRIu64 " integers, list capacity %" PRIu64 "\n", chunk, run.list_cap);

    test_kernels(run, st.current, top, chunk);
    say("Starting search...\n");

    time_t last_progress = time(nullptr), ckpt_last = time(nullptr);
    check(ww_clear_results(run.ctx), "clear results");
    int in_flight = 0;
    for (uint64_t start = st.current, stop; start < top; start = stop) {
        stop = start + chunk;
        if (stop > top || stop < start) stop = top;
        time_t now = time(nullptr);
        if (now - last_progress > 1) {
            printf("Tests done: %.4f%%\n", (double)(start - bottom) / (double)(top - bottom) * 100.0);
            fflush(stdout);
            last_progress = now;
        }
        if (run.fused) {
            check(ww_search_fused(run.ctx, start, stop), "persistent kernel");
        } else { /* -testprimes needs the PRP list of the two-kernel form */
            check(ww_clear_prp_count(run.ctx), "clear PRP count");
            check(ww_getsegprimes(run.ctx, start, stop), "getsegprimes");
            if (test_prps) printf("GPU primes match CPU (%u PRPs)\n", check_prps(run, start, stop));
            check(ww_wieferich_wallsunsun(run.ctx), "wieferich/wallsunsun");
        }
        if (++in_flight == 3) { /* bound the queue depth as main.cpp:1685-1691 */
            check(ww_finish(run.ctx), "finish");
            in_flight = 0;
        }
        now = time(nullptr);
        int ckpt_diff = (int)(now - ckpt_last);
        if (ckpt_diff > 60 || stop == top) {
            check(ww_finish(run.ctx), "finish");
            in_flight = 0;
            get_results(run, ckpt_diff, st);
            st.current = stop;
            if (stop == top) {
                char buf[64];
                snprintf(buf, sizeof buf, "%016" PRIX64 "\n", st.cksm);
                report_solution(buf);
                float avg = (float)(st.psum / st.pcnt);
                say("Workunit complete. Average speed: %.2f Million primes per second.\n", avg);
                printf("Checksum: %016" PRIX64 "\n", st.cksm);
            }
            write_state(bottom, top, st, write_a_next);
            printf("Checkpoint, current: %" PRIu64 "\n", st.current);
            ckpt_last = now;
            check(ww_clear_results(run.ctx), "clear results");
        }
    }
    ww_close(run.ctx);
    return 0;
}
