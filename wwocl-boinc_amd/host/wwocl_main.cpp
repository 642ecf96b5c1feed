/*
 * wwocl_main.cpp -- the drop-in host: the reference's command line, checkpoint files and
 * results file, driving the gfx950 path through the C ABI (include/ww_api.h).
 *
 *   wwocl START END [-testprimes] [-compute] [-debug] [-ws N] [-tune]
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
 * "init_data.xml not found" message. The work-size probe (profileGPU, main.cpp:959-1107)
 * times the chunk body at the current position and scales the chunk to the reference's
 * kernel-time targets: 10 ms, -ws N ms (1..100), or 100 ms with -compute. The checksum
 * does not depend on chunking. The library's autotuner (sieve bounds, geometry, launch
 * shape) runs with -tune and writes a per-device tuning cache, which later starts load
 * ($WW_TUNE_CACHE, else ~/.cache/wwocl/tuning-<target>-<CUs>cu.json). Validation lives in
 * wwocl_checks.cpp.
 */
#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "ww_api.h"
#include "wwocl_checks.h"

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

/* Print on both streams and exit with EXIT_FAILURE (the reference's error pattern). */
#define die(...) wwhost::fatal(true, __VA_ARGS__)

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

/* The library's context as the checks' view of the device. */
class LibView : public wwhost::DeviceView {
public:
    explicit LibView(ww_ctx *c) : c_(c) {}
    int read_counters(ww_counters *out) override { return ww_read_counters(c_, out); }
    int read_results(ww_result *out, uint32_t cap, uint32_t *n) override
    {
        return ww_read_results(c_, out, cap, n);
    }
    int read_prps(uint64_t *out, uint64_t cap, uint64_t *n) override
    {
        return ww_read_prps(c_, out, cap, n);
    }

private:
    ww_ctx *c_;
};

struct Run {
    ww_ctx *ctx = nullptr;
    uint64_t list_cap = 0;
    int debug = 0;
    bool fused = true; /* chunks through the persistent kernel (no PRP list) */
};

/* One chunk [start, stop) of the search through the chunk body in use. */
void run_chunk(Run &run, uint64_t start, uint64_t stop)
{
    if (run.fused) {
        check(ww_search_fused(run.ctx, start, stop), "persistent kernel");
    } else {
        check(ww_clear_prp_count(run.ctx), "clear PRP count");
        check(ww_getsegprimes(run.ctx, start, stop), "getsegprimes");
        check(ww_wieferich_wallsunsun(run.ctx), "wieferich/wallsunsun");
    }
}

/* Largest chunk the chunk body takes: 2^34 integers for the persistent kernel; for the
   two-kernel form the reference's limit (main.cpp:1077-1079), whose list the context is
   sized for. */
uint64_t max_chunk(const Run &run) { return run.fused ? (1ull << 34) : 4294900000ull; }

/* profileGPU (main.cpp:959-1107): the first estimate is the reference's (compute units x
   2.2e6, or x 2.2e7 with -compute), timed at the current position (at least 1e8) once to
   warm up and three times more; the chunk is that estimate scaled to the kernel-time
   target. Leaves the counters cleared. */
uint64_t profile_chunk(Run &run, int cus, uint64_t current, double target_ms, int compute)
{
    uint64_t est = (uint64_t)cus * (compute ? 22000000ull : 2200000ull);
    est = std::min(est, max_chunk(run));
    uint64_t start = std::max<uint64_t>(current, 100000000ull);
    if (start > UINT64_MAX - est) start = UINT64_MAX - est;
    double ms = 0;
    for (int rep = 0; rep < 4; ++rep) {
        check(ww_clear_results(run.ctx), "clear results");
        check(ww_kernel_times(run.ctx, nullptr, nullptr, nullptr), "kernel times");
        run_chunk(run, start, start + est);
        double a = 0, b = 0;
        check(ww_kernel_times(run.ctx, &a, &b, nullptr), "kernel times");
        if (rep) ms += a + b; /* the first is ignored, as the reference does */
    }
    check(ww_clear_results(run.ctx), "clear results");
    check(ww_clear_prp_count(run.ctx), "clear PRP count");
    const double avg = ms / 3.0;
    double chunk = avg > 0 ? (double)est * (target_ms / avg) : (double)est;
    uint64_t c = chunk >= (double)max_chunk(run) ? max_chunk(run) : (uint64_t)chunk;
    if (c < (1ull << 20)) c = 1ull << 20;
    fprintf(stderr, "Kernel profile: %0.3f ms. Estimated / Actual worksize: %" PRIu64 " / %" PRIu64 "\n",
            avg, est, c);
    if (run.debug)
        printf("Kernel profile: %0.3f ms. Estimated / Actual worksize: %" PRIu64 " / %" PRIu64 "\n",
               avg, est, c);
    return c;
}

/* -debug (main.cpp:1632-1661): time one chunk at the current position in each kernel. The
   two-kernel form splits the sieve + PRP kernel (with its pre-sieve) from the fused
   Wieferich + Wall-Sun-Sun kernel; the persistent kernel, which the search uses unless
   -testprimes, does all of it in one launch. Leaves the counters cleared. */
void profile_kernels(Run &run, uint64_t current, uint64_t top, uint64_t chunk)
{
    printf("Profiling kernels...\n");
    uint64_t stop = current + std::min<uint64_t>(chunk, 4294900000ull);
    if (stop > top || stop < current) stop = top;
    double prime_ms = 0, ww_ms = 0, fused_ms = 0;
    check(ww_kernel_times(run.ctx, nullptr, nullptr, nullptr), "kernel times");
    check(ww_clear_results(run.ctx), "clear results");
    check(ww_clear_prp_count(run.ctx), "clear PRP count");
    check(ww_getsegprimes(run.ctx, current, stop), "getsegprimes");
    check(ww_wieferich_wallsunsun(run.ctx), "wieferich/wallsunsun");
    check(ww_kernel_times(run.ctx, &prime_ms, &ww_ms, nullptr), "kernel times");
    check(ww_clear_results(run.ctx), "clear results");
    check(ww_search_fused(run.ctx, current, stop), "persistent kernel");
    check(ww_kernel_times(run.ctx, nullptr, &fused_ms, nullptr), "kernel times");
    check(ww_clear_results(run.ctx), "clear results");
    check(ww_clear_prp_count(run.ctx), "clear PRP count");
    const double total = prime_ms + ww_ms;
    const int prime_pct = total > 0 ? (int)(prime_ms / total * 100.0) : 0;
    const int ww_pct = total > 0 ? (int)(ww_ms / total * 100.0) : 0;
    printf("Percent of time spent in each kernel:\n");
    printf("\tPrime Generator:\t%d %%\t(%0.1f ms)\n\tWieferich+WallSunSun Test:\t%d %%\t(%0.1f ms)\n",
           prime_pct, prime_ms, ww_pct, ww_ms);
    printf("\tPersistent kernel (sieve, PRP and both tests in one launch):\t(%0.1f ms)\n", fused_ms);
}

/* testKernels (main.cpp:590-738): the device path and the CPU path over
   [start, start + min(2e6, chunk)) must agree on PRP count, checksum and result count.
   The CPU side (ww_cpu_range) is a formulation of its own, not the device arithmetic. */
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
    LibView view(run.ctx);
    const uint32_t gpu_prps = wwhost::check_prps(view, start, stop);
    ww_counters k;
    check(ww_read_counters(run.ctx, &k), "read counters");
    wwhost::compare_selftest(gpu_prps, k, cpu);
    fprintf(stderr, "GPU prime generator kernel test passed. (%u PRPs)\n", gpu_prps);
    if (run.debug) printf("GPU prime generator kernel test passed. (%u PRPs)\n", gpu_prps);
    fprintf(stderr, "GPU Wieferich/WallSunSun kernel tests passed. (%" PRIu64 " checksum, %u results)\n",
            k.checksum, k.resultcount);
    if (run.debug)
        printf("GPU Wieferich/WallSunSun kernel tests passed. (%" PRIu64 " checksum, %u results)\n",
               k.checksum, k.resultcount);
    if (run.fused) {
        /* the chunk body the search uses must agree with the CPU path too */
        check(ww_clear_results(run.ctx), "clear results");
        check(ww_search_fused(run.ctx, start, stop), "persistent kernel");
        ww_counters f;
        check(ww_read_counters(run.ctx, &f), "read counters");
        wwhost::compare_selftest_fused(f, cpu);
        if (run.debug) printf("GPU persistent kernel test passed.\n");
    }
}

/* getResults (main.cpp:768-956): validation in wwocl_checks.cpp, then the results file,
   the -debug report and the speed bookkeeping. */
void get_results(Run &run, int ckpt_time, State &st)
{
    LibView view(run.ctx);
    const wwhost::CheckpointReport rep = wwhost::collect_results(view, run.list_cap, kNumResults);
    const ww_counters &k = rep.counters;
    st.cksm += k.checksum;
    if (k.resultcount > 0) report_solution(rep.text);
    if (run.debug) {
        printf("PRPs tested since last checkpoint: %" PRIu64 "\n", k.totalprimecount);
        printf("Max number of results:             %u\n", kNumResults);
        printf("Number of results:                 %u\n", k.resultcount);
        printf("PRPs removed from results:         %u\n", rep.dropped);
        printf("Largest kernel PRP count:          %u\n", k.primecount_max);
        printf("Max kernel PRP count for malloc:   %" PRIu64 "\n", run.list_cap);
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

/* The tuning cache of this device. */
std::string tuning_cache_path(const ww_device_info_t &info)
{
    const char *env = getenv("WW_TUNE_CACHE");
    if (env && *env) return env;
    const char *home = getenv("HOME");
    std::string dir = std::string(home && *home ? home : ".") + "/.cache";
    mkdir(dir.c_str(), 0755);
    dir += "/wwocl";
    mkdir(dir.c_str(), 0755);
    std::string target;
    for (const char *c = info.arch; *c && *c != ':'; ++c) target += *c;
    return dir + "/tuning-" + target + "-" + std::to_string(info.compute_units) + "cu.json";
}

/* -tune: the library's autotuner on 2^31 integers at the current position (at least 1e8),
   then the cache; otherwise the cache when there is one. */
void apply_tuning(Run &run, const ww_device_info_t &info, uint64_t current, bool tune)
{
    const std::string path = tuning_cache_path(info);
    if (tune) {
        uint64_t lo = std::max<uint64_t>(current, 100000000ull);
        if (lo > UINT64_MAX - (1ull << 31)) lo = UINT64_MAX - (1ull << 31);
        ww_tuning best;
        uint32_t n = 0;
        fprintf(stderr, "Tuning the GPU kernels...\n");
        check(ww_autotune(run.ctx, lo, lo + (1ull << 31), nullptr, &best, nullptr, 0, &n), "autotune");
        say("Tuned (%u trials): deep bound %u, pre-sieve bound %u, %u waves per workgroup, "
            "%d workgroups per CU, claims of %u, %u-bit sieve segments: %.3e primes/s\n",
            n, best.deep_bound, best.presieve_bound, best.waves_per_group, best.blocks_per_cu,
            best.claim, best.segment_bits, best.prps_per_s);
        if (ww_tuning_save(run.ctx, path.c_str()) == WW_OK)
            fprintf(stderr, "Tuning saved to %s\n", path.c_str());
        else
            fprintf(stderr, "Tuning not saved: %s\n", ww_last_error());
        return;
    }
    if (access(path.c_str(), R_OK) != 0) return;
    ww_tuning t;
    if (ww_tuning_load(run.ctx, path.c_str(), &t) == WW_OK)
        fprintf(stderr, "Tuning loaded from %s\n", path.c_str());
    else
        fprintf(stderr, "Tuning cache ignored: %s\n", ww_last_error());
}

void usage(const char *prog)
{
    printf("Usage: %s START END -testprimes -compute -debug -ws #\nWhere 127 <= START < END < 2^64\n", prog);
    printf("-testprimes is optional and will compare all GPU generated primes with CPU generated primes (very slow)\n\n");
    printf("-compute is optional and will fully load the gpu with large kernels (100 ms each).\n\n");
    printf("-debug is optional and will display more information in standalone mode.\n\n");
    printf("-ws # is optional worksize for adjusting the GPU load higher for faster computation or lower for better screen refresh rate.\n");
    printf(" If using compute mode, -ws is ignored.  Usage: -ws 10 for default (10 ms kernels), -ws 1 for lowest GPU load, -ws 100 for highest GPU load.\n\n");
    printf("-tune is optional and tunes the kernels for this GPU, saving the result for later runs.\n\n");
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

    int test_prps = 0, debug = 0, compute = 0, ws = 0, tune = 0;
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
        } else if (strcmp(argv[w], "-tune") == 0) {
            tune = 1;
        }
    }
    /* command-line kernel time in ms, ignored in compute mode (main.cpp:1424-1444) */
    if (!compute) {
        for (int w = 3; w + 1 < argc; ++w) {
            if (strcmp(argv[w], "-ws") != 0) continue;
            int v = 0;
            if (sscanf(argv[w + 1], "%d", &v) == 1 && v > 0 && v <= 100) {
                fprintf(stderr, "Kernel work size of %d specified on command line.\n", v);
                if (debug) printf("Kernel work size of %d specified on command line.\n", v);
                ws = v;
            }
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
    /* the PRP list: the two-kernel form (-testprimes, and the -debug split) needs one for a
       whole chunk; the persistent kernel needs none, so then it only holds the self-test's
       2e6-integer range (worst case: the smallest p, where primes are densest) */
    const uint64_t list_span = (test_prps || debug) ? 4294900000ull : 2000000ull;
    run.list_cap = std::max(ww_list_bound(1ull << 32, (1ull << 32) + list_span),
                            ww_list_bound(WW_MIN_START, WW_MIN_START + list_span));
    check(ww_open(device, run.list_cap, kNumResults, &run.ctx), "ww_open");
    apply_tuning(run, info, st.current, tune != 0);

    const double target_ms = compute ? 100.0 : (ws ? (double)ws : 10.0);
    const uint64_t chunk = profile_chunk(run, info.compute_units, st.current, target_ms, compute);
    if (debug) {
        printf("chunk size: %" PRIu64 " integers, list capacity %" PRIu64 "\n", chunk, run.list_cap);
        profile_kernels(run, st.current, top, chunk);
    }

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
            run_chunk(run, start, stop);
        } else { /* -testprimes checks the PRP list of the two-kernel form (main.cpp:1731-1735) */
            LibView view(run.ctx);
            check(ww_clear_prp_count(run.ctx), "clear PRP count");
            check(ww_getsegprimes(run.ctx, start, stop), "getsegprimes");
            printf("GPU primes match CPU (%u PRPs)\n", wwhost::check_prps(view, start, stop));
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
