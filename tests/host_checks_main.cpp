/*
 * host_checks_main.cpp -- CPU driver for the drop-in host's validation (wwocl_checks.cpp):
 * injects counters, result records and PRP lists through a fake device view and reports
 * what the checks do. Built and run by tests/test_host_checks.py (with AddressSanitizer
 * and UndefinedBehaviorSanitizer), no GPU involved.
 *
 *   host_checks SCENARIO [ARGS...]
 * A failure check prints "FATAL both=<0|1>" and the message, and exits with 3 (the test
 * hook throws where the program's hook exits); scenario default_hook keeps the program's
 * own hook, which exits with EXIT_FAILURE.
 */
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../wwocl-boinc_amd/host/wwocl_checks.h"

/* The library's error text is not linked in here. */
extern "C" const char *ww_last_error(void) { return "injected read failure"; }

namespace {

struct Fatal {
    bool both;
    std::string msg;
};

void throwing_hook(bool both, const std::string &msg) { throw Fatal{both, msg}; }

class FakeView : public wwhost::DeviceView {
public:
    ww_counters k{};
    std::vector<ww_result> recs;
    std::vector<uint64_t> list;
    bool fail_reads = false;

    int read_counters(ww_counters *out) override
    {
        if (fail_reads) return WW_E_HIP;
        *out = k;
        return WW_OK;
    }
    int read_results(ww_result *out, uint32_t cap, uint32_t *n) override
    {
        if (fail_reads) return WW_E_HIP;
        uint32_t m = (uint32_t)recs.size() < cap ? (uint32_t)recs.size() : cap;
        for (uint32_t i = 0; i < m; ++i) out[i] = recs[i];
        *n = m;
        return WW_OK;
    }
    int read_prps(uint64_t *out, uint64_t cap, uint64_t *n) override
    {
        if (fail_reads) return WW_E_HIP;
        uint64_t m = list.size() < cap ? list.size() : cap;
        for (uint64_t i = 0; i < m; ++i) out[i] = list[i];
        *n = list.size();
        return WW_OK;
    }
};

ww_result rec(uint64_t p, int32_t code, int32_t rem, int32_t quot)
{
    ww_result r;
    r.p = p;
    r.result = code;
    r.rem = rem;
    r.quot = quot;
    r.pad = 0;
    return r;
}

uint64_t num(const char *s) { return strtoull(s, nullptr, 10); }

/* Records "p code rem quot" one per line, from a file. */
std::vector<ww_result> read_records(const char *path)
{
    std::vector<ww_result> v;
    std::ifstream in(path);
    unsigned long long p;
    int c, r, q;
    while (in >> p >> c >> r >> q) v.push_back(rec(p, c, r, q));
    return v;
}

std::vector<uint64_t> read_list(const char *path)
{
    std::vector<uint64_t> v;
    std::ifstream in(path);
    unsigned long long p;
    while (in >> p) v.push_back(p);
    return v;
}

int run(int argc, char **argv)
{
    const std::string sc = argc > 1 ? argv[1] : "";
    FakeView dev;
    if (sc == "results") {
        /* results RECORDS_FILE RESULTCOUNT PRIMECOUNT_MAX LIST_CAP NUM_RESULTS */
        if (argc != 7) return 2;
        dev.recs = read_records(argv[2]);
        dev.k.resultcount = (uint32_t)num(argv[3]);
        dev.k.primecount_max = (uint32_t)num(argv[4]);
        dev.k.checksum = 0x0123456789abcdefull;
        dev.k.totalprimecount = 12345;
        const wwhost::CheckpointReport rep =
            wwhost::collect_results(dev, num(argv[5]), (uint32_t)num(argv[6]));
        printf("OK dropped=%u checksum=%016" PRIX64 "\n%s", rep.dropped, rep.counters.checksum,
               rep.text.c_str());
        return 0;
    }
    if (sc == "prps") {
        /* prps LIST_FILE LO HI */
        if (argc != 5) return 2;
        dev.list = read_list(argv[2]);
        const uint32_t n = wwhost::check_prps(dev, num(argv[3]), num(argv[4]));
        printf("OK prps=%u\n", n);
        return 0;
    }
    if (sc == "selftest") {
        /* selftest GPU_PRPS GPU_CHECKSUM GPU_RESULTS CPU_PRPS CPU_CHECKSUM CPU_RESULTS */
        if (argc != 8) return 2;
        ww_counters k{};
        k.checksum = num(argv[3]);
        k.resultcount = (uint32_t)num(argv[4]);
        const uint64_t cpu[3] = {num(argv[5]), num(argv[6]), num(argv[7])};
        wwhost::compare_selftest((uint32_t)num(argv[2]), k, cpu);
        printf("OK\n");
        return 0;
    }
    if (sc == "selftest_fused") {
        /* selftest_fused GPU_TESTED GPU_CHECKSUM GPU_RESULTS CPU_PRPS CPU_CHECKSUM CPU_RESULTS */
        if (argc != 8) return 2;
        ww_counters k{};
        k.totalprimecount = num(argv[2]);
        k.checksum = num(argv[3]);
        k.resultcount = (uint32_t)num(argv[4]);
        const uint64_t cpu[3] = {num(argv[5]), num(argv[6]), num(argv[7])};
        wwhost::compare_selftest_fused(k, cpu);
        printf("OK\n");
        return 0;
    }
    if (sc == "read_failure") {
        dev.fail_reads = true;
        (void)wwhost::collect_results(dev, 100, 100);
        return 0;
    }
    return 2;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "default_hook") == 0) {
        wwhost::fatal(true, "Error: injected failure through the program's own hook.\n");
    }
    wwhost::set_fatal_hook(throwing_hook);
    try {
        const int rc = run(argc, argv);
        if (rc == 2) fprintf(stderr, "bad scenario or arguments\n");
        return rc;
    } catch (const Fatal &f) {
        printf("FATAL both=%d\n%s", f.both ? 1 : 0, f.msg.c_str());
        return 3;
    }
}
