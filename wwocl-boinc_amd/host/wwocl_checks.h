/*
 * wwocl_checks.h -- the drop-in host's result validation and self-test comparison,
 * separated from the device so that every failure branch can be driven on the CPU with
 * injected counters, records and PRP lists.
 *
 *   collect_results   getResults, main.cpp:768-956
 *   check_prps        checkPRPS, main.cpp:451-523
 *   compare_selftest  testKernels' comparisons, main.cpp:703-728
 *
 * A failure prints the reference's message and exits with EXIT_FAILURE, as the
 * reference does; the exit goes through a replaceable hook (tests throw instead).
 */
#ifndef WWOCL_CHECKS_H
#define WWOCL_CHECKS_H

#include <cstdint>
#include <string>

#include "ww_api.h"

namespace wwhost {

/* The device state the checks read: the library's context in the program, injected data
   in tests. Each call returns WW_OK or a WW_E_* code. */
class DeviceView {
public:
    virtual ~DeviceView() = default;
    virtual int read_counters(ww_counters *out) = 0;
    /* up to cap records; *n = records copied */
    virtual int read_results(ww_result *out, uint32_t cap, uint32_t *n) = 0;
    /* up to cap list entries; *n = the list length */
    virtual int read_prps(uint64_t *out, uint64_t cap, uint64_t *n) = 0;
};

/* Where a fatal message goes: `both` = stdout and stderr (else stderr only), then the
   process ends with EXIT_FAILURE unless the hook does not return (a test's hook throws). */
typedef void (*FatalHook)(bool both, const std::string &msg);
void set_fatal_hook(FatalHook hook);

/* Print (both streams or stderr only), then exit with EXIT_FAILURE. */
[[noreturn]] void fatal(bool both, const char *fmt, ...);

/* One checkpoint's records: counters read, overflow checks, records sorted and formatted. */
struct CheckpointReport {
    ww_counters counters;
    std::string text;      /* results-file lines, in the reference's order and formats */
    uint32_t dropped = 0;  /* base-2 pseudoprimes removed from the results */
};

/* The results-file line of one record, validated on the CPU: primes are formatted
   (main.cpp:869-905), composite base-2 PRPs are dropped (main.cpp:907-911), anything else
   is a computation error (main.cpp:912-916). Returns false when dropped. */
bool format_record(const ww_result &r, std::string &line, uint32_t &dropped);

/* getResults without the host bookkeeping (main.cpp:768-856): reads the counters, stops on
   a list longer than list_cap or more than num_results records, then formats them. */
CheckpointReport collect_results(DeviceView &dev, uint64_t list_cap, uint32_t num_results);

/* checkPRPS (main.cpp:451-523): every list entry is prime or a base-2 PRP with no factor
   <= 113, and the primes are exactly those of [lo, hi), found on the CPU. Returns the list
   length. */
uint32_t check_prps(DeviceView &dev, uint64_t lo, uint64_t hi);

/* testKernels' comparisons (main.cpp:703-728): the device's PRP count, checksum and record
   count against the CPU path's {PRP count, checksum, result count}. */
void compare_selftest(uint32_t gpu_prps, const ww_counters &gpu, const uint64_t cpu[3]);

/* The same check of a chunk run through the persistent kernel (its tested count stands for
   the PRP count). */
void compare_selftest_fused(const ww_counters &gpu, const uint64_t cpu[3]);

} // namespace wwhost

#endif
