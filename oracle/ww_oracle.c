/*
 * ww_oracle.c -- CPU oracle for the Wieferich / Wall-Sun-Sun search path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg load it; the product never does.
 *
 * A plain-C restatement of the reference CPU path (citations per function). The
 * arithmetic is exact integer arithmetic on unsigned __int128; every function is
 * pinned by tests/test_oracle.py against Python big-integer closed forms, the
 * reference's own wwcpu.cpp/prime.h compiled into oracle/_ref, the known answers of
 * tests.txt and brute force on small inputs.
 */
#include "ww_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define SPECIAL_THRESHOLD 1000u /* WW.h:9 */

/* ---------------------------------------------------------------------------------
 * 64-bit Montgomery arithmetic, R = 2^64 (prime.h:7-59)
 * ------------------------------------------------------------------------------- */

/* p^-1 mod 2^64 by Newton's iteration (prime.h:7-12). */
uint64_t or_invert(uint64_t p)
{
    uint64_t x = 1, prev = 0;
    while (x != prev) {
        prev = x;
        x *= 2 - p * x;
    }
    return x;
}

/* a*b/R mod p for a, b < p (prime.h:15-31). */
uint64_t or_mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q)
{
    u128 t = (u128)a * b;
    uint64_t lo = (uint64_t)t, hi = (uint64_t)(t >> 64);
    uint64_t m = lo * q;
    uint64_t mp = (uint64_t)(((u128)m * p) >> 64);
    uint64_t r = hi - mp;
    return hi < mp ? r + p : r;
}

/* a + b mod p (prime.h:34-43). */
static uint64_t mod_add(uint64_t a, uint64_t b, uint64_t p)
{
    uint64_t c = (a >= p - b) ? p : 0;
    return a + b - c;
}

/* Strong probable-prime test of odd n > 1 to `base` (prime.h:46-112, with the
   Montgomery constants of montInit). */
int or_strong_prp(uint64_t base, uint64_t n)
{
    uint64_t q = or_invert(n);
    uint64_t one = (0 - n) % n;          /* R mod n */
    uint64_t minus_one = n - one;        /* Montgomery form of n-1 */
    uint64_t four = mod_add(mod_add(one, one, n), mod_add(one, one, n), n);
    uint64_t r2 = four;                  /* 4^(2^5) = 2^64 -> R^2 mod n */
    for (int i = 0; i < 5; ++i) r2 = or_mont_mul(r2, r2, n, q);

    uint64_t nm1 = n - 1;
    int t = __builtin_ctzll(nm1);
    uint64_t d = n >> t;                 /* (n-1)/2^t: the shift drops the low 1 bit */
    int lead = 63 - __builtin_clzll(d);  /* bits below the leading one of d */
    uint64_t bit = lead ? 1ull << (lead - 1) : 0;  /* d = 1 (n = 2^t + 1): no bits */

    uint64_t mb = or_mont_mul(base, r2, n, q);
    uint64_t a = mb;
    for (; bit; bit >>= 1) {
        a = or_mont_mul(a, a, n, q);
        if (d & bit) a = or_mont_mul(a, mb, n, q);
    }
    if (a == one || a == minus_one) return 1;
    for (int s = 1; s < t; ++s) {
        a = or_mont_mul(a, a, n, q);
        if (a == minus_one) return 1;
    }
    return 0;
}

/* Odd and free of the prime factors 3..113 (main.cpp:653-682, getsegprimes.cl:258-261). */
int or_is_candidate(uint64_t n)
{
    static const unsigned small[29] = {3,  5,  7,  11, 13, 17, 19, 23, 29, 31,
                                       37, 41, 43, 47, 53, 59, 61, 67, 71, 73,
                                       79, 83, 89, 97, 101, 103, 107, 109, 113};
    if ((n & 1) == 0) return 0;
    for (int i = 0; i < 29; ++i)
        if (n % small[i] == 0) return 0;
    return 1;
}

/* isPRP (main.cpp:245-291): candidate and spsp(2). */
int or_is_prp(uint64_t n)
{
    return or_is_candidate(n) && or_strong_prp(2, n);
}

/* isPrime (main.cpp:301-324): Miller-Rabin to bases 2..23 below 3825123056546413051,
   bases 2..37 above. */
int or_is_prime(uint64_t n)
{
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if ((n & 1) == 0) return 0;
    int nb = n < 3825123056546413051ull ? 9 : 12;
    for (int i = 0; i < nb; ++i)
        if (!or_strong_prp(bases[i], n)) return 0;
    return 1;
}

/* ---------------------------------------------------------------------------------
 * mod-p^2 residues as two base-p digits (x0 + p*x1), Montgomery factor 2^64
 * (wwcpu.cpp:19-250; Dorais & Klyve, J. Integer Seq. 14 (2011), "double precision"
 * Montgomery variant).
 * ------------------------------------------------------------------------------- */

typedef struct { uint64_t d0, d1; } m2p;

/* x - y mod p and its borrow (wwcpu.cpp:89-102). */
static uint64_t sub_mod_b(uint64_t x, uint64_t y, uint64_t p, uint64_t *borrow)
{
    *borrow = x < y;
    return x - y + (x < y ? p : 0);
}
/* x + y mod p and its carry (wwcpu.cpp:73-86). */
static uint64_t add_mod_c(uint64_t x, uint64_t y, uint64_t p, uint64_t *carry)
{
    int c = x >= p - y;
    *carry = (uint64_t)c;
    return x + y - (c ? p : 0);
}

/* 2^64 mod p^2 in digits (wwcpu.cpp:109-125). */
static m2p m2p_one(uint64_t p)
{
    m2p r;
    if ((p >> 32) == 0) {
        uint64_t p2 = p * p;
        uint64_t rr = (0 - p2) % p2;
        r.d0 = rr % p;
        r.d1 = rr / p;
    } else {
        uint64_t mp = 0 - p;
        r.d0 = mp % p;
        r.d1 = mp / p + 1;
    }
    return r;
}

void or_m2p_one(uint64_t p, uint64_t out[2])
{
    m2p r = m2p_one(p);
    out[0] = r.d0;
    out[1] = r.d1;
}

/* 2x (wwcpu.cpp:128-137) and x + y (wwcpu.cpp:241-250). */
static m2p m2p_add(m2p x, m2p y, uint64_t p)
{
    uint64_t c;
    m2p r;
    r.d0 = add_mod_c(x.d0, y.d0, p, &c);
    r.d1 = mod_add(x.d1 + c, y.d1, p);
    return r;
}

/* Shared tail of square / mul: given the low 128-bit product t = x0*y0 and the
   cross term cross = x0*y1 + x1*y0 (up to 129 bits: cross_hi_carry), reduce
   (wwcpu.cpp:143-166, 213-236). */
static m2p m2p_reduce(u128 t, u128 cross_lo128, int cross_carry, uint64_t p, uint64_t q)
{
    uint64_t u0 = q * (uint64_t)t;
    uint64_t t1 = (uint64_t)(t >> 64);
    uint64_t v1 = (uint64_t)(((u128)p * u0) >> 64);
    u128 tp = cross_lo128 + (u128)u0;
    int carry = cross_carry | (tp < cross_lo128);
    uint64_t tp_h = (uint64_t)(tp >> 64);
    uint64_t tpc = (carry || tp_h >= p) ? p : 0;
    uint64_t up0 = q * (uint64_t)tp;
    uint64_t t1p = tp_h - tpc;
    uint64_t v1p = (uint64_t)(((u128)p * up0) >> 64);
    uint64_t b;
    m2p r;
    r.d0 = sub_mod_b(t1, v1, p, &b);
    r.d1 = (t1p < v1p + b) ? t1p - (v1p + b) + p : t1p - (v1p + b);
    return r;
}

static m2p m2p_square(m2p x, uint64_t p, uint64_t q)
{
    u128 t = (u128)x.d0 * x.d0;
    u128 x01 = (u128)x.d0 * x.d1;
    u128 two = x01 + x01;
    return m2p_reduce(t, two, two < x01, p, q);
}

static m2p m2p_mul(m2p x, m2p y, uint64_t p, uint64_t q)
{
    u128 t = (u128)x.d0 * y.d0;
    u128 a = (u128)x.d0 * y.d1;
    u128 b = (u128)x.d1 * y.d0;
    u128 s = a + b;
    return m2p_reduce(t, s, s < b, p, q);
}

/* REDC back to ordinary digits (wwcpu.cpp:189-206). */
static m2p m2p_get(m2p x, uint64_t p, uint64_t q)
{
    uint64_t u0 = q * x.d0;
    uint64_t v1 = (uint64_t)(((u128)p * u0) >> 64);
    uint64_t tp = x.d1 + u0;
    uint64_t up0 = q * tp;
    uint64_t t1p = tp < x.d1 ? 1 : 0;
    uint64_t v1p = (uint64_t)(((u128)p * up0) >> 64);
    uint64_t b;
    m2p r;
    r.d0 = sub_mod_b(0, v1, p, &b);
    r.d1 = (t1p < v1p + b) ? t1p - (v1p + b) + p : t1p - (v1p + b);
    return r;
}

/* s = 2^((p-1)/2) mod p^2 as digits (wwcpu.cpp:253-284). The reference's single-digit
   fast path (wwcpu.cpp:270-275) yields the same values as the general square. */
void or_wieferich(uint64_t p, uint64_t out[2])
{
    uint64_t q = or_invert(p);
    m2p one = m2p_one(p);
    uint64_t e = p >> 1;
    uint64_t bit = 1ull << (62 - __builtin_clzll(e));
    m2p a = m2p_add(one, one, p);   /* 2 */
    a = m2p_add(a, a, p);           /* 2^2 */
    if (e & bit) a = m2p_add(a, a, p);
    for (bit >>= 1; bit; bit >>= 1) {
        a = m2p_square(a, p, q);
        if (e & bit) a = m2p_add(a, a, p);
    }
    m2p s = m2p_get(a, p, q);
    out[0] = s.d0;
    out[1] = s.d1;
}

/* Single-digit product of residues with a zero high digit (wwcpu.cpp:170-186).
   As written there the product is x*x: the second operand is not read
   (wwcpu.cpp:173), unlike the device kernel (wallsunsun.cl:190). */
static m2p m2p_mul_s(uint64_t x, uint64_t y, uint64_t p, uint64_t q)
{
    (void)y;
    return m2p_mul((m2p){x, 0}, (m2p){x, 0}, p, q);
}

/* F_{p - (p/5)} mod p^2 as digits by the (F_n, F_{n-1}) doubling ladder
   (wwcpu.cpp:306-372).
   The first loop is the reference's single-digit phase (wwcpu.cpp:323-344): it runs
   while the high digit of F_n is zero and squares F_{n-1} from its low digit only,
   so when F_{n-1} has a nonzero high digit there it departs from the true
   Fibonacci value. That happens only for p < 2^32 (for larger p the high digit of
   2^64 mod p^2 is floor(2^64/p) >= 1 and the phase never runs); the checksum
   follows the reference, so the oracle restates the phase as written. */
void or_wallsunsun(uint64_t p, uint64_t out[2])
{
    uint64_t q = or_invert(p);
    m2p one = m2p_one(p);
    int pm5 = (int)(p % 5);
    int kro = (pm5 == 2 || pm5 == 3) ? -1 : 1;
    uint64_t e = p - (uint64_t)(int64_t)kro - 1;
    uint64_t bit = 1ull << (62 - __builtin_clzll(e));
    m2p b = one, c = {0, 0};
    for (; bit && b.d1 == 0; bit >>= 1) {
        m2p b2 = m2p_mul_s(b.d0, b.d0, p, q);
        m2p c2 = m2p_mul_s(c.d0, c.d0, p, q);
        m2p a = m2p_add(b, c, p);
        m2p ac = m2p_add(a, c, p);
        b = ac.d1 == 0 ? m2p_mul_s(b.d0, ac.d0, p, q) : m2p_mul(b, ac, p, q);
        c = m2p_add(b2, c2, p);
        if (e & bit) {
            m2p s = m2p_add(b, c, p);
            c = b;
            b = s;
        }
    }
    for (; bit; bit >>= 1) {
        m2p b2 = m2p_square(b, p, q);
        m2p c2 = m2p_square(c, p, q);
        m2p a = m2p_add(b, c, p);
        b = m2p_mul(b, m2p_add(a, c, p), p, q);   /* F_2n = F_n (F_n+1 + F_n-1) */
        c = m2p_add(b2, c2, p);                   /* F_2n-1 = F_n^2 + F_n-1^2 */
        if (e & bit) {
            m2p s = m2p_add(b, c, p);
            c = b;
            b = s;
        }
    }
    m2p f = m2p_get(m2p_add(b, c, p), p, q);
    out[0] = f.d0;
    out[1] = f.d1;
}

/* wieferichCPU's return value and checksum slice (wwcpu.cpp:286-303). */
int or_wieferich_cpu_flag(uint64_t p, uint64_t *checksum)
{
    uint64_t s[2];
    or_wieferich(p, s);
    *checksum += s[1] & 0xffffffffu;
    if ((s[1] == 0 && s[0] == 1) || (s[1] == p - 1 && s[0] == p - 1)) return 1;
    if (s[1] <= SPECIAL_THRESHOLD) return 1;
    if (s[1] >= p - SPECIAL_THRESHOLD) return 1;
    return 0;
}

/* wallsunsunCPU's return value and checksum slice (wwcpu.cpp:377-395). */
int or_wallsunsun_cpu_flag(uint64_t p, uint64_t *checksum)
{
    uint64_t c[2];
    or_wallsunsun(p, c);
    *checksum += c[1] & 0xffffffffu;
    if (c[1] == 0) return 1;
    if (c[1] <= SPECIAL_THRESHOLD) return 1;
    if (c[1] >= p - SPECIAL_THRESHOLD) return 1;
    return 0;
}

/* Device-side Wieferich classification (wieferich.cl:247-293). Returns 1 and fills
   rec when a record is produced. All comparisons are unsigned 64-bit, as in OpenCL. */
int or_classify_wieferich(uint64_t p, uint64_t s0, uint64_t s1, or_result *rec)
{
    int32_t rem, quot;
    if ((s1 == 0 && s0 == 1) || (s1 == p - 1 && s0 == p - 1)) {
        rem = 1;
        quot = 0;
    } else if (s1 <= SPECIAL_THRESHOLD) {
        if (s0 == 1) { rem = 1; quot = (int32_t)s1; }
        else if (s0 == p - 1) { rem = -1; quot = (int32_t)(s1 + 1); }
        else return 0;
    } else if (s1 >= p - SPECIAL_THRESHOLD) {
        if (s0 == 1) { rem = 1; quot = (int32_t)(uint32_t)(s1 - p); }
        else if (s0 == p - 1) { rem = -1; quot = (int32_t)(uint32_t)(s1 + 1 - p); }
        else return 0;
    } else {
        return 0;
    }
    rec->p = p;
    rec->result = 1093;
    rec->rem = rem;
    rec->quot = quot;
    rec->pad = 0;
    return 1;
}

/* Device-side Wall-Sun-Sun classification (wallsunsun.cl:325-349). */
int or_classify_wallsunsun(uint64_t p, uint64_t c21, or_result *rec)
{
    int32_t code, quot;
    if (c21 == 0) { code = 1; quot = 0; }
    else if (c21 <= SPECIAL_THRESHOLD) { code = 2; quot = (int32_t)c21; }
    else if (c21 >= p - SPECIAL_THRESHOLD) { code = 2; quot = (int32_t)(uint32_t)(c21 - p); }
    else return 0;
    rec->p = p;
    rec->result = code;
    rec->rem = 0;
    rec->quot = quot;
    rec->pad = 0;
    return 1;
}

/* findWheelOffset (main.cpp:528-587): first n >= start coprime to 30 and its index
   into the gap table {4,2,4,2,4,6,2,6} (which starts at residue 7). */
void or_find_wheel_offset(uint64_t *start, int32_t *index)
{
    static const int32_t gaps[8] = {4, 2, 4, 2, 4, 6, 2, 6};
    uint64_t n = (*start / 6) * 6 - 1;
    int up = 1;               /* next step of the 6k+-1 walk is +2 when set */
    while (n < *start || n % 5 == 0) {
        n += up ? 2 : 4;
        up = !up;
    }
    *start = n;
    int32_t idx = -1;
    while (idx < 0) {
        if (up) {
            n += 2;
            up = 0;
            if (n % 5 == 0) { n -= 2; idx = 5; }
        } else {
            n += 4;
            up = 1;
            if (n % 5 == 0) { n -= 4; idx = 7; }
        }
    }
    while (n != *start) {
        if (--idx < 0) idx = 7;
        n -= (uint64_t)gaps[idx];
    }
    *index = idx;
}

/* ---------------------------------------------------------------------------------
 * Range loops
 * ------------------------------------------------------------------------------- */

/* The CPU range loop of testKernels (main.cpp:648-697) plus device-style records. */
uint64_t or_range(uint64_t lo, uint64_t hi, or_summary *sum, or_result *out, uint64_t cap)
{
    or_summary s = {0, 0, 0, 0};
    uint64_t nrec = 0;
    uint64_t n = lo | 1;
    if (n < lo) n = hi; /* lo == 2^64-1 edge: nothing below hi */
    for (; n < hi; n += 2) {
        if (!or_is_candidate(n) || !or_strong_prp(2, n)) {
            if (n + 2 < n) break;
            continue;
        }
        s.prp_count++;
        uint64_t w[2], f[2];
        or_wieferich(n, w);
        or_wallsunsun(n, f);
        s.checksum += (w[1] & 0xffffffffu) + (f[1] & 0xffffffffu);
        uint64_t dummy = 0;
        s.cpu_result_count += (uint64_t)or_wieferich_cpu_flag(n, &dummy);
        s.cpu_result_count += (uint64_t)or_wallsunsun_cpu_flag(n, &dummy);
        or_result rec;
        if (or_classify_wieferich(n, w[0], w[1], &rec)) {
            if (out && nrec < cap) out[nrec] = rec;
            nrec++;
        }
        if (or_classify_wallsunsun(n, f[1], &rec)) {
            if (out && nrec < cap) out[nrec] = rec;
            nrec++;
        }
        if (n + 2 < n) break;
    }
    s.result_count = nrec;
    if (sum) *sum = s;
    return nrec < cap ? nrec : cap;
}

typedef struct {
    uint64_t lo, hi;
    or_summary sum;
} slice_job;

static void *slice_worker(void *arg)
{
    slice_job *j = (slice_job *)arg;
    or_range(j->lo, j->hi, &j->sum, NULL, 0);
    return NULL;
}

void or_range_mt(uint64_t lo, uint64_t hi, int threads, or_summary *sum)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    slice_job jobs[256];
    pthread_t tid[256];
    uint64_t width = hi - lo, step = width / (uint64_t)threads;
    for (int i = 0; i < threads; ++i) {
        jobs[i].lo = lo + step * (uint64_t)i;
        jobs[i].hi = (i == threads - 1) ? hi : lo + step * (uint64_t)(i + 1);
        pthread_create(&tid[i], NULL, slice_worker, &jobs[i]);
    }
    or_summary s = {0, 0, 0, 0};
    for (int i = 0; i < threads; ++i) {
        pthread_join(tid[i], NULL);
        s.prp_count += jobs[i].sum.prp_count;
        s.checksum += jobs[i].sum.checksum;
        s.result_count += jobs[i].sum.result_count;
        s.cpu_result_count += jobs[i].sum.cpu_result_count;
    }
    *sum = s;
}

uint64_t or_prp_list(uint64_t lo, uint64_t hi, uint64_t *out, uint64_t cap)
{
    uint64_t k = 0;
    uint64_t n = lo | 1;
    if (n < lo) return 0;
    for (; n < hi; n += 2) {
        if (or_is_candidate(n) && or_strong_prp(2, n)) {
            if (k < cap) out[k] = n;
            k++;
        }
        if (n + 2 < n) break;
    }
    return k;
}

/* Segmented sieve of Eratosthenes over [lo, hi): odd-only window bitmap; base primes
   up to sqrt(hi) are themselves produced segment by segment, so hi may reach 2^64-1.
   Returns the number of primes found (entries beyond cap are counted, not stored),
   or UINT64_MAX on allocation failure. */
uint64_t or_sieve_primes(uint64_t lo, uint64_t hi, uint64_t *out, uint64_t cap)
{
    if (hi <= lo) return 0;
    uint64_t k = 0;
    if (lo <= 2 && hi > 2) {
        if (k < cap) out[k] = 2;
        k++;
    }
    uint64_t first = lo | 1;           /* first odd >= lo */
    if (first < 3) first = 3;
    if (first >= hi || first < lo) return k;
    uint64_t nodd = (hi - first + 1) / 2; /* odd numbers first, first+2, ... < hi */
    uint8_t *comp = (uint8_t *)calloc(nodd, 1);
    if (!comp) return UINT64_MAX;

    /* limit = floor(sqrt(hi - 1)) */
    uint64_t limit = 0;
    {
        uint64_t x = hi - 1;
        for (int b = 31; b >= 0; --b) {       /* bitwise integer square root */
            uint64_t c = limit | (1ull << b);
            if ((u128)c * c <= x) limit = c;
        }
    }
    /* base primes: small sieve up to sqrt(limit), then segments over [3, limit] */
    uint64_t slim = 1;
    while ((slim + 1) * (slim + 1) <= limit) slim++;
    uint8_t *small = (uint8_t *)calloc(slim + 1, 1);
    if (!small) { free(comp); return UINT64_MAX; }
    for (uint64_t i = 2; i * i <= slim; ++i)
        if (!small[i])
            for (uint64_t j = i * i; j <= slim; j += i) small[j] = 1;
    const uint64_t SEG = 1u << 20;
    uint8_t *seg = (uint8_t *)malloc(SEG);
    if (!seg) { free(comp); free(small); return UINT64_MAX; }
    for (uint64_t s0 = 3; s0 <= limit; s0 += SEG) {
        uint64_t s1 = s0 + SEG - 1 < limit ? s0 + SEG - 1 : limit;
        memset(seg, 0, (size_t)(s1 - s0 + 1));
        for (uint64_t i = 2; i <= slim && i * i <= s1; ++i) {
            if (small[i]) continue;
            uint64_t j = ((s0 + i - 1) / i) * i;
            if (j < i * i) j = i * i;
            for (; j <= s1; j += i) seg[j - s0] = 1;
        }
        for (uint64_t v = s0 | 1; v <= s1; v += 2) {
            if (seg[v - s0]) continue;
            /* v is an odd base prime: cross off odd multiples m >= max(v*v, first) */
            u128 sq = (u128)v * v;
            uint64_t m;
            if (sq >= (u128)first) {
                if (sq >= (u128)hi) continue;
                m = (uint64_t)sq;
            } else {
                uint64_t r = first % v;
                m = r ? first + (v - r) : first;
                if ((m & 1) == 0) m += v;
                if (m < first) continue; /* wrapped */
            }
            for (; m < hi; m += 2 * v) {
                comp[(m - first) / 2] = 1;
                if (m + 2 * v < m) break;
            }
        }
    }
    for (uint64_t i = 0; i < nodd; ++i) {
        if (!comp[i]) {
            if (k < cap) out[k] = first + 2 * i;
            k++;
        }
    }
    free(seg);
    free(small);
    free(comp);
    return k;
}
