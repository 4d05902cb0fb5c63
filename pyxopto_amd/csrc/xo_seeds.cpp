/*
 * xo_seeds.cpp — host side of the MWC random number streams.
 *
 *  - xo_safeprime_multipliers regenerates the multiplier column of the
 *    reference's seed table (xopto/data/primes/safeprimes_base32_500k.npz):
 *    counting down from a = 4294967118, a is kept when a*2^32-1 and
 *    (a*2^32-2)/2 = a*2^31-1 are both prime (xopto/src/rng/find_primes.c:43-93
 *    searches the same sequence with GMP). Primality is decided exactly for
 *    64-bit integers by Miller-Rabin with the first twelve prime bases, which
 *    has no pseudoprime below 3.3e24 (Sorenson & Webster 2015), after trial
 *    division by small primes. Blocks of candidates are tested in parallel
 *    with std::thread.
 *  - xo_init_rng restates init_RNG (xopto/src/rng/rng.cpp:64-103).
 *  - xo_rng_host is the RngKernel test harness
 *    (xopto/mcbase/kernel/mcbase.template.c:1640-1646) on the host.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/xomc.h"
#include "xo_internal.h"

namespace {

typedef unsigned __int128 u128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)((u128)a * b % m);
}

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1u) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

bool mr_witness_passes(uint64_t n, uint64_t a, uint64_t d, int s) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) return true;
    for (int i = 1; i < s; ++i) {
        x = mulmod(x, x, n);
        if (x == n - 1) return true;
    }
    return false;
}

bool is_prime_u64(uint64_t n) {
    static const uint32_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint32_t p : bases) {
        if (n % p == 0) return n == p;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1u)) {
        d >>= 1;
        ++s;
    }
    for (uint32_t a : bases)
        if (!mr_witness_passes(n, a, d, s)) return false;
    return true;
}

std::vector<uint32_t> small_odd_primes(int count) {
    std::vector<uint32_t> out;
    for (uint32_t p = 3; (int)out.size() < count; p += 2) {
        bool ok = true;
        for (uint32_t q = 3; q * q <= p; q += 2)
            if (p % q == 0) { ok = false; break; }
        if (ok) out.push_back(p);
    }
    return out;
}

std::mutex g_table_mutex;
std::vector<uint32_t> g_table; /* multipliers found so far, in order */
uint64_t g_next_a = XO_FIRST_PRIME_A;

/* Extend g_table to at least n entries. Caller holds g_table_mutex. */
void extend_table(uint32_t n) {
    static const std::vector<uint32_t> small = small_odd_primes(1500);
    unsigned nthreads = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    const uint64_t block = 1u << 16;
    std::vector<uint8_t> flags(block);
    while (g_table.size() < n) {
        const uint64_t a0 = g_next_a;
        auto work = [&](unsigned t) {
            for (uint64_t i = t; i < block; i += nthreads) {
                uint64_t a = a0 - i;
                uint8_t ok = 0;
                if (a >= 2) {
                    uint64_t n2 = (a << 32) - 1u, n1 = (a << 31) - 1u;
                    ok = 1;
                    for (uint32_t p : small) {
                        if (n2 % p == 0 || n1 % p == 0) { ok = 0; break; }
                    }
                    if (ok) ok = is_prime_u64(n2) && is_prime_u64(n1);
                }
                flags[i] = ok;
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
        for (uint64_t i = 0; i < block && g_table.size() < n; ++i)
            if (flags[i]) g_table.push_back((uint32_t)(a0 - i));
        g_next_a = a0 - block;
        if (g_next_a < 2) break;
    }
}

} // namespace

namespace xo {

/* Install a precomputed prefix of the table (from the package's cache file);
 * entries beyond what is already known are appended. */
int table_preload(const uint32_t *a, uint32_t n) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (n <= g_table.size()) return 0;
    if (n > 0 && a[0] != XO_FIRST_PRIME_A) return -1;
    g_table.assign(a, a + n);
    g_next_a = (uint64_t)a[n - 1] - 1u;
    return 0;
}

int table_get(uint32_t n, uint32_t *out) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (g_table.size() < n) extend_table(n);
    if (g_table.size() < n) return -1;
    std::memcpy(out, g_table.data(), sizeof(uint32_t) * n);
    return 0;
}

} // namespace xo

extern "C" {

int xo_safeprime_multipliers(uint32_t n, uint32_t *a_out) {
    if (n > XO_NUM_PRIMES) {
        xo::set_error("xo_safeprime_multipliers: at most %u multipliers", XO_NUM_PRIMES);
        return -1;
    }
    if (xo::table_get(n, a_out) != 0) {
        xo::set_error("xo_safeprime_multipliers: search failed");
        return -1;
    }
    return 0;
}

int xo_safeprime_preload(const uint32_t *a, uint32_t n) {
    if (xo::table_preload(a, n) != 0) {
        xo::set_error("xo_safeprime_preload: table does not start at %u", XO_FIRST_PRIME_A);
        return -1;
    }
    return 0;
}

/* init_RNG, xopto/src/rng/rng.cpp:64-103. */
int xo_init_rng(uint32_t n, uint64_t xinit, uint64_t *x, uint32_t *a) {
    if (n > XO_MAX_SEEDS) {
        xo::set_error("Cannot generate more than %u seeds ... requested %u seeds!",
                      XO_MAX_SEEDS, n);
        return -1;
    }
    std::vector<uint32_t> fora(n + 1u);
    if (xo::table_get(n + 1u, fora.data()) != 0) {
        xo::set_error("xo_init_rng: safe-prime table unavailable");
        return -1;
    }
    const uint32_t begin = fora[0];
    if ((xinit == 0ull) | (((uint32_t)(xinit >> 32)) >= (begin - 1u)) |
        (((uint32_t)xinit) >= 0xffffffffu))
        return 1;
    for (uint32_t i = 0; i < n; ++i) {
        a[i] = fora[i + 1];
        x[i] = 0;
        while ((x[i] == 0) | (((uint32_t)(x[i] >> 32)) >= (fora[i + 1] - 1u)) |
               (((uint32_t)x[i]) >= 0xffffffffu)) {
            xinit = (xinit & 0xffffffffull) * begin + (xinit >> 32);
            x[i] = (uint32_t)std::floor(
                (((double)((uint32_t)xinit)) / (double)0x100000000ull) *
                fora[i + 1]);
            x[i] = x[i] << 32;
            xinit = (xinit & 0xffffffffull) * begin + (xinit >> 32);
            x[i] += (uint32_t)xinit;
        }
    }
    return 0;
}

int xo_rng_host(uint64_t *x, uint32_t a, uint32_t n, float *out) {
    uint64_t s = *x;
    for (uint32_t i = 0; i < n; ++i) {
        s = (s & 0xffffffffull) * a + (s >> 32);
        out[i] = (float)(uint32_t)s / 4294967296.0f;
    }
    *x = s;
    return 0;
}

} // extern "C"
