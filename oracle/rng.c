/* Counter-V1 RNG restatement (test infrastructure).
 * Reference: rust/crates/vckss-core/src/rng.rs and rust/RNG_CONTRACT.md:1-68. */
#include <stdarg.h>
#include <stdio.h>

#include "orc_internal.h"

static _Thread_local char orc_error_text[512];

const char *orc_last_error(void) { return orc_error_text; }

int orc_fail(int code, const char *phase, const char *fmt, ...) {
    va_list args;
    va_start(args, fmt);
    int used = snprintf(orc_error_text, sizeof orc_error_text, "%s: ", phase);
    if (used < 0) used = 0;
    vsnprintf(orc_error_text + used, sizeof orc_error_text - (size_t)used, fmt, args);
    va_end(args);
    return code;
}

/* rng.rs:14-19 */
#define PHILOX_M0 0xd2b74407b1ce6e93ULL
#define PHILOX_M1 0xca5a826395121157ULL
#define PHILOX_W0 0x9e3779b97f4a7c15ULL
#define PHILOX_W1 0xbb67ae8584caa73bULL
#define KEY_XOR 0xd1b54a32d192ed03ULL
#define MAX_WORDS_PER_ATOM (1ULL << 20)
#define MAX_EXACT (1ULL << 53)

/* rng.rs:309-315 */
static uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* rng.rs:73-93 (raw_block) with philox_round rng.rs:317-327 */
void orc_philox_block(uint64_t seed, uint64_t tag, uint64_t probe_block,
                      uint64_t entity, uint64_t word, uint64_t out[4]) {
    uint64_t c0 = entity, c1 = probe_block, c2 = word, c3 = tag;
    uint64_t k0 = splitmix64(seed), k1 = splitmix64(seed ^ KEY_XOR);
    for (int round = 0; round < 10; ++round) {
        unsigned __int128 p0 = (unsigned __int128)PHILOX_M0 * c0;
        unsigned __int128 p1 = (unsigned __int128)PHILOX_M1 * c2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        if (round + 1 != 10) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* rng.rs:95-99 */
uint64_t orc_rng_word(uint64_t seed, uint64_t tag, uint64_t probe,
                      uint64_t entity, uint64_t word) {
    uint64_t block[4];
    orc_philox_block(seed, tag, probe / 4, entity, word, block);
    return block[probe % 4];
}

/* rng.rs:121-179 */
int orc_rademacher_sum(uint64_t seed, uint64_t tag, uint64_t probe,
                       uint64_t entity, uint64_t trials, int64_t *out) {
    if (trials == 0 || trials > MAX_EXACT)
        return orc_fail(ORC_RNG_CONTRACT, "counter_rng",
                        "Rademacher trial count must lie in [1, 2^53]");
    uint64_t words = (trials + 63) / 64;
    if (words > MAX_WORDS_PER_ATOM)
        return orc_fail(ORC_RESOURCE_LIMIT, "counter_rng",
                        "atom needs %llu words", (unsigned long long)words);
    uint64_t complete = trials / 64, remainder = trials % 64, positive = 0;
    for (uint64_t w = 0; w < complete; ++w)
        positive += (uint64_t)__builtin_popcountll(orc_rng_word(seed, tag, probe, entity, w));
    if (remainder) {
        uint64_t mask = (1ULL << remainder) - 1;
        positive += (uint64_t)__builtin_popcountll(
            orc_rng_word(seed, tag, probe, entity, complete) & mask);
    }
    *out = (int64_t)(2 * positive) - (int64_t)trials;
    return ORC_OK;
}

/* rng.rs:202-305: output column-major by probe, out[col * n + row]. */
int orc_rademacher_fill(uint64_t seed, uint64_t tag, uint64_t first_probe,
                        uint64_t columns, const uint64_t *entity,
                        const uint64_t *trials, uint64_t n, int64_t *out) {
    if (columns == 0 || n == 0)
        return orc_fail(ORC_INVALID_INPUT, "counter_rng",
                        "semantic atom request has incompatible dimensions");
    uint64_t last = first_probe + columns - 1;
    uint64_t first_block = first_probe / 4, last_block = last / 4;
    for (uint64_t row = 0; row < n; ++row) {
        uint64_t t = trials[row];
        if (t == 0 || t > MAX_EXACT)
            return orc_fail(ORC_RNG_CONTRACT, "counter_rng",
                            "Rademacher trial count must lie in [1, 2^53]");
        if ((t + 63) / 64 > MAX_WORDS_PER_ATOM)
            return orc_fail(ORC_RESOURCE_LIMIT, "counter_rng", "atom word limit");
        uint64_t complete = t / 64, remainder = t % 64;
        for (uint64_t pb = first_block; pb <= last_block; ++pb) {
            uint64_t positive[4] = {0, 0, 0, 0}, block[4];
            for (uint64_t w = 0; w < complete; ++w) {
                orc_philox_block(seed, tag, pb, entity[row], w, block);
                for (int lane = 0; lane < 4; ++lane)
                    positive[lane] += (uint64_t)__builtin_popcountll(block[lane]);
            }
            if (remainder) {
                uint64_t mask = (1ULL << remainder) - 1;
                orc_philox_block(seed, tag, pb, entity[row], complete, block);
                for (int lane = 0; lane < 4; ++lane)
                    positive[lane] += (uint64_t)__builtin_popcountll(block[lane] & mask);
            }
            for (int lane = 0; lane < 4; ++lane) {
                uint64_t probe = pb * 4 + (uint64_t)lane;
                if (probe < first_probe || probe > last) continue;
                out[(probe - first_probe) * n + row] =
                    (int64_t)(2 * positive[lane]) - (int64_t)t;
            }
        }
    }
    return ORC_OK;
}
