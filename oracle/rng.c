/*
 * rng.c -- restatement of the random generator behind the reference's seeded
 * synthetic BED files (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * src/bench.rs:75-104 (generate_bed_file) draws from rand 0.8.5 `StdRng`
 * (src/bench.rs:9-12, Cargo.lock rand 0.8.5 / rand_chacha 0.3.1 /
 * rand_core 0.6.4).  None of these crates is vendored under the reference, so
 * their published algorithms are restated here:
 *   - StdRng in rand 0.8 is ChaCha12 (D. J. Bernstein's ChaCha with 12 rounds,
 *     original 64-bit block counter in words 12-13 and a 64-bit stream id in
 *     words 14-15, both zero after seeding).  Its u32 output is the keystream
 *     read as little-endian words, block after block; u64 output is two
 *     consecutive words, low first.
 *   - rand_core 0.6 `SeedableRng::seed_from_u64` fills the 32-byte seed with
 *     PCG32 outputs (state advanced first with the PCG multiplier and
 *     increment, then the XSH-RR output function), little-endian.
 *   - `Rng::gen_range` on u32 (UniformInt::sample_single_inclusive, rand
 *     0.8.5): range = hi-lo+1, zone = (range << clz(range)) - 1, draw v,
 *     64-bit widening multiply v*range; accept when the low word <= zone and
 *     return lo + high word.  An exclusive range a..b samples a..=b-1.
 *   - `SliceRandom::choose` on a slice of length n <= u32::MAX draws
 *     gen_range(0..n as u32).
 * Parity is anchored on the reference's own goldens: tests/integration.rs:73-74
 * (158 hits, 413,652 bp) and :113-114 (nnz 10, total 58) are only reproduced
 * when this generator and the sweep are both right; the ChaCha core is pinned
 * separately by the published ChaCha20 all-zero key/nonce test vector.
 */
#include "oracle.h"

#include <string.h>

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

#define QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

void orc_chacha_block(const uint32_t key[8], uint64_t ctr, uint64_t stream,
                      int rounds, uint32_t out[16]) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                       key[0], key[1], key[2], key[3],
                       key[4], key[5], key[6], key[7],
                       (uint32_t)ctr, (uint32_t)(ctr >> 32),
                       (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

void orc_rng_from_seed(orc_rng* r, const uint8_t seed[32], int rounds) {
    for (int i = 0; i < 8; i++)
        r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                    ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    r->ctr = 0;
    r->idx = 16;
    r->rounds = rounds;
}

/* rand_core 0.6.4 seed_from_u64: PCG32 expansion of the u64 into 32 bytes. */
void orc_rng_seed_u64(orc_rng* r, uint64_t state) {
    const uint64_t MUL = 6364136223846793005ull;
    const uint64_t INC = 11634580027462260723ull;
    uint8_t seed[32];
    for (int i = 0; i < 8; i++) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        uint32_t x = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
        seed[4 * i] = (uint8_t)x;
        seed[4 * i + 1] = (uint8_t)(x >> 8);
        seed[4 * i + 2] = (uint8_t)(x >> 16);
        seed[4 * i + 3] = (uint8_t)(x >> 24);
    }
    orc_rng_from_seed(r, seed, 12);
}

uint32_t orc_rng_next_u32(orc_rng* r) {
    if (r->idx >= 16) {
        orc_chacha_block(r->key, r->ctr, 0, r->rounds, r->buf);
        r->ctr++;
        r->idx = 0;
    }
    return r->buf[r->idx++];
}

uint64_t orc_rng_next_u64(orc_rng* r) {
    uint64_t lo = orc_rng_next_u32(r);
    uint64_t hi = orc_rng_next_u32(r);
    return lo | (hi << 32);
}

uint32_t orc_rng_range_u32_incl(orc_rng* r, uint32_t lo, uint32_t hi) {
    uint32_t range = hi - lo + 1u;
    if (range == 0) return orc_rng_next_u32(r);
    uint32_t zone = (range << __builtin_clz(range)) - 1u;
    for (;;) {
        uint32_t v = orc_rng_next_u32(r);
        uint64_t m = (uint64_t)v * (uint64_t)range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}

/* src/bench.rs:75-104, unsorted branch (:94-101): per interval draw the
 * chromosome (choose over the 24 names of :15-19), the length in
 * min_len..=max_len, then start in 0..genome_size.saturating_sub(len). */
void orc_generate_bed(uint64_t seed, size_t n, uint32_t genome_size,
                      uint32_t min_len, uint32_t max_len, uint8_t* chrom,
                      uint32_t* start, uint32_t* end) {
    orc_rng r;
    orc_rng_seed_u64(&r, seed);
    for (size_t i = 0; i < n; i++) {
        uint32_t c = orc_rng_range_u32_incl(&r, 0, 23);
        uint32_t len = orc_rng_range_u32_incl(&r, min_len, max_len);
        uint32_t hi = genome_size > len ? genome_size - len : 0;
        uint32_t s = orc_rng_range_u32_incl(&r, 0, hi - 1u);
        chrom[i] = (uint8_t)c;
        start[i] = s;
        end[i] = s + len;
    }
}
