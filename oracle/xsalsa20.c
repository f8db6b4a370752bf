/*
 * Oracle: XSalsa20 keystream and the samplers built on it.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference draws its randomness from a Salsa20 core in counter mode
 * (RNG::generateRandom_device, include/RNG.cuh:18-51; VecCrypt,
 * include/RNG_impl.cuh:70-177) whose key and nonce are re-drawn from
 * std::random_device on every call and whose sigma constant is never
 * initialised, so its output cannot be reproduced or pinned.  This engine uses
 * a real, seeded XSalsa20 (D. J. Bernstein, "Extending the Salsa20 nonce",
 * 2011; "Salsa20 specification", 2005): HSalsa20(key, nonce[0:16]) -> subkey,
 * then Salsa20/20(subkey, nonce[16:24], block counter).
 *
 * Sampler stream layout (shared with the HIP samplers): the stream of
 * (seed, tag, entry) is XSalsa20 with key = seed and nonce = tag (8 bytes
 * little endian) || 8 zero bytes || entry (8 bytes little endian); batch entry b
 * of a batched operation reads stream entry b.  Element e of a sampler that consumes w
 * bytes per element reads keystream bytes [e*w, e*w + w):
 *   uniform mod q   w = 16, element = limb*N + k, value = (hi:lo) mod q
 *                   (bias < 2^-67 for 60-bit q)      (cf. convert_uniform_xq,
 *                                                      include/Sampler_impl.cuh:13)
 *   ternary         w = 8, element = k, value = (u64 mod 3) - 1
 *   zero-one (v)    w = 1, element = k, byte < 128 -> 0, < 192 -> 1, else -1
 *                   (the probabilities of convert_ZO, include/Sampler_impl.cuh:89)
 *   CBD, k = 21     w = 8, element = k, popcount(x & M) - popcount((x>>21) & M),
 *                   M = 2^21-1, sigma = sqrt(10.5) ~ 3.24 (replaces the
 *                   reference's inverse-CDF Gaussian, include/Sampler_impl.cuh:28)
 * Coefficients that do not depend on the limb are written to every limb of
 * the requested range as residues (negative values as q - |v|).
 *
 * Pinned by (tests/test_oracle_xsalsa20.py): the published NaCl XSalsa20
 * vector (first 32 bytes for key "firstkey" / its nonce), an independent
 * pure-Python XSalsa20 of the specification on random keys, and
 * distribution tests (counts, moments, chi-square) for the samplers.
 */
#include <stdlib.h>
#include <string.h>

#include "ctx_internal.h"

static inline uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }
static inline uint32_t ld32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static inline void st32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

static const uint32_t SIGMA[4] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};

/* 20 rounds (10 double rounds) of the Salsa20 core; with add_input the input
 * words are added back (the Salsa20 hash), without it the raw state is kept
 * (HSalsa20). */
void or_salsa20_core(const uint32_t in[16], uint32_t out[16], int add_input) {
    uint32_t x[16];
    memcpy(x, in, sizeof(x));
    for (int r = 0; r < 10; r++) {
        /* column round */
        x[4] ^= rotl(x[0] + x[12], 7);
        x[8] ^= rotl(x[4] + x[0], 9);
        x[12] ^= rotl(x[8] + x[4], 13);
        x[0] ^= rotl(x[12] + x[8], 18);
        x[9] ^= rotl(x[5] + x[1], 7);
        x[13] ^= rotl(x[9] + x[5], 9);
        x[1] ^= rotl(x[13] + x[9], 13);
        x[5] ^= rotl(x[1] + x[13], 18);
        x[14] ^= rotl(x[10] + x[6], 7);
        x[2] ^= rotl(x[14] + x[10], 9);
        x[6] ^= rotl(x[2] + x[14], 13);
        x[10] ^= rotl(x[6] + x[2], 18);
        x[3] ^= rotl(x[15] + x[11], 7);
        x[7] ^= rotl(x[3] + x[15], 9);
        x[11] ^= rotl(x[7] + x[3], 13);
        x[15] ^= rotl(x[11] + x[7], 18);
        /* row round */
        x[1] ^= rotl(x[0] + x[3], 7);
        x[2] ^= rotl(x[1] + x[0], 9);
        x[3] ^= rotl(x[2] + x[1], 13);
        x[0] ^= rotl(x[3] + x[2], 18);
        x[6] ^= rotl(x[5] + x[4], 7);
        x[7] ^= rotl(x[6] + x[5], 9);
        x[4] ^= rotl(x[7] + x[6], 13);
        x[5] ^= rotl(x[4] + x[7], 18);
        x[11] ^= rotl(x[10] + x[9], 7);
        x[8] ^= rotl(x[11] + x[10], 9);
        x[9] ^= rotl(x[8] + x[11], 13);
        x[10] ^= rotl(x[9] + x[8], 18);
        x[12] ^= rotl(x[15] + x[14], 7);
        x[13] ^= rotl(x[12] + x[15], 9);
        x[14] ^= rotl(x[13] + x[12], 13);
        x[15] ^= rotl(x[14] + x[13], 18);
    }
    for (int i = 0; i < 16; i++) out[i] = add_input ? x[i] + in[i] : x[i];
}

void or_hsalsa20(const uint8_t key[32], const uint8_t nonce16[16], uint8_t subkey[32]) {
    uint32_t s[16], z[16];
    s[0] = SIGMA[0];
    s[5] = SIGMA[1];
    s[10] = SIGMA[2];
    s[15] = SIGMA[3];
    for (int i = 0; i < 4; i++) {
        s[1 + i] = ld32(key + 4 * i);
        s[11 + i] = ld32(key + 16 + 4 * i);
        s[6 + i] = ld32(nonce16 + 4 * i);
    }
    or_salsa20_core(s, z, 0);
    static const int pick[8] = {0, 5, 10, 15, 6, 7, 8, 9};
    for (int i = 0; i < 8; i++) st32(subkey + 4 * i, z[pick[i]]);
}

static void salsa20_block(const uint8_t subkey[32], const uint8_t nonce8[8], uint64_t ctr,
                          uint8_t out[64]) {
    uint32_t s[16], z[16];
    s[0] = SIGMA[0];
    s[5] = SIGMA[1];
    s[10] = SIGMA[2];
    s[15] = SIGMA[3];
    for (int i = 0; i < 4; i++) {
        s[1 + i] = ld32(subkey + 4 * i);
        s[11 + i] = ld32(subkey + 16 + 4 * i);
    }
    s[6] = ld32(nonce8);
    s[7] = ld32(nonce8 + 4);
    s[8] = (uint32_t)ctr;
    s[9] = (uint32_t)(ctr >> 32);
    or_salsa20_core(s, z, 1);
    for (int i = 0; i < 16; i++) st32(out + 4 * i, z[i]);
}

void or_xsalsa20_stream(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0,
                        uint8_t *out, size_t len) {
    uint8_t sub[32], blk[64];
    or_hsalsa20(key, nonce, sub);
    uint64_t ctr = block0;
    for (size_t off = 0; off < len; off += 64, ctr++) {
        salsa20_block(sub, nonce + 16, ctr, blk);
        size_t take = len - off < 64 ? len - off : 64;
        memcpy(out + off, blk, take);
    }
}

typedef struct {
    uint8_t sub[32];
    uint8_t n8[8];
    uint64_t cached;
    uint8_t blk[64];
} stream_t;

static void stream_init(stream_t *s, const uint8_t seed[32], uint64_t tag, uint64_t entry) {
    uint8_t nonce[24];
    memset(nonce, 0, sizeof(nonce));
    for (int i = 0; i < 8; i++) nonce[i] = (uint8_t)(tag >> (8 * i));
    for (int i = 0; i < 8; i++) nonce[16 + i] = (uint8_t)(entry >> (8 * i));
    or_hsalsa20(seed, nonce, s->sub);
    memcpy(s->n8, nonce + 16, 8);
    s->cached = (uint64_t)-1;
}
static const uint8_t *stream_at(stream_t *s, uint64_t byte_off) {
    uint64_t b = byte_off >> 6;
    if (b != s->cached) {
        salsa20_block(s->sub, s->n8, b, s->blk);
        s->cached = b;
    }
    return s->blk + (byte_off & 63);
}
static uint64_t ld64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

void or_sample_uniform(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                       int mod0, int limbs, uint64_t *out) {
    stream_t s;
    stream_init(&s, seed, tag, entry);
    size_t N = c->n;
    for (int i = 0; i < limbs; i++) {
        uint64_t q = c->mod[mod0 + i];
        for (size_t k = 0; k < N; k++) {
            const uint8_t *p = stream_at(&s, 16 * ((uint64_t)i * N + k));
            u128 v = ((u128)ld64(p + 8) << 64) | ld64(p);
            out[(size_t)i * N + k] = (uint64_t)(v % q);
        }
    }
}

static void spread(const or_ctx *c, const int64_t *v, int mod0, int limbs, uint64_t *out) {
    size_t N = c->n;
    for (int i = 0; i < limbs; i++) {
        uint64_t q = c->mod[mod0 + i];
        for (size_t k = 0; k < N; k++)
            out[(size_t)i * N + k] = v[k] < 0 ? q - (uint64_t)(-v[k]) : (uint64_t)v[k];
    }
}

void or_sample_ternary(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                       int mod0, int limbs, uint64_t *out) {
    stream_t s;
    stream_init(&s, seed, tag, entry);
    size_t N = c->n;
    int64_t *v = (int64_t *)malloc(N * sizeof(int64_t));
    for (size_t k = 0; k < N; k++) v[k] = (int64_t)(ld64(stream_at(&s, 8 * k)) % 3) - 1;
    spread(c, v, mod0, limbs, out);
    free(v);
}

void or_sample_zo(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                  int mod0, int limbs, uint64_t *out) {
    stream_t s;
    stream_init(&s, seed, tag, entry);
    size_t N = c->n;
    int64_t *v = (int64_t *)malloc(N * sizeof(int64_t));
    for (size_t k = 0; k < N; k++) {
        uint8_t b = *stream_at(&s, k);
        v[k] = b < 128 ? 0 : (b < 192 ? 1 : -1);
    }
    spread(c, v, mod0, limbs, out);
    free(v);
}

void or_sample_cbd(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                   int mod0, int limbs, uint64_t *out) {
    stream_t s;
    stream_init(&s, seed, tag, entry);
    size_t N = c->n;
    int64_t *v = (int64_t *)malloc(N * sizeof(int64_t));
    const uint64_t M = (1ULL << 21) - 1;
    for (size_t k = 0; k < N; k++) {
        uint64_t x = ld64(stream_at(&s, 8 * k));
        v[k] = (int64_t)__builtin_popcountll(x & M) - (int64_t)__builtin_popcountll((x >> 21) & M);
    }
    spread(c, v, mod0, limbs, out);
    free(v);
}
