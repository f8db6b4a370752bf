// Seeded XSalsa20 keystream and the samplers fused onto it.
//
// Replaces RNG::generateRandom_device + VecCrypt (include/RNG.cuh:18-51,
// include/RNG_impl.cuh:70-177) and the Sampler conversion kernels
// (include/Sampler_impl.cuh:13-107).  The reference re-seeds from
// std::random_device on every call and never sets the Salsa20 constant; this
// is a real XSalsa20 (Bernstein 2011): HSalsa20(key, nonce[0:16]) on the host,
// then one Salsa20/20 block per thread, converted to samples in registers (no
// byte buffer round trip through HBM).
//
// Stream of (seed, tag, entry): key = seed, nonce = tag (8 bytes LE) || 8 zero
// bytes || entry (8 bytes LE), so one HSalsa20 subkey serves a whole batch and
// batch entry b reads stream entry b.  Element e of a sampler with w bytes per element reads keystream bytes
// [e*w, e*w+w): uniform mod q (w = 16, element limb*N + k, value (hi:lo) mod q),
// ternary (w = 8, (u64 mod 3) - 1), zero-one (w = 1, byte < 128 -> 0,
// < 192 -> 1, else -1), CBD k = 21 (w = 8, popcount difference of two 21-bit
// fields).  Limb-independent samples are written to every requested limb.
#include <cstring>

#include "kbfv_internal.h"
#include "kbfv_ops.h"

namespace {

constexpr uint32_t SIG0 = 0x61707865u, SIG1 = 0x3320646eu, SIG2 = 0x79622d32u, SIG3 = 0x6b206574u;

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }

#define KB_QR(a, b, c, d)           \
    b ^= rotl32(a + d, 7);          \
    c ^= rotl32(b + a, 9);          \
    d ^= rotl32(c + b, 13);         \
    a ^= rotl32(d + c, 18)

// 10 double rounds; x in/out
__host__ __device__ __forceinline__ void salsa_rounds(uint32_t x[16]) {
    for (int r = 0; r < 10; ++r) {
        KB_QR(x[0], x[4], x[8], x[12]);
        KB_QR(x[5], x[9], x[13], x[1]);
        KB_QR(x[10], x[14], x[2], x[6]);
        KB_QR(x[15], x[3], x[7], x[11]);
        KB_QR(x[0], x[1], x[2], x[3]);
        KB_QR(x[5], x[6], x[7], x[4]);
        KB_QR(x[10], x[11], x[8], x[9]);
        KB_QR(x[15], x[12], x[13], x[14]);
    }
}

__device__ __forceinline__ void salsa_block(const kb_stream_key &k, uint64_t ctr, uint32_t out[16]) {
    uint32_t s[16] = {SIG0,     k.sub[0], k.sub[1], k.sub[2], k.sub[3], SIG1,
                      k.n8[0],  k.n8[1],  (uint32_t)ctr, (uint32_t)(ctr >> 32), SIG2,
                      k.sub[4], k.sub[5], k.sub[6], k.sub[7], SIG3};
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = s[i];
    salsa_rounds(out);
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] += s[i];
}

uint32_t ld32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

__global__ void k_keystream(kb_stream_key k, uint64_t block0, uint8_t *__restrict__ out, size_t len) {
    const long blk = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t off = (size_t)blk * 64;
    if (off >= len) return;
    uint32_t z[16];
    salsa_block(k, block0 + blk, z);
    const size_t take = len - off < 64 ? len - off : 64;
    for (size_t i = 0; i < take; ++i) out[off + i] = (uint8_t)(z[i >> 2] >> (8 * (i & 3)));
}

__device__ __forceinline__ uint64_t w64(const uint32_t *z, int i) {
    return (uint64_t)z[2 * i] | ((uint64_t)z[2 * i + 1] << 32);
}

// one Salsa20 block per thread; batch entry bt reads stream entry entry0 + bt
template <int KIND>
__global__ void k_sample(kb_dev_tables t, kb_stream_key key, uint64_t entry0,
                         uint64_t *__restrict__ out, long ostride, int mod0, int limbs, int n,
                         long blocks_per) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int bt = blockIdx.y;
    if (gid >= blocks_per) return;
    const uint64_t entry = entry0 + (uint64_t)bt;
    key.n8[0] = (uint32_t)entry;
    key.n8[1] = (uint32_t)(entry >> 32);
    uint32_t z[16];
    salsa_block(key, (uint64_t)gid, z);
    uint64_t *o = out + (long)bt * ostride;
    if constexpr (KIND == 0) {
        // 4 samples of 16 bytes; element e = gid*4 + s -> limb e / n, coefficient e % n
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const long e = gid * 4 + s;
            const int l = (int)(e / n);
            const int m = mod0 + l;
            kbfv_u128 v{w64(z, 2 * s), w64(z, 2 * s + 1)};
            o[e] = kb_barrett128(v, t.mod[m], t.mu[2 * m], t.mu[2 * m + 1]);
        }
    } else if constexpr (KIND == 2) {
#pragma unroll 4
        for (int s = 0; s < 64; ++s) {
            const long k = gid * 64 + s;
            const uint32_t byte = (z[s >> 2] >> (8 * (s & 3))) & 0xffu;
            const int v = byte < 128 ? 0 : (byte < 192 ? 1 : -1);
            for (int l = 0; l < limbs; ++l) {
                const uint64_t q = t.mod[mod0 + l];
                o[(long)l * n + k] = v < 0 ? q - 1 : (uint64_t)v;
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const long k = gid * 8 + s;
            const uint64_t x = w64(z, s);
            int v;
            if constexpr (KIND == 1) {
                v = (int)(x % 3) - 1;
            } else {
                const uint64_t M = (1ULL << 21) - 1;
                v = __popcll(x & M) - __popcll((x >> 21) & M);
            }
            for (int l = 0; l < limbs; ++l) {
                const uint64_t q = t.mod[mod0 + l];
                o[(long)l * n + k] = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
            }
        }
    }
}

}  // namespace

kb_stream_key kb_stream_setup(const uint8_t seed[32], uint64_t tag) {
    uint8_t nonce[24];
    memset(nonce, 0, sizeof(nonce));
    for (int i = 0; i < 8; ++i) nonce[i] = (uint8_t)(tag >> (8 * i));
    uint32_t s[16];
    s[0] = SIG0, s[5] = SIG1, s[10] = SIG2, s[15] = SIG3;
    for (int i = 0; i < 4; ++i) {
        s[1 + i] = ld32(seed + 4 * i);
        s[11 + i] = ld32(seed + 16 + 4 * i);
        s[6 + i] = ld32(nonce + 4 * i);
    }
    salsa_rounds(s);  // HSalsa20: no feed-forward
    kb_stream_key k;
    const int pick[8] = {0, 5, 10, 15, 6, 7, 8, 9};
    for (int i = 0; i < 8; ++i) k.sub[i] = s[pick[i]];
    k.n8[0] = ld32(nonce + 16);
    k.n8[1] = ld32(nonce + 20);
    return k;
}

int kb_xsalsa20(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0, uint8_t *out, size_t len,
                hipStream_t st) {
    KB_REQUIRE(key && nonce && out, "null argument");
    if (len == 0) return KBFV_OK;
    uint32_t s[16];
    s[0] = SIG0, s[5] = SIG1, s[10] = SIG2, s[15] = SIG3;
    for (int i = 0; i < 4; ++i) {
        s[1 + i] = ld32(key + 4 * i);
        s[11 + i] = ld32(key + 16 + 4 * i);
        s[6 + i] = ld32(nonce + 4 * i);
    }
    salsa_rounds(s);
    kb_stream_key k;
    const int pick[8] = {0, 5, 10, 15, 6, 7, 8, 9};
    for (int i = 0; i < 8; ++i) k.sub[i] = s[pick[i]];
    k.n8[0] = ld32(nonce + 16);
    k.n8[1] = ld32(nonce + 20);
    const long blocks = (long)((len + 63) / 64);
    hipLaunchKernelGGL(k_keystream, dim3((unsigned)((blocks + 255) / 256)), dim3(256), 0, st, k, block0, out, len);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_sample(const kbfv_context *c, int kind, const uint8_t seed[32], uint64_t tag, uint64_t entry0,
              int mod0, int limbs, uint64_t *out, long ostride, int batch, hipStream_t st) {
    KB_REQUIRE(c && seed && out && batch > 0 && batch <= 65535, "bad sampler arguments");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod, "bad limb range");
    KB_REQUIRE(kind >= 0 && kind <= 3, "unknown sampler");
    const kb_stream_key key = kb_stream_setup(seed, tag);
    const long n = c->n;
    if (ostride <= 0) ostride = (long)limbs * n;
    long per;
    if (kind == 0)
        per = (long)limbs * n / 4;
    else if (kind == 2)
        per = n / 64;
    else
        per = n / 8;
    const dim3 grid((unsigned)((per + 255) / 256), (unsigned)batch);
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_sample<0>, grid, dim3(256), 0, st, c->t, key, entry0, out, ostride, mod0, limbs, c->n, per); break;
        case 1: hipLaunchKernelGGL(k_sample<1>, grid, dim3(256), 0, st, c->t, key, entry0, out, ostride, mod0, limbs, c->n, per); break;
        case 2: hipLaunchKernelGGL(k_sample<2>, grid, dim3(256), 0, st, c->t, key, entry0, out, ostride, mod0, limbs, c->n, per); break;
        default: hipLaunchKernelGGL(k_sample<3>, grid, dim3(256), 0, st, c->t, key, entry0, out, ostride, mod0, limbs, c->n, per); break;
    }
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
