// Device-side modular arithmetic for gfx950 (CDNA4, wave64).
//
// All residues are uint64 below 2^60 (the 60-bit chains and the plain
// modulus), so lazy values up to 4q fit in 62 bits.  CDNA4 has no 64x64
// multiply; the compiler expands 64-bit products into 32-bit
// v_mad_u64_u32 / v_mul_hi_u32 / v_mul_lo_u32 sequences.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct __align__(16) kbfv_tw {
    uint64_t w;   // twiddle
    uint64_t ws;  // Shoup companion floor(w * 2^64 / q)
};

struct kbfv_u128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ uint64_t kb_mulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// a mod q for a in [0, 2q)
__device__ __forceinline__ uint64_t kb_csub(uint64_t a, uint64_t q) {
    uint64_t t = a - q;
    return t < a ? t : a;  // a >= q  <=>  a - q does not wrap
}

// Shoup product x*w mod q, lazily in [0, 2q); valid for any 64-bit x.
__device__ __forceinline__ uint64_t kb_shoup_lazy(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    return x * w - kb_mulhi(x, ws) * q;
}
__device__ __forceinline__ uint64_t kb_shoup(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    return kb_csub(kb_shoup_lazy(x, w, ws, q), q);
}

// full 64x64 -> 128 product
__device__ __forceinline__ kbfv_u128 kb_mul128(uint64_t a, uint64_t b) {
    kbfv_u128 r;
    r.lo = a * b;
    r.hi = kb_mulhi(a, b);
    return r;
}
__device__ __forceinline__ void kb_mac128(kbfv_u128 &acc, uint64_t a, uint64_t b) {
    uint64_t lo = a * b;
    uint64_t hi = kb_mulhi(a, b);
    acc.lo += lo;
    acc.hi += hi + (acc.lo < lo ? 1u : 0u);
}
__device__ __forceinline__ void kb_add128(kbfv_u128 &acc, uint64_t v) {
    acc.lo += v;
    acc.hi += (acc.lo < v ? 1u : 0u);
}
// 128-bit wrap-around subtraction of a 64-bit value
__device__ __forceinline__ void kb_sub128(kbfv_u128 &acc, uint64_t v) {
    uint64_t borrow = acc.lo < v ? 1u : 0u;
    acc.lo -= v;
    acc.hi -= borrow;
}

// Barrett reduction of any 128-bit value with mu = floor(2^128 / q) = (mu_hi, mu_lo).
// The quotient estimate floor(a * mu / 2^128) is at least floor(a / q) - 1, so a
// single conditional subtraction yields the canonical residue in [0, q).
__device__ __forceinline__ uint64_t kb_barrett128(kbfv_u128 a, uint64_t q, uint64_t mu_lo,
                                                  uint64_t mu_hi) {
    uint64_t t = kb_mulhi(a.lo, mu_lo);
    uint64_t m1lo = a.lo * mu_hi, m1hi = kb_mulhi(a.lo, mu_hi);
    uint64_t m2lo = a.hi * mu_lo, m2hi = kb_mulhi(a.hi, mu_lo);
    uint64_t s = t + m1lo;
    uint64_t c1 = s < t ? 1u : 0u;
    uint64_t s2 = s + m2lo;
    uint64_t c2 = s2 < s ? 1u : 0u;
    uint64_t qe = a.hi * mu_hi + m1hi + m2hi + c1 + c2;
    uint64_t r = a.lo - qe * q;
    return kb_csub(r, q);
}

// canonical a*b mod q via Barrett
__device__ __forceinline__ uint64_t kb_mulmod(uint64_t a, uint64_t b, uint64_t q, uint64_t mu_lo,
                                              uint64_t mu_hi) {
    return kb_barrett128(kb_mul128(a, b), q, mu_lo, mu_hi);
}

__device__ __forceinline__ uint64_t kb_addmod(uint64_t a, uint64_t b, uint64_t q) {
    return kb_csub(a + b, q);
}
__device__ __forceinline__ uint64_t kb_submod(uint64_t a, uint64_t b, uint64_t q) {
    return kb_csub(a + q - b, q);
}

// double -> uint64 truncation saturating at 2^64 - 1 for non-negative x (the
// conversion the reference's nvcc build performs for static_cast<uint64_t>)
__device__ __forceinline__ uint64_t kb_f64_to_u64_sat(double x) {
    return x >= 18446744073709551616.0 ? ~0ULL : (uint64_t)x;
}
// uint64 -> double with one rounding (error at most 2^11)
__device__ __forceinline__ double kb_u64_to_f64(uint64_t x) {
    return fma((double)(uint32_t)(x >> 32), 4294967296.0, (double)(uint32_t)x);
}

// Shoup lazy product from 32-bit halves: r = x*w - floor(x*ws / 2^64) * q in
// [0, 2q) for any x < 2^64, with nq = 2^64 - q so that both low products
// accumulate in one 64-bit chain (x*w + h*nq = x*w - h*q mod 2^64).
__device__ __forceinline__ uint64_t kb_mad32(uint32_t a, uint32_t b, uint64_t c) {
    return (uint64_t)a * b + c;
}
// a*b mod 2^32 as one opaque v_mul_lo_u32.  Written as plain C, the compiler
// folds the cross products below back into 64x64-bit multiplies and expands
// those again with extra v_mad_u64_u32 steps (about twice the instructions).
__device__ __forceinline__ uint32_t kb_mul_lo32(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_lo_u32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint64_t kb_shoup_lazy_nq(uint64_t x, uint64_t w, uint64_t ws, uint64_t nq) {
    const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
    const uint32_t s0 = (uint32_t)ws, s1 = (uint32_t)(ws >> 32);
    const uint64_t t = kb_mad32(x1, s0, __umulhi(x0, s0));
    const uint64_t u = kb_mad32(x0, s1, (uint32_t)t);
    const uint64_t h = kb_mad32(x1, s1, (t >> 32) + (u >> 32));  // floor(x*ws / 2^64)
    const uint32_t h0 = (uint32_t)h, h1 = (uint32_t)(h >> 32);
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
    const uint32_t n0 = (uint32_t)nq, n1 = (uint32_t)(nq >> 32);
    const uint64_t r = kb_mad32(h0, n0, kb_mad32(x0, w0, 0));
    const uint32_t cross = kb_mul_lo32(x0, w1) + kb_mul_lo32(x1, w0) + kb_mul_lo32(h0, n1) +
                           kb_mul_lo32(h1, n0);
    return r + ((uint64_t)cross << 32);
}
// a >= b ? a - b : a, for a, b < 2^62 (the sign of the difference decides)
__device__ __forceinline__ uint64_t kb_csub_s(uint64_t a, uint64_t b) {
    const uint64_t t = a - b;
    return (int64_t)t < 0 ? a : t;
}
