// Exact 60-bit x 60-bit modular dot products on the int8 matrix cores of gfx950.
//
// A residue x < 2^62 is written in balanced base 2^8:
//     x = sum_u d_u 256^u,  d_u in [-128, 127],  u = 0..7,
// whose eight signed bytes are the bytes of (x + B) ^ B with B = 0x8080808080808080
// (adding 128 to every byte lets the carries settle the balanced form in one add).
// For a dot product sum_k x_k m_k the weight-w partial sum
//     P_w = sum_k sum_{u+v=w} d_u(x_k) d_v(m_k),   w = 0..14,
// is one int8 MFMA chain when the K dimension runs over (k, u): the x operand
// holds the digit bytes of x_k as they are, and the m operand holds, per u, the
// digit v = w - u of m_k.  With r = bytereverse(digits(m_k)) that operand is just
// r shifted by 8|w - 7| bits (left for w > 7, right for w < 7), zero bytes where
// w - u leaves [0, 7].  |P_w| <= 8 K 2^14, so int32 accumulation is exact for
// K <= 2^14; the sum is rebuilt as
//     sum_w P_w 2^(8w) = (S2 2^48 + S1) 2^48 + S0,
//     S2 = P_14..P_12, S1 = P_11..P_6, S0 = P_5..P_0 (Horner in 2^8, int64),
// which is exact in 128-bit wrap-around arithmetic because the true value is a
// non-negative integer below 2^127.
//
// Constant matrices use an expanded form instead: since x = sum_u d_u 2^(8u),
//     sum_k x_k m_k = sum_(k,u) d_ku (2^(8u) m_k mod p)   (mod p),
// so K runs over (k, u) with the data digits as they are and the matrix side
// holds the balanced digits v = 0..7 of the precomputed 2^(8u) m_k mod p: 8
// MFMA passes (one per v) instead of 15 and no zero bytes.  With biased
// accumulators A_v in [0, 2^23] (KB_MX_ACC_BIAS) the sum is
//     V = sum_v A_v 2^(8v) + off,   off = (-bias sum_v 2^(8v)) mod p,
// a non-negative integer below 2^80 congruent to the dot product; the four
// 32-bit words A_(2k+1) 2^8 + A_2k give V exactly mod 2^64 and, in double
// precision, a quotient estimate that is one of floor(V/p) - 1, floor(V/p)
// for p >= 2^58, so one multiply-subtract and one conditional subtraction
// finish the reduction.
//
// MFMA v_mfma_i32_16x16x64_i8 lane layout (checked on the device): lane l holds
// row (A) / column (B) l % 16 and K bytes 16 (l / 16) .. 16 (l / 16) + 15; the
// result lane l holds rows 4 (l / 16) + i, i = 0..3, of column l % 16.  Here a
// lane's 16 K bytes are the digits of two consecutive k.
#pragma once
#include <cstdint>

typedef int kb_v4i __attribute__((ext_vector_type(4)));

#define KB_DIGIT_BIAS 0x8080808080808080ULL

// Expanded-matrix products start each digit pass's int32 accumulator at this
// bias: a pass sums at most 256 products of two digits in [-128, 127], so the
// biased sums lie in [0, 2^23] and pairs of them combine as unsigned 32-bit
// words.  The table's per-row offset cancels the bias modulo p.
#define KB_MX_ACC_BIAS (1 << 22)
// words per output row after the expanded matrix
#define KB_MX_ROW 6

__device__ __forceinline__ uint64_t kb_digits(uint64_t x) { return (x + KB_DIGIT_BIAS) ^ KB_DIGIT_BIAS; }

__device__ __forceinline__ kb_v4i kb_pack2(uint64_t a, uint64_t b) {
    kb_v4i r;
    r.x = (int)(uint32_t)a;
    r.y = (int)(uint32_t)(a >> 32);
    r.z = (int)(uint32_t)b;
    r.w = (int)(uint32_t)(b >> 32);
    return r;
}

// operand of weight w from the byte-reversed digits r (w is a constant after unrolling)
__device__ __forceinline__ uint64_t kb_weight_op(uint64_t r, int w) {
    return w >= 7 ? r << (8 * (w - 7)) : r >> (8 * (7 - w));
}

__device__ __forceinline__ kb_v4i kb_mfma_i8(kb_v4i a, kb_v4i b, kb_v4i c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// (hi:lo) = (hi:lo) * 2^48 + s  (s signed)
__device__ __forceinline__ void kb_horner48(uint64_t &lo, uint64_t &hi, int64_t s) {
    hi = (hi << 48) | (lo >> 16);
    lo <<= 48;
    const uint64_t nl = lo + (uint64_t)s;
    hi += (uint64_t)(nl < lo) + (s < 0 ? ~0ULL : 0ULL);
    lo = nl;
}

