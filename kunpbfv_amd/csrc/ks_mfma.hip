// Key-switch inner product on the int8 matrix cores.
//
// Replaces external_product_T_kernel (include/ExternalProduct.cuh:11-112) for
// batches of ciphertexts: for T limb k and coefficient n,
//     out[ct][h][b] = sum_{d < D} cin[ct][d] key[h][d][b]   mod t_k,
// one small product per coefficient (rows (h, b): 2 x ri_blocks, K = dnum,
// columns: the ciphertexts of the batch), exact through the digit scheme of
// kbfv_mfma.h and one Barrett reduction per output, so the result equals the
// scalar kernel's bit for bit.
//
// One wave per coefficient at no more than 128 VGPRs: a workgroup's 8 waves
// take 8 consecutive coefficients of one T limb, two workgroups share each CU,
// and the two workgroups that cover one 128-byte row of 16 coefficients are
// dispatched to the same XCD.  A wave keeps its coefficient's key rows as
// byte-reversed digits (kbfv_mfma.h, shifted per weight pass) and streams the
// ciphertexts through in tiles of 16: 15 weight passes of MFMAs into biased
// accumulators (every P_w + 2^22 lies in [0, 2^23]), paired into 32-bit words
// on the fly, then per output the 128-bit sum rebuilt from its 8 words, the
// bias removed, and one Barrett reduction.
#include "kbfv_internal.h"
#include "kbfv_mfma.h"
#include "kbfv_ops.h"

namespace {

constexpr int IM_WAVES = 8;               // coefficients per workgroup (one per wave)
constexpr int IM_THREADS = 64 * IM_WAVES;  // 512
constexpr int IM_CT = 16;                 // ciphertexts per tile (one MFMA column tile)
// output staging [r][ct][pos] in LDS: pos stride 1, an odd ct stride and a
// row stride of 4 mod 8 keep the per-wave writes and the row-wise reads
// (nearly) free of bank conflicts
constexpr int IM_OC = IM_WAVES + 1, IM_OR = 16 * IM_OC + 4, IM_OBUF = 32 * IM_OR;

// Raw buffer access: a resource (base in scalar registers) plus a 32-bit byte
// offset per lane, so an address costs one VGPR.  Every offset used here stays
// inside its allocation by construction (clamped indices); num_records is left
// at its maximum.
typedef unsigned int kb_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t kb_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ uint64_t kb_bload(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    const kb_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
    return ((uint64_t)v.y << 32) | v.x;
}
__device__ __forceinline__ void kb_bstore(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, uint64_t x) {
    kb_u32x2 v;
    v.x = (uint32_t)x;
    v.y = (uint32_t)(x >> 32);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, byte_off, 0, 0);
}

// sum_{w < 15} KB_MX_ACC_BIAS 2^(8w) mod 2^128: what the biased accumulators add
constexpr unsigned __int128 ip_bias_sum() {
    unsigned __int128 s = 0;
    for (int w = 0; w < 15; ++w) s += (unsigned __int128)KB_MX_ACC_BIAS << (8 * w);
    return s;
}
constexpr uint64_t IM_BIAS_LO = (uint64_t)ip_bias_sum();
constexpr uint64_t IM_BIAS_HI = (uint64_t)(ip_bias_sum() >> 64);

// KC: 8-digit chunks (dnum <= 8 KC); NJ: 16-row tiles of the 2 x ri_blocks key rows
template <int KC, int NJ>
__global__ void __launch_bounds__(IM_THREADS, 4)
    k_ip_mfma(const uint64_t *__restrict__ cin, const uint64_t *__restrict__ key, uint64_t *__restrict__ out,
              const uint64_t *__restrict__ mods, const uint64_t *__restrict__ mus, int D, int B, int KD,
              int KB, int tn, int oT, int n, int batch) {
    const int lane = threadIdx.x & 63;
    const int p16 = lane & 15, g = lane >> 4;
    const int groups = n / IM_WAVES;
    // XCD-aware order: workgroups b and b + 8 (dispatched to the same XCD) take
    // the two halves of one 128-byte row of 16 coefficients, so each line is
    // fetched into one L2 only
    const int gidx = (blockIdx.x / 16) * 16 + (blockIdx.x % 8) * 2 + (blockIdx.x / 8) % 2;
    const int kk = gidx / groups;
    // this wave's coefficient (wave-uniform, so the base pointers below live in
    // scalar registers and every lane address is a 32-bit offset)
    const int pos = (gidx % groups) * IM_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long N = n, TN = (long)tn * n;
    const uint32_t tn32 = (uint32_t)TN;  // 8 TN * 16 * 32 fits 31 bits (checked by the launcher)
    const int R2 = 2 * B;
    const uint64_t q = mods[oT + kk], mlo = mus[2 * (oT + kk)], mhi = mus[2 * (oT + kk) + 1];

    // key layout [2][KD][KB][tn][N] (its first D digits and B blocks are used),
    // row r = h B + b; lane holds rows 16 nt + p16, digits d = 8c + 2g + e
    // (zero outside 2B x D)
    uint64_t rk[NJ][KC][2];
    {
        const uint64_t *kb = key + kk * N + pos;
#pragma unroll
        for (int nt = 0; nt < NJ; ++nt)
#pragma unroll
            for (int c = 0; c < KC; ++c)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int r = 16 * nt + p16, d = 8 * c + 2 * g + e;
                    const bool live = r < R2 && d < D;
                    const int rr = live ? r : 0, dd = live ? d : 0;
                    const int h = rr / B, b = rr - h * B;
                    const uint64_t v = kb[(size_t)(uint32_t)((h * KD + dd) * KB + b) * TN];
                    rk[nt][c][e] = __builtin_bswap64(kb_digits(live ? v : 0));
                }
    }

    // ciphertext digits, cin layout [batch][D][tn][N], staged per tile of 16
    // ciphertexts through LDS as [ct][d][pos ^ (ct mod 8)]: 64-byte rows of 8
    // coefficients come in, and the lanes of a wave that read one pos (its
    // coefficient) for 16 ciphertexts spread over the banks
    const int ntiles = (batch + IM_CT - 1) / IM_CT;
    __shared__ uint64_t ibuf[IM_CT * 32 * IM_WAVES];
    __shared__ uint64_t obuf[IM_OBUF];
    const int wave = pos % IM_WAVES;
    for (int tile = 0; tile < ntiles; ++tile) {
        {
            int t = threadIdx.x;
            asm volatile("" : "+v"(t));  // recomputed per tile rather than held
            const int pp = t % IM_WAVES;
            const __amdgpu_buffer_rsrc_t xb = kb_rsrc(cin + (long)tile * IM_CT * D * TN + kk * N + (pos - wave));
            // rows j = 32 ct + d; rows d >= D stay unwritten (they meet zero key
            // digits) and ciphertexts past the batch repeat the last one
            for (int j = t / IM_WAVES; j < IM_CT * 32; j += IM_THREADS / IM_WAVES) {
                const int ctl = j >> 5, d = j & 31;
                if (d < D) {
                    const int cl = min(ctl, batch - 1 - tile * IM_CT);
                    ibuf[j * IM_WAVES + (pp ^ (ctl % IM_WAVES))] =
                        kb_bload(xb, ((uint32_t)(cl * D + d) * tn32 + pp) * 8u);
                }
            }
        }
        __syncthreads();
        kb_v4i bop[KC];
        {
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int p16 = ln & 15, g = ln >> 4;
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const uint64_t *src = ibuf + (p16 * 32 + 8 * c + 2 * g) * IM_WAVES + (wave ^ (p16 % IM_WAVES));
                bop[c] = kb_pack2(kb_digits(src[0]), kb_digits(src[IM_WAVES]));
            }
        }
        uint64_t *const ob = obuf;

#pragma unroll
        for (int nt = 0; nt < NJ; ++nt) {
            // W[i][k] = P'_(2k+1) 2^8 + P'_2k (k < 7), W[i][7] = P'_14
            uint32_t W[4][8];
#pragma unroll
            for (int w = 14; w >= 0; --w) {
                kb_v4i acc = kb_v4i{KB_MX_ACC_BIAS, KB_MX_ACC_BIAS, KB_MX_ACC_BIAS, KB_MX_ACC_BIAS};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    // shift here, not earlier: hoisted shifted copies exhaust registers
                    asm volatile("" : "+v"(rk[nt][c][0]), "+v"(rk[nt][c][1]));
                    const kb_v4i a = kb_pack2(kb_weight_op(rk[nt][c][0], w), kb_weight_op(rk[nt][c][1], w));
                    acc = kb_mfma_i8(a, bop[c], acc);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (w == 14 || (w & 1))
                        W[i][w >> 1] = (uint32_t)acc[i];
                    else
                        W[i][w >> 1] = (W[i][w >> 1] << 8) + (uint32_t)acc[i];
                    // pin the word here: sunk to the epilogue it would keep all 15
                    // accumulators alive
                    asm volatile("" : "+v"(W[i][w >> 1]));
                }
                // keep each weight pass together (no hoisting of later passes' operands)
                __builtin_amdgcn_sched_barrier(0);
            }
            // result lane: ciphertext ct, key rows 16 nt + 4 g + i
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int p16 = ln & 15, g = ln >> 4;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = 16 * nt + 4 * g + i;
                // V' = sum_k W_k 2^(16k) mod 2^128, then V = V' - bias (exact: 0 <= V < 2^127)
                const uint64_t a0 = (uint64_t)W[i][0] + ((uint64_t)W[i][1] << 16);
                const uint64_t a1 = (uint64_t)W[i][2] + ((uint64_t)W[i][3] << 16);
                const uint64_t a2 = (uint64_t)W[i][4] + ((uint64_t)W[i][5] << 16);
                const uint64_t a3 = (uint64_t)W[i][6] + ((uint64_t)W[i][7] << 16);
                const uint64_t lo = a0 + (a1 << 32);
                uint64_t hi = (a1 >> 32) + a2 + (a3 << 32) + (uint64_t)(lo < a0);
                hi -= IM_BIAS_HI + (uint64_t)(lo < IM_BIAS_LO);
                const uint64_t v = kb_barrett128(kbfv_u128{lo - IM_BIAS_LO, hi}, q, mlo, mhi);
                ob[r * IM_OR + p16 * IM_OC + wave] = v;
                // one output at a time (interleaving them costs registers, not time)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // the tile's outputs leave in 64-byte rows (8 coefficients).  Both
        // buffers are reused by the next tile only after its first barrier, by
        // which every thread has finished reading them here and above
        __syncthreads();
        {
            int t = threadIdx.x;
            asm volatile("" : "+v"(t));  // recomputed per tile rather than held
            const int pp = t % IM_WAVES, ctl = (t / IM_WAVES) % IM_CT;
            const int ctg = tile * IM_CT + ctl;
            const __amdgpu_buffer_rsrc_t og = kb_rsrc(out + (long)tile * IM_CT * R2 * TN + kk * N + (pos - wave));
            for (int r = t / (IM_WAVES * IM_CT); r < R2; r += IM_THREADS / (IM_WAVES * IM_CT))
                if (ctg < batch)
                    kb_bstore(og, (uint32_t)(ctl * R2 + r) * tn32 * 8u + pp * 8u, ob[r * IM_OR + ctl * IM_OC + pp]);
        }
    }
}

}  // namespace

int kb_inner_product_mfma(const kbfv_context *c, const uint64_t *cin, const uint64_t *key, uint64_t *out,
                          int batch, hipStream_t st) {
    const int D = c->dnum, B = c->ri_blocks;
    if (D > 32 || 2 * B > 32 || (c->tn * c->n / IM_WAVES) % 16) return 1;
    if ((long)c->tn * c->n * 8 * IM_CT * 32 >= (1L << 31)) return 1;  // 32-bit byte offsets
    const int kc = (D + 7) / 8, nj = (2 * B + 15) / 16;
    const dim3 grid((unsigned)(c->tn * (c->n / IM_WAVES)));
#define KB_IM(KC, NJ)                                                                                   \
    if (kc == KC && nj == NJ) {                                                                          \
        hipLaunchKernelGGL((k_ip_mfma<KC, NJ>), grid, dim3(IM_THREADS), 0, st, cin, key, out, c->t.mod, \
                           c->t.mu, D, B, c->key_dnum, c->key_blocks, c->tn, c->oT, c->n, batch);        \
        KB_LAUNCH_CHECK();                                                                               \
        return KBFV_OK;                                                                                  \
    }
    KB_IM(1, 1)
    KB_IM(2, 1)
    KB_IM(3, 1)
    KB_IM(4, 1)
    KB_IM(1, 2)
    KB_IM(2, 2)
    KB_IM(3, 2)
    KB_IM(4, 2)
#undef KB_IM
    return 1;
}
