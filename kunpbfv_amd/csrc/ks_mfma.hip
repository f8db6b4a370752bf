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
// A workgroup takes 8 consecutive coefficients of one T limb, one wave per
// coefficient.  The key of those coefficients is staged through LDS once and
// kept in registers as byte-reversed digits; digit tiles of 32 ciphertexts are
// then streamed through LDS (coalesced 64-byte rows in, transposed so a lane
// reads its 16 operand bytes at once) and the outputs leave through LDS the
// same way.  The key is read from HBM once per call, the digits and outputs
// once each.
#include "kbfv_internal.h"
#include "kbfv_mfma.h"
#include "kbfv_ops.h"

namespace {

constexpr int IM_POS = 8;                    // coefficients per workgroup (one per wave)
constexpr int IM_THREADS = 64 * IM_POS;      // 512
constexpr int IM_CT = 32;                    // ciphertexts per tile (2 MFMA column tiles)
constexpr int IM_ROWS = 32;                  // key rows (h, b) held: 2 x ri_blocks <= 32
constexpr int IM_KP = 32;                    // K (digits) held: dnum <= 32
constexpr int IM_RS = IM_KP + 2;             // LDS row stride: 16 lanes' 16-byte reads hit distinct banks
constexpr int IM_PS = IM_ROWS * IM_RS + 2;   // LDS stride between coefficients, [pos][row][d]
constexpr int IM_CS = IM_ROWS * IM_POS + 1;  // LDS stride between ciphertexts, [ct][row][pos]
constexpr int IM_LDS = (IM_POS * IM_PS > IM_CT * IM_CS) ? IM_POS * IM_PS : IM_CT * IM_CS;
// per-thread share of a digit tile (IM_CT x IM_KP x IM_POS values)
constexpr int IM_PER = IM_CT * IM_KP * IM_POS / IM_THREADS;  // 16

// KC: 8-digit chunks (dnum <= 8 KC); NJ: 16-row tiles of the 2 x ri_blocks key rows
template <int KC, int NJ>
__global__ void __launch_bounds__(IM_THREADS, 2)
    k_ip_mfma(const uint64_t *__restrict__ cin, const uint64_t *__restrict__ key, uint64_t *__restrict__ out,
              const uint64_t *__restrict__ mods, const uint64_t *__restrict__ mus, int D, int B, int tn,
              int oT, int n, int batch) {
    __shared__ uint64_t lds[IM_LDS];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;  // = this wave's coefficient within the tile
    const int p16 = lane & 15, g = lane >> 4;
    const int tiles = n / IM_POS;
    const int kk = blockIdx.x / tiles;
    const int n0 = (blockIdx.x % tiles) * IM_POS;
    const long N = n;
    const int R2 = 2 * B;
    const uint64_t q = mods[oT + kk], mlo = mus[2 * (oT + kk)], mhi = mus[2 * (oT + kk) + 1];

    // ---- key of the 8 coefficients -> LDS [pos][row][d] -> registers (reversed digits)
    // key layout [2][D][B][tn][N]: row r = h*B + b
    for (int e = tid; e < R2 * D * IM_POS; e += IM_THREADS) {
        const int pos = e % IM_POS;
        const int rest = e / IM_POS;
        const int b = rest % B;
        const int hd = rest / B;  // h*D + d
        const int d = hd % D, h = hd / D;
        const uint64_t v = key[((long)hd * B + b) * tn * N + (long)kk * N + n0 + pos];
        lds[pos * IM_PS + (h * B + b) * IM_RS + d] = v;
    }
    __syncthreads();
    uint64_t rk[NJ][KC][2];
#pragma unroll
    for (int nt = 0; nt < NJ; ++nt)
#pragma unroll
        for (int c = 0; c < KC; ++c)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int r = 16 * nt + p16, d = 8 * c + 2 * g + e;
                const uint64_t v = (r < R2 && d < D) ? lds[wave * IM_PS + r * IM_RS + d] : 0;
                rk[nt][c][e] = __builtin_bswap64(kb_digits(v));
            }

    // ---- digit tiles: cin layout [batch][D][tn][N]
    const int ntiles = (batch + IM_CT - 1) / IM_CT;
    uint64_t pre[IM_PER];
    auto fetch = [&](int tile) {
#pragma unroll
        for (int r = 0; r < IM_PER; ++r) {
            const int e = tid + r * IM_THREADS;  // (ct, d, pos), pos fastest
            const int pos = e % IM_POS;
            const int d = (e / IM_POS) % IM_KP;
            const int ct = tile * IM_CT + e / (IM_POS * IM_KP);
            const bool live = d < D && ct < batch;  // branch-free: clamp, load, then select
            const uint64_t v = cin[(((long)(live ? ct : 0) * D + (live ? d : 0)) * tn + kk) * N + n0 + pos];
            pre[r] = live ? v : 0;
        }
    };
    fetch(0);
    for (int tile = 0; tile < ntiles; ++tile) {
        __syncthreads();  // the previous tile's outputs have left LDS
#pragma unroll
        for (int r = 0; r < IM_PER; ++r) {
            const int e = tid + r * IM_THREADS;
            const int pos = e % IM_POS;
            const int d = (e / IM_POS) % IM_KP;
            const int ctl = e / (IM_POS * IM_KP);
            lds[pos * IM_PS + ctl * IM_RS + d] = pre[r];
        }
        __syncthreads();
        kb_v4i bop[2][KC];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const uint64_t *src = lds + wave * IM_PS + (16 * mt + p16) * IM_RS + 8 * c + 2 * g;
                bop[mt][c] = kb_pack2(kb_digits(src[0]), kb_digits(src[1]));
            }
        __syncthreads();  // the tile's digits are in registers; LDS now takes outputs
        if (tile + 1 < ntiles) fetch(tile + 1);

        // the key operands stay loop-carried: without this the compiler hoists
        // all 15 shifted copies out of the tile loop and runs out of registers
#pragma unroll
        for (int nt = 0; nt < NJ; ++nt)
#pragma unroll
            for (int c = 0; c < KC; ++c) asm volatile("" : "+v"(rk[nt][c][0]), "+v"(rk[nt][c][1]));
#pragma unroll
        for (int nt = 0; nt < NJ; ++nt) {
            int64_t S[2][4];
            uint64_t Hl[2][4], Hh[2][4];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) S[mt][i] = 0, Hl[mt][i] = 0, Hh[mt][i] = 0;
#pragma unroll
            for (int w = 14; w >= 0; --w) {
                kb_v4i acc[2] = {kb_v4i{0, 0, 0, 0}, kb_v4i{0, 0, 0, 0}};
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    const kb_v4i a = kb_pack2(kb_weight_op(rk[nt][c][0], w), kb_weight_op(rk[nt][c][1], w));
                    acc[0] = kb_mfma_i8(a, bop[0][c], acc[0]);
                    acc[1] = kb_mfma_i8(a, bop[1][c], acc[1]);
                }
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        S[mt][i] = (int64_t)((uint64_t)S[mt][i] << 8) + (int64_t)acc[mt][i];
                        if (w == 12 || w == 6 || w == 0) {
                            kb_horner48(Hl[mt][i], Hh[mt][i], S[mt][i]);
                            S[mt][i] = 0;
                        }
                    }
                // keep each weight pass together (no hoisting of later passes' operands)
                __builtin_amdgcn_sched_barrier(0);
            }
            // result lane: ciphertext 16 mt + p16, key rows 16 nt + 4 g + i
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = 16 * nt + 4 * g + i;
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
                    lds[(16 * mt + p16) * IM_CS + r * IM_POS + wave] =
                        kb_barrett128(kbfv_u128{Hl[mt][i], Hh[mt][i]}, q, mlo, mhi);
            }
        }
        __syncthreads();
        // outputs [batch][2][B][tn][N]: row (ct, r) -> 8 consecutive coefficients
        for (int e = tid; e < IM_CT * R2 * IM_POS; e += IM_THREADS) {
            const int pos = e % IM_POS;
            const int r = (e / IM_POS) % R2;
            const int ctl = e / (IM_POS * R2);
            const int ct = tile * IM_CT + ctl;
            if (ct < batch)
                out[(((long)ct * R2 + r) * tn + kk) * N + n0 + pos] = lds[ctl * IM_CS + r * IM_POS + pos];
        }
    }
}

}  // namespace

int kb_inner_product_mfma(const kbfv_context *c, const uint64_t *cin, const uint64_t *key, uint64_t *out,
                          int batch, hipStream_t st) {
    const int D = c->dnum, B = c->ri_blocks;
    if (D > IM_KP || 2 * B > IM_ROWS || c->n % IM_POS) return 1;
    const int kc = (D + 7) / 8, nj = (2 * B + 15) / 16;
    const dim3 grid((unsigned)(c->tn * (c->n / IM_POS)));
#define KB_IM(KC, NJ)                                                                                   \
    if (kc == KC && nj == NJ) {                                                                          \
        hipLaunchKernelGGL((k_ip_mfma<KC, NJ>), grid, dim3(IM_THREADS), 0, st, cin, key, out, c->t.mod, \
                           c->t.mu, D, B, c->tn, c->oT, c->n, batch);                                    \
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
