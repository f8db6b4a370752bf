// HPS base conversion and scale-and-round as int8 MFMA products.
//
// Both are (coefficient x constant matrix) products with exact 128-bit sums
// before one Barrett reduction per output (include/poly_arithmetic.cuh:1139-1178
// base_convert_matmul_hps_unroll2_kernel, :1263-1317 scaleAndRound_HPS_QR_R).
// The product runs transposed, out^T[j][p] = sum_i M^T[j][i] X^T[i][p], so the
// matrix is the MFMA A operand (rows = output limbs j) and the coefficients the
// B operand (columns = positions p); a result lane then holds 16 consecutive
// positions of 4 output limbs and its stores are 128-byte runs.
//
// The matrix enters in expanded form (kbfv_mfma.h): x_i = sum_u d_iu 2^(8u)
// gives sum_i x_i M_ij = sum_(i,u) d_iu (2^(8u) M_ij mod p_j)  (mod p_j), so
// K runs over (i, u) and the products split into 8 digit passes of the
// expanded entries -- 8 MFMA passes and 8 int32 partial sums per output.
//
// Base conversion Q -> R (and R -> Q):  y_i = [x_i (Q/q_i)^-1]_{q_i};
//   alpha = llround(sum_i y_i / q_i)   (double, input order, explicit fma);
//   out_j = [sum_i y_i [Q/q_i]_{r_j} + alpha [-Q]_{r_j}]_{r_j}
// with alpha riding in the product as one more K row.
// Scale-and-round QR -> R:  nu = 0.5 + sum_{i<qn} x_i frac_i (double);
//   out_j = [sum_{i<qn} x_i M_ij + x_{qn+j} c_j + (uint64)nu]_{r_j}.
// The double estimates are the reference's, in its order; every other step is
// exact, so outputs equal the scalar kernels' bit for bit.
#include "kbfv_internal.h"
#include "kbfv_mfma.h"
#include "kbfv_ops.h"

namespace {

constexpr int RG_MT = 4;                 // 16-position tiles per wave
constexpr int RG_POS = 16 * RG_MT;       // positions per wave (one per lane in the estimate)
constexpr int RG_WAVES = 1;              // waves per workgroup
constexpr int RG_BLOCK = RG_POS * RG_WAVES;
constexpr int RG_KMAX = 32;              // K rows of the largest matrix
static_assert(RG_KMAX * RG_POS * 2 >= 4 * 4 * RG_MT * 64, "W words must fit the wave's input tile");

// KC: 8-row chunks of K (K <= 8 KC);  NJ: 16-row tiles of output limbs.
// SR = false: base conversion (Shoup-scaled inputs, alpha row at K index in_n);
// SR = true: scale-and-round (raw Q inputs, R-limb diagonal term and nu added after).
// mat: [16 NJ][8][8 KC] packed digits of 2^(8u) M_ij mod p_j, then [16 NJ][2]
// offsets (the expanded-matrix form of kbfv_mfma.h).
template <int KC, int NJ, bool SR>
__global__ void __launch_bounds__(64 * RG_WAVES, 2)
    k_rns_gemm(const uint64_t *__restrict__ x, long in_stride, uint64_t *__restrict__ y, long out_stride,
               const uint64_t *__restrict__ mat, const uint64_t *__restrict__ mods_out,
               const kbfv_tw *__restrict__ hatinv, const uint64_t *__restrict__ mods_in,
               const double *__restrict__ est, int in_n, int out_n, int n) {
    constexpr int KP = 8 * KC, JP = 16 * NJ;
    // ys: the wave's inputs for the estimate, then (reused) its 32-bit words W
    __shared__ uint64_t ys[RG_WAVES][RG_KMAX][RG_POS];
    __shared__ uint64_t alpha_s[RG_WAVES][RG_POS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t *const wl = (uint32_t *)&ys[wave][0][0];  // W words [k][i][mt][lane]
    const int p16 = lane & 15, g = lane >> 4;
    const int tiles = n / RG_BLOCK;
    const long poly = blockIdx.x / tiles;
    const int n0 = (blockIdx.x % tiles) * RG_BLOCK + wave * RG_POS;
    const uint64_t *xp = x + poly * in_stride + n0;
    uint64_t *yp = y + poly * out_stride + n0;

    // inputs: lane (position p16 of tile mt, group g) holds K rows 8c + 2g + e
    // all loads first (rows past in_n read row 0 and are zeroed afterwards), so
    // that every load is in flight before the first use
    uint64_t yv[RG_MT][KC][2];
    kbfv_tw hv[KC][2];
    uint64_t qv[KC][2];
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int i = 8 * c + 2 * g + e;
            const int ii = i < in_n ? i : 0;
#pragma unroll
            for (int mt = 0; mt < RG_MT; ++mt) yv[mt][c][e] = xp[(long)ii * n + 16 * mt + p16];
            if (!SR) hv[c][e] = hatinv[ii], qv[c][e] = mods_in[ii];
        }
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const bool live = 8 * c + 2 * g + e < in_n;
#pragma unroll
            for (int mt = 0; mt < RG_MT; ++mt) {
                uint64_t v = yv[mt][c][e];
                if (!SR) v = kb_shoup(v, hv[c][e].w, hv[c][e].ws, qv[c][e]);
                yv[mt][c][e] = live ? v : 0;
            }
        }
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int i = 8 * c + 2 * g + e;
            if (i < in_n) {
#pragma unroll
                for (int mt = 0; mt < RG_MT; ++mt) ys[wave][i][16 * mt + p16] = yv[mt][c][e];
            }
        }
    __syncthreads();
    // the reference's double estimate, one position per lane, in input order
    {
        double acc = SR ? 0.5 : 0.0;
        for (int i = 0; i < in_n; ++i) acc = fma((double)ys[wave][i][lane], est[i], acc);
        alpha_s[wave][lane] = SR ? kb_f64_to_u64_sat(acc) : (uint64_t)(int)llround(acc);
    }
    __syncthreads();

    kb_v4i bop[RG_MT][KC];
#pragma unroll
    for (int mt = 0; mt < RG_MT; ++mt) {
        const uint64_t al = alpha_s[wave][16 * mt + p16];
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            uint64_t d0 = yv[mt][c][0], d1 = yv[mt][c][1];
            if (!SR) {
                if (8 * c + 2 * g == in_n) d0 = al;
                if (8 * c + 2 * g + 1 == in_n) d1 = al;
            }
            bop[mt][c] = kb_pack2(kb_digits(d0), kb_digits(d1));
        }
    }
    const uint64_t *rows = mat + (long)JP * 8 * KP;
    // one 16-limb output tile at a time: 8 digit passes of the expanded matrix
    // from biased accumulators, paired into the 32-bit words W3..W0, then the
    // double-precision quotient estimate, one multiply-subtract and stores
#pragma unroll 1
    for (int nt = 0; nt < NJ; ++nt) {
        const kb_v4i *mrow = (const kb_v4i *)(mat + (long)(16 * nt + p16) * 8 * KP + 8 * 0 + 2 * g);
        // W[k] = A_(2k+1) 2^8 + A_2k goes to LDS as soon as its pair of passes is
        // done, which keeps the pass loop's registers to the operands
        uint32_t W[RG_MT][4];
        // E[j][v][8c + 2g .. +1]: 16 bytes, row stride 8 KP words, v stride KP
        // words; the next pass's operands load while this pass multiplies, and
        // the scheduling fences keep deeper prefetch from exhausting registers
        kb_v4i mcur[KC], mnxt[KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) mcur[c] = mrow[(7 * KP + 8 * c) / 2];
#pragma unroll
        for (int v = 7; v >= 0; --v) {
            if (v > 0) {
#pragma unroll
                for (int c = 0; c < KC; ++c) mnxt[c] = mrow[((v - 1) * KP + 8 * c) / 2];
            }
            kb_v4i acc[RG_MT];
#pragma unroll
            for (int mt = 0; mt < RG_MT; ++mt)
                acc[mt] = kb_v4i{KB_MX_ACC_BIAS, KB_MX_ACC_BIAS, KB_MX_ACC_BIAS, KB_MX_ACC_BIAS};
#pragma unroll
            for (int c = 0; c < KC; ++c)
#pragma unroll
                for (int mt = 0; mt < RG_MT; ++mt) acc[mt] = kb_mfma_i8(mcur[c], bop[mt][c], acc[mt]);
#pragma unroll
            for (int mt = 0; mt < RG_MT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (v & 1) {
                        W[mt][i] = (uint32_t)acc[mt][i];
                    } else {
                        W[mt][i] = (W[mt][i] << 8) + (uint32_t)acc[mt][i];
                        wl[(((v >> 1) * 4 + i) * RG_MT + mt) * 64 + lane] = W[mt][i];
                    }
                }
#pragma unroll
            for (int c = 0; c < KC; ++c) mcur[c] = mnxt[c];
            __builtin_amdgcn_sched_barrier(0);
        }
        // the words come back from LDS below, not from registers
        asm volatile("" ::: "memory");
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 16 * nt + 4 * g + i;
            if (j >= out_n) continue;
            const uint64_t *rw = rows + KB_MX_ROW * j;
            const uint64_t pj = mods_out[j], off = rw[0], np = rw[1];
            const double offd = __longlong_as_double((long long)rw[2]);
            const double invp = __longlong_as_double((long long)rw[3]);
#pragma unroll
            for (int mt = 0; mt < RG_MT; ++mt) {
                const uint32_t w3 = wl[((3 * 4 + i) * RG_MT + mt) * 64 + lane];
                const uint32_t w2 = wl[((2 * 4 + i) * RG_MT + mt) * 64 + lane];
                const uint32_t w1 = wl[((1 * 4 + i) * RG_MT + mt) * 64 + lane];
                const uint32_t w0 = wl[((0 * 4 + i) * RG_MT + mt) * 64 + lane];
                // V = w3 2^48 + w2 2^32 + w1 2^16 + w0 + off: exactly mod 2^64, and in double
                uint64_t lo = (uint64_t)w0 + off + ((uint64_t)w1 << 16) +
                              ((uint64_t)(w2 + (w3 << 16)) << 32);
                double vd = fma(fma(fma((double)w3, 65536.0, (double)w2), 65536.0, (double)w1), 65536.0,
                                (double)w0) + offd;
                if (SR) {
                    // diagonal term x_(qn+j) c_j (Shoup, [0, 2p)) and the rounded estimate
                    const uint64_t d = kb_shoup_lazy_nq(xp[(long)(in_n + j) * n + 16 * mt + p16], rw[4], rw[5], np);
                    const uint64_t al = alpha_s[wave][16 * mt + p16];
                    lo += d + al;
                    vd = vd + kb_u64_to_f64(d) + kb_u64_to_f64(al);
                }
                // |vd - V| < 2^30 and V < 2^80, so for p >= 2^58 the estimate minus
                // 2^-20 truncates to floor(V/p) or one less: r = V - qe p in [0, 2p)
                const uint32_t qe = (uint32_t)fmax(fma(vd, invp, -0x1p-20), 0.0);
                const uint64_t r = kb_mad32(qe, (uint32_t)np, lo) +
                                   ((uint64_t)kb_mul_lo32(qe, (uint32_t)(np >> 32)) << 32);
                yp[(long)j * n + 16 * mt + p16] = kb_csub(r, pj);
            }
        }
        // the next tile overwrites the words
        asm volatile("" ::: "memory");
    }
}

template <bool SR>
int launch_rns_gemm(int kc, int nj, dim3 grid, hipStream_t st, const uint64_t *x, long in_stride,
                    uint64_t *y, long out_stride, const uint64_t *mat, const uint64_t *mods_out,
                    const kbfv_tw *hatinv, const uint64_t *mods_in, const double *est, int in_n,
                    int out_n, int n) {
#define KB_RG(KC, NJ)                                                                                 \
    if (kc == KC && nj == NJ) {                                                                        \
        hipLaunchKernelGGL((k_rns_gemm<KC, NJ, SR>), grid, dim3(64 * RG_WAVES), 0, st, x, in_stride, y, \
                           out_stride, mat, mods_out, hatinv, mods_in, est, in_n, out_n, n);           \
        return 0;                                                                                      \
    }
    KB_RG(1, 1)
    KB_RG(2, 1)
    KB_RG(3, 1)
    KB_RG(4, 1)
    KB_RG(1, 2)
    KB_RG(2, 2)
    KB_RG(3, 2)
    KB_RG(4, 2)
#undef KB_RG
    return 1;
}

}  // namespace

// K rows (including the alpha row of base conversion) and output limbs a matrix
// of the MFMA path can hold; larger bases use the scalar kernels.
bool kb_rns_gemm_fits(int k_rows, int out_n) { return k_rows <= 32 && out_n <= 32; }

int kb_bconv_mfma(const kbfv_context *c, bool q_to_r, const uint64_t *x, long in_stride, uint64_t *y,
                  long out_stride, int polys, hipStream_t st, bool noscale) {
    const int in_n = q_to_r ? c->qn : c->rn, out_n = q_to_r ? c->rn : c->qn;
    const int kc = (in_n + 1 + 7) / 8, nj = (out_n + 15) / 16;
    const kb_dev_tables &t = c->t;
    const int in0 = q_to_r ? c->oQ : c->oR, out0 = q_to_r ? c->oR : c->oQ;
    const dim3 grid((unsigned)((long)polys * (c->n / RG_BLOCK)));
    if (launch_rns_gemm<false>(kc, nj, grid, st, x, in_stride, y, out_stride, q_to_r ? t.mx_qr : t.mx_rq,
                               t.mod + out0, q_to_r ? (noscale ? t.q_hatinv_n : t.q_hatinv) : t.r_hatinv,
                               t.mod + in0,
                               q_to_r ? t.q_inv : t.r_inv, in_n, out_n, c->n))
        return 1;
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_scale_round_mfma(const kbfv_context *c, const uint64_t *x, uint64_t *y, int polys, hipStream_t st) {
    const int kc = (c->qn + 7) / 8, nj = (c->rn + 15) / 16;
    const kb_dev_tables &t = c->t;
    const dim3 grid((unsigned)((long)polys * (c->n / RG_BLOCK)));
    if (launch_rns_gemm<true>(kc, nj, grid, st, x, (long)(c->qn + c->rn) * c->n, y, (long)c->rn * c->n,
                              t.mx_sr, t.mod + c->oR, nullptr, nullptr, t.sr_frac, c->qn, c->rn, c->n))
        return 1;
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
