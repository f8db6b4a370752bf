// RNS kernels of the HPS HomMult path, BFV encryption scaling and decryption.
//
// Replaces (include/poly_arithmetic.cuh): the element-wise batch kernels
// (:26-360), bfv_add_timesQ_overt_kernel (:902), the hps_decrypt_scale_and_round
// kernels (:929-1076), bconv_mult_unroll2 + base_convert_matmul_hps_unroll2
// (:1078-1178), tensor_prod_2x2_rns_poly (:1180) and scaleAndRound_HPS_QR_R
// (:1263-1317).  The double-precision estimates keep the reference's operation
// order with explicit fma (the reference's nvcc build contracts `acc += x*c`);
// everything else is exact modular arithmetic with canonical outputs.
//
// Base conversion and scale-and-round are (coefficient x small matrix)
// products: a workgroup takes 64 coefficients, stages the scaled inputs in LDS
// (one pass over HBM), and its four waves each produce a quarter of the output
// limbs with the matrix row read through the scalar cache (the row index is
// wave-uniform).
#include "kbfv_internal.h"
#include "kbfv_ops.h"

namespace {

constexpr int BC_TILE = 64;   // coefficients per workgroup
constexpr int BC_WAVES = 4;   // output-limb groups per workgroup
constexpr int BC_MAXIN = 64;  // max input limbs staged in LDS

// ------------------------------------------------------------------ element-wise
template <int OP>
__global__ void k_poly_binop(kb_dev_tables t, const uint64_t *__restrict__ a,
                             const uint64_t *__restrict__ b, uint64_t *__restrict__ o, int mod0,
                             int limbs, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int l = (int)((i / n) % limbs);
    const int m = mod0 + l;
    const uint64_t q = t.mod[m];
    const uint64_t x = a[i], y = b[i];
    uint64_t r;
    if constexpr (OP == 0)
        r = kb_addmod(x, y, q);
    else if constexpr (OP == 1)
        r = kb_submod(x, y, q);
    else
        r = kb_mulmod(x, y, q, t.mu[2 * m], t.mu[2 * m + 1]);
    o[i] = r;
}

// ------------------------------------------------------------------ HPS base conversion
// x: [polys][in_n][N] (stride in_stride), y: [polys][out_n][N] (stride out_stride)
// Phase 1 (all waves): y_i = [x_i * (Q/q_i)^-1]_{q_i} into LDS.  Phase 2 (wave 0):
// the reference's double estimate of the overflow count, summed in input order.
// Phase 3 (all waves, one output limb per wave at a time): 128-bit MACs with
// the matrix row read through the scalar cache.
__global__ void __launch_bounds__(BC_TILE *BC_WAVES)
    k_bconv(const uint64_t *__restrict__ x, long in_stride, uint64_t *__restrict__ y, long out_stride,
            const uint64_t *__restrict__ mods, const uint64_t *__restrict__ mus,
            const kbfv_tw *__restrict__ hatinv, const uint64_t *__restrict__ hat_mod_out,
            const uint64_t *__restrict__ alpha_mod_out, const double *__restrict__ inv, int in0,
            int in_n, int out0, int out_n, int n) {
    __shared__ uint64_t ys[BC_MAXIN][BC_TILE];
    __shared__ int alpha_s[BC_TILE];
    const int tiles = n / BC_TILE;
    const int poly = blockIdx.x / tiles;
    const int k0 = (blockIdx.x % tiles) * BC_TILE;
    const uint64_t *xp = x + poly * in_stride;
    uint64_t *yp = y + poly * out_stride;
    const int lane = threadIdx.x % BC_TILE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / BC_TILE);
    for (int i = wave; i < in_n; i += BC_WAVES) {
        const uint64_t qi = mods[in0 + i];
        const kbfv_tw h = hatinv[i];
        ys[i][lane] = kb_shoup(xp[(long)i * n + k0 + lane], h.w, h.ws, qi);
    }
    __syncthreads();
    if (wave == 0) {
        double frac = 0.0;
        for (int i = 0; i < in_n; ++i) frac = fma((double)ys[i][lane], inv[i], frac);
        alpha_s[lane] = (int)llround(frac);
    }
    __syncthreads();
    const int alpha = alpha_s[lane];
    for (int j = wave; j < out_n; j += BC_WAVES) {
        const int m = out0 + j;
        const uint64_t pj = mods[m];
        const uint64_t *row = hat_mod_out + (long)j * in_n;
        kbfv_u128 acc{0, 0};
        int i = 0;
        for (; i + 4 <= in_n; i += 4) {
            kb_mac128(acc, ys[i][lane], row[i]);
            kb_mac128(acc, ys[i + 1][lane], row[i + 1]);
            kb_mac128(acc, ys[i + 2][lane], row[i + 2]);
            kb_mac128(acc, ys[i + 3][lane], row[i + 3]);
        }
        for (; i < in_n; ++i) kb_mac128(acc, ys[i][lane], row[i]);
        const uint64_t v = kb_barrett128(acc, pj, mus[2 * m], mus[2 * m + 1]);
        yp[(long)j * n + k0 + lane] = kb_submod(v, alpha_mod_out[(long)alpha * out_n + j], pj);
    }
}

// ------------------------------------------------------------------ tensor
// Q part from the NTT-domain ciphertexts, R part from the extended polys.
// ct: [batch][2][qn][N]; r: [batch][2][rn][N]; out: [batch][3][qn+rn][N]
// grid (n / 256, qn + rn, batch): coefficient, limb l, ciphertext b
__global__ void k_tensor_qr(kb_dev_tables t, const uint64_t *__restrict__ ct1,
                            const uint64_t *__restrict__ ct2, const uint64_t *__restrict__ r1,
                            const uint64_t *__restrict__ r2, uint64_t *__restrict__ out, int qn,
                            int rn, int oQ, int n) {
    const int sn = qn + rn;
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int l = blockIdx.y;
    const long b = blockIdx.z;
    const int m = oQ + l;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    uint64_t a1, b1, a2, b2;
    if (l < qn) {
        const long base = b * 2 * qn * (long)n + (long)l * n + k;
        a1 = ct1[base];
        b1 = ct1[base + (long)qn * n];
        a2 = ct2[base];
        b2 = ct2[base + (long)qn * n];
    } else {
        const long base = b * 2 * rn * (long)n + (long)(l - qn) * n + k;
        a1 = r1[base];
        b1 = r1[base + (long)rn * n];
        a2 = r2[base];
        b2 = r2[base + (long)rn * n];
    }
    const uint64_t d0 = kb_mulmod(a1, a2, q, mlo, mhi);
    const uint64_t d2 = kb_mulmod(b1, b2, q, mlo, mhi);
    // Karatsuba middle term, as the reference (include/poly_arithmetic.cuh:1208-1211)
    uint64_t d1 = kb_mulmod(a1 + b1, a2 + b2, q, mlo, mhi);
    d1 = kb_csub(kb_csub(d1 + 2 * q - d0 - d2, q), q);
    const long ob = b * 3 * sn * (long)n + (long)l * n + k;
    out[ob] = d0;
    out[ob + (long)sn * n] = d1;
    out[ob + 2L * sn * n] = d2;
}

// generic tensor over a contiguous [batch][2][limbs][N] pair -> [batch][3][limbs][N]
__global__ void k_tensor(kb_dev_tables t, const uint64_t *__restrict__ ct1,
                         const uint64_t *__restrict__ ct2, uint64_t *__restrict__ out, int mod0,
                         int limbs, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n);
    const long rest = i / n;
    const int l = (int)(rest % limbs);
    const long b = rest / limbs;
    const int m = mod0 + l;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const long L = (long)limbs * n;
    const long base = b * 2 * L + (long)l * n + k;
    const uint64_t a1 = ct1[base], b1 = ct1[base + L], a2 = ct2[base], b2 = ct2[base + L];
    const uint64_t d0 = kb_mulmod(a1, a2, q, mlo, mhi);
    const uint64_t d2 = kb_mulmod(b1, b2, q, mlo, mhi);
    uint64_t d1 = kb_mulmod(a1 + b1, a2 + b2, q, mlo, mhi);
    d1 = kb_csub(kb_csub(d1 + 2 * q - d0 - d2, q), q);
    const long ob = b * 3 * L + (long)l * n + k;
    out[ob] = d0;
    out[ob + L] = d1;
    out[ob + 2 * L] = d2;
}

// ------------------------------------------------------------------ scale-and-round QR -> R
// x: [polys][qn+rn][N] coefficient form -> y: [polys][rn][N]
__global__ void __launch_bounds__(BC_TILE *BC_WAVES)
    k_scale_round(kb_dev_tables t, const uint64_t *__restrict__ x, uint64_t *__restrict__ y, int qn,
                  int rn, int oR, int n) {
    __shared__ uint64_t xs[BC_MAXIN][BC_TILE];
    __shared__ uint64_t alpha_s[BC_TILE];
    const int tiles = n / BC_TILE;
    const int poly = blockIdx.x / tiles;
    const int k0 = (blockIdx.x % tiles) * BC_TILE;
    const int sn = qn + rn;
    const uint64_t *xp = x + (long)poly * sn * n;
    uint64_t *yp = y + (long)poly * rn * n;
    const int lane = threadIdx.x % BC_TILE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / BC_TILE);
    for (int i = wave; i < qn; i += BC_WAVES) xs[i][lane] = xp[(long)i * n + k0 + lane];
    __syncthreads();
    if (wave == 0) {
        double nu = 0.5;
        for (int i = 0; i < qn; ++i) nu = fma((double)xs[i][lane], t.sr_frac[i], nu);
        alpha_s[lane] = kb_f64_to_u64_sat(nu);
    }
    __syncthreads();
    const uint64_t alpha = alpha_s[lane];
    for (int j = wave; j < rn; j += BC_WAVES) {
        const int m = oR + j;
        const uint64_t rj = t.mod[m];
        const uint64_t *row = t.sr_modr + (long)j * (qn + 1);
        kbfv_u128 acc{0, 0};
        int i = 0;
        for (; i + 4 <= qn; i += 4) {
            kb_mac128(acc, xs[i][lane], row[i]);
            kb_mac128(acc, xs[i + 1][lane], row[i + 1]);
            kb_mac128(acc, xs[i + 2][lane], row[i + 2]);
            kb_mac128(acc, xs[i + 3][lane], row[i + 3]);
        }
        for (; i < qn; ++i) kb_mac128(acc, xs[i][lane], row[i]);
        kb_mac128(acc, xp[(long)(qn + j) * n + k0 + lane], row[qn]);
        const uint64_t v = kb_barrett128(acc, rj, t.mu[2 * m], t.mu[2 * m + 1]);
        const uint64_t a = kb_barrett128(kbfv_u128{alpha, 0}, rj, t.mu[2 * m], t.mu[2 * m + 1]);
        yp[(long)j * n + k0 + lane] = kb_addmod(v, a, rj);
    }
}

// ------------------------------------------------------------------ HPS decrypt
// x: [batch][qn][N] coefficient form -> m: [batch][N]
template <bool LARGE, bool LAZY>
__global__ void k_hps_decrypt(kb_dev_tables t, const uint64_t *__restrict__ x,
                              uint64_t *__restrict__ m, int qn, uint64_t plain, int hf, int n,
                              long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long b = i / n;
    const int k = (int)(i % n);
    const uint64_t *xp = x + b * (long)qn * n + k;
    double dsum = 0.0;
    uint64_t isum = 0;
    const double tinv = 1. / (double)plain;
    for (int l = 0; l < qn; ++l) {
        const uint64_t tmp = xp[(long)l * n];
        const kbfv_tw c = t.dec_modt[l];
        if constexpr (!LARGE) {
            dsum = fma((double)tmp, t.dec_frac[l], dsum);
            isum += LAZY ? tmp * c.w : kb_shoup(tmp, c.w, c.ws, plain);
        } else {
            const kbfv_tw cb = t.dec_B_modt[l];
            const uint64_t hi = tmp >> hf, lo = tmp & ((1ULL << hf) - 1);
            dsum = fma((double)lo, t.dec_frac[l], dsum);
            dsum = fma((double)hi, t.dec_B_frac[l], dsum);
            if constexpr (LAZY) {
                isum += lo * c.w;
                isum += hi * cb.w;
            } else {
                isum += kb_shoup(lo, c.w, c.ws, plain);
                isum += kb_shoup(hi, cb.w, cb.ws, plain);
            }
        }
    }
    dsum += (double)isum;
    const uint64_t quot = (uint64_t)(dsum * tinv);
    dsum -= (double)(plain * quot);
    m[i] = (uint64_t)llround(dsum);
}

// ------------------------------------------------------------------ Delta * m
// m: [batch][N] (mod t) -> out: [batch][qn][N] = ceil(Q m / t) mod q_i
__global__ void k_delta_m(kb_dev_tables t, const uint64_t *__restrict__ m, uint64_t *__restrict__ out,
                          int qn, int oQ, uint64_t plain, uint64_t negq, uint64_t negq_s, int n,
                          long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n);
    const long rest = i / n;
    const int l = (int)(rest % qn);
    const long b = rest / qn;
    const uint64_t q = t.mod[oQ + l];
    const uint64_t mq = kb_shoup(m[b * n + k], negq, negq_s, plain);
    const kbfv_tw ti = t.tinv_mod_q[l];
    out[i] = kb_shoup(mq, ti.w, ti.ws, q);
}

// x mod t for decode of values that may equal t (llround of an upper tie)
__global__ void k_reduce_plain(uint64_t *__restrict__ v, uint64_t plain, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) v[i] = v[i] % plain;
}

inline unsigned blocks_for(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------- launchers
int kb_poly_binop(const kbfv_context *c, int op, const uint64_t *a, const uint64_t *b, uint64_t *o,
                  int mod0, int limbs, int polys, hipStream_t st) {
    KB_REQUIRE(c && a && b && o, "null argument");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod && polys > 0, "bad limb range");
    const long total = (long)polys * limbs * c->n;
    const unsigned g = blocks_for(total, 256);
    if (op == 0)
        hipLaunchKernelGGL(k_poly_binop<0>, dim3(g), dim3(256), 0, st, c->t, a, b, o, mod0, limbs, c->n, total);
    else if (op == 1)
        hipLaunchKernelGGL(k_poly_binop<1>, dim3(g), dim3(256), 0, st, c->t, a, b, o, mod0, limbs, c->n, total);
    else
        hipLaunchKernelGGL(k_poly_binop<2>, dim3(g), dim3(256), 0, st, c->t, a, b, o, mod0, limbs, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_bconv(const kbfv_context *c, bool q_to_r, const uint64_t *x, long in_stride, uint64_t *y,
             long out_stride, int polys, hipStream_t st, bool noscale) {
    KB_REQUIRE(c && x && y && polys > 0, "bad bconv arguments");
    KB_REQUIRE(c->rn <= BC_MAXIN, "too many limbs for base conversion");
    KB_REQUIRE(q_to_r || !noscale, "unscaled input only for Q -> R");
    if (c->t.mx_qr && kb_rns_gemm_fits((q_to_r ? c->qn : c->rn) + 1, q_to_r ? c->rn : c->qn)) {
        const int r = kb_bconv_mfma(c, q_to_r, x, in_stride, y, out_stride, polys, st, noscale);
        if (r != 1) return r;
    }
    const int tiles = c->n / BC_TILE;
    const kb_dev_tables &t = c->t;
    // y_i = [x_i (Q/q_i)^-1]_q_i is a Shoup product, exact for any 64-bit x_i:
    // an unscaled input takes the constant with n^-1 folded in
    if (q_to_r)
        hipLaunchKernelGGL(k_bconv, dim3(tiles * polys), dim3(BC_TILE * BC_WAVES), 0, st, x, in_stride, y,
                           out_stride, t.mod, t.mu, noscale ? t.q_hatinv_n : t.q_hatinv, t.q_hat_mod_r,
                           t.alphaQ_mod_r, t.q_inv, c->oQ, c->qn, c->oR, c->rn, c->n);
    else
        hipLaunchKernelGGL(k_bconv, dim3(tiles * polys), dim3(BC_TILE * BC_WAVES), 0, st, x, in_stride, y,
                           out_stride, t.mod, t.mu, t.r_hatinv, t.r_hat_mod_q, t.alphaR_mod_q, t.r_inv,
                           c->oR, c->rn, c->oQ, c->qn, c->n);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_tensor_qr(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, const uint64_t *r1,
                 const uint64_t *r2, uint64_t *out, int batch, hipStream_t st) {
    KB_REQUIRE(batch <= 65535, "batch too large for one tensor launch");
    hipLaunchKernelGGL(k_tensor_qr, dim3(c->n / 256, c->qn + c->rn, batch), dim3(256), 0, st, c->t, ct1, ct2,
                       r1, r2, out, c->qn, c->rn, c->oQ, c->n);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_tensor(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out, int mod0,
              int limbs, int batch, hipStream_t st) {
    KB_REQUIRE(c && ct1 && ct2 && out && batch > 0, "bad tensor arguments");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod, "bad limb range");
    const long total = (long)batch * limbs * c->n;
    hipLaunchKernelGGL(k_tensor, dim3(blocks_for(total, 256)), dim3(256), 0, st, c->t, ct1, ct2, out, mod0,
                       limbs, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_scale_round(const kbfv_context *c, const uint64_t *x, uint64_t *y, int polys, hipStream_t st) {
    KB_REQUIRE(c && x && y && polys > 0, "bad scale-round arguments");
    KB_REQUIRE(c->qn <= BC_MAXIN, "too many limbs for scale-and-round");
    if (c->t.mx_sr && kb_rns_gemm_fits(c->qn, c->rn)) {
        const int r = kb_scale_round_mfma(c, x, y, polys, st);
        if (r != 1) return r;
    }
    const int tiles = c->n / BC_TILE;
    hipLaunchKernelGGL(k_scale_round, dim3(tiles * polys), dim3(BC_TILE * BC_WAVES), 0, st, c->t, x, y,
                       c->qn, c->rn, c->oR, c->n);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_hps_decrypt(const kbfv_context *c, const uint64_t *x, uint64_t *m, int batch, hipStream_t st) {
    KB_REQUIRE(c && x && m && batch > 0, "bad decrypt arguments");
    const long total = (long)batch * c->n;
    const bool large = c->qMSB + c->sizeQMSB >= 52;
    const int hf = c->qMSB >> 1;
    const bool lazy = large ? (hf + c->tMSB + c->sizeQMSB) < 52 : (c->qMSB + c->tMSB + c->sizeQMSB) < 52;
    const dim3 g(blocks_for(total, 256)), b(256);
    if (large && lazy)
        hipLaunchKernelGGL((k_hps_decrypt<true, true>), g, b, 0, st, c->t, x, m, c->qn, c->plain, hf, c->n, total);
    else if (large)
        hipLaunchKernelGGL((k_hps_decrypt<true, false>), g, b, 0, st, c->t, x, m, c->qn, c->plain, hf, c->n, total);
    else if (lazy)
        hipLaunchKernelGGL((k_hps_decrypt<false, true>), g, b, 0, st, c->t, x, m, c->qn, c->plain, hf, c->n, total);
    else
        hipLaunchKernelGGL((k_hps_decrypt<false, false>), g, b, 0, st, c->t, x, m, c->qn, c->plain, hf, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_delta_m(const kbfv_context *c, const uint64_t *m, uint64_t *out, int batch, hipStream_t st) {
    const long total = (long)batch * c->qn * c->n;
    hipLaunchKernelGGL(k_delta_m, dim3(blocks_for(total, 256)), dim3(256), 0, st, c->t, m, out, c->qn, c->oQ,
                       c->plain, c->negQ_mod_t, c->negQ_mod_t_shoup, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_reduce_plain(const kbfv_context *c, uint64_t *v, long count, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_plain, dim3(blocks_for(count, 256)), dim3(256), 0, st, v, c->plain, count);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
