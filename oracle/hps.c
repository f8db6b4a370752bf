/*
 * Oracle: HPS tensoring pieces, BFV encode-scale and HPS decryption.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   BaseConverter::bConv_HPS                    include/BasisConv_QR.cuh:191-208
 *   bconv_mult_unroll2_kernel                   include/poly_arithmetic.cuh:1078-1099
 *   base_convert_matmul_hps_unroll2_kernel      include/poly_arithmetic.cuh:1101-1178
 *   tensor_prod_2x2_rns_poly                    include/poly_arithmetic.cuh:1180-1219
 *   scaleAndRound_HPS_QR_R_kernel               include/poly_arithmetic.cuh:1263-1307
 *   hps_decrypt_scale_and_round_kernel_*        include/poly_arithmetic.cuh:929-1076,
 *       dispatch                                include/BFVcontext.cuh:1680-1717
 *   bfv_add_timesQ_overt_kernel                 include/poly_arithmetic.cuh:902-927
 *
 * The double-precision estimates are restated with the reference's operation
 * order; `acc += (double)x * c` is written fma((double)x, c, acc).
 *
 * Pinned by (tests/test_oracle_hps.py): exact CRT in Python big integers.
 * bConv must equal the centered lift of x mod Q reduced mod r_j (or that lift
 * +-Q when the double estimate sits within 2^-30 of a rounding tie);
 * scale-and-round must equal floor(t*X/Q + 1/2) mod r_j exactly when the Q
 * residues are below 2^20 (the double estimate is then exact to 2^-27) and
 * t*X/Q is further than 2^-20 from a half-integer, and the same up to one small
 * offset common to every r_j on full-size residues (the estimate's 60-bit
 * rounding); decrypt must equal round(t*x/Q) mod t (exactly on small residues
 * away from ties); the tensor equals the mod-q products; Delta-m equals
 * ceil(Q*m/t) mod q_i.  tests/test_oracle_mutations.py shows the checks reject
 * a dropped rounding half and a rounded (not truncated) nu.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ctx_internal.h"

static uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}
static uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

/* generic HPS fast base conversion with the double-estimated overflow count */
static void bconv(const or_ctx *c, const uint64_t *x, int in0, int in_n, const uint64_t *hatinv,
                  const uint64_t *hat_mod_out /*[out_n][in_n]*/,
                  const uint64_t *alpha_mod_out /*[in_n+1][out_n]*/, const double *inv,
                  uint64_t *y, int out0, int out_n) {
    size_t N = c->n;
    uint64_t *t = (uint64_t *)malloc(sizeof(uint64_t) * in_n);
    for (size_t k = 0; k < N; k++) {
        double frac = 0.0;
        for (int i = 0; i < in_n; i++) {
            uint64_t qi = c->mod[in0 + i];
            t[i] = or_mulmod(x[(size_t)i * N + k], hatinv[i], qi);
            frac = fma((double)t[i], inv[i], frac);
        }
        long long alpha = llround(frac);
        for (int j = 0; j < out_n; j++) {
            uint64_t pj = c->mod[out0 + j];
            u128 acc = 0;
            for (int i = 0; i < in_n; i++) acc += (u128)t[i] * hat_mod_out[(size_t)j * in_n + i];
            uint64_t v = (uint64_t)(acc % pj);
            y[(size_t)j * N + k] = sub_mod(v, alpha_mod_out[(size_t)alpha * out_n + j], pj);
        }
    }
    free(t);
}

void or_bconv_q_to_r(const or_ctx *c, const uint64_t *x_q, uint64_t *y_r) {
    bconv(c, x_q, c->oQ, c->qn, c->q_hatinv, c->q_hat_mod_r, c->alphaQ_mod_r, c->q_inv, y_r, c->oR,
          c->rn);
}
void or_bconv_r_to_q(const or_ctx *c, const uint64_t *x_r, uint64_t *y_q) {
    bconv(c, x_r, c->oR, c->rn, c->r_hatinv, c->r_hat_mod_q, c->alphaR_mod_q, c->r_inv, y_q, c->oQ,
          c->qn);
}

/* ct1, ct2: [2][limbs][N] (a, b); out: [3][limbs][N] = (a1a2, a1b2+a2b1, b1b2) */
void or_tensor(const or_ctx *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out, int limbs,
               int mod0) {
    size_t N = c->n, L = (size_t)limbs * N;
    for (int i = 0; i < limbs; i++) {
        uint64_t q = c->mod[mod0 + i];
        for (size_t k = 0; k < N; k++) {
            size_t o = (size_t)i * N + k;
            uint64_t a1 = ct1[o], b1 = ct1[L + o], a2 = ct2[o], b2 = ct2[L + o];
            uint64_t d0 = or_mulmod(a1, a2, q), d2 = or_mulmod(b1, b2, q);
            uint64_t d1 = add_mod(or_mulmod(a1, b2, q), or_mulmod(a2, b1, q), q);
            out[o] = d0;
            out[L + o] = d1;
            out[2 * L + o] = d2;
        }
    }
}

/* x_qr: [qn + rn][N] coefficient form; y_r: [rn][N] */
void or_scale_round(const or_ctx *c, const uint64_t *x_qr, uint64_t *y_r) {
    size_t N = c->n;
    int qn = c->qn, rn = c->rn;
    for (size_t k = 0; k < N; k++) {
#if OR_MUTATION == 1 /* planted mistake (tests only): the rounding half dropped */
        double nu = 0.0;
#else
        double nu = 0.5;
#endif
        for (int i = 0; i < qn; i++) nu = fma((double)x_qr[(size_t)i * N + k], c->sr_frac[i], nu);
        /* the reference's static_cast<uint64> of nu is a saturating conversion
         * (nvcc's cvt.rzi.u64.f64): nu can exceed 2^64 when every x_i is near q_i */
#if OR_MUTATION == 4 /* planted mistake (tests only): nu rounded instead of truncated */
        uint64_t alpha = (uint64_t)llround(nu);
#else
        uint64_t alpha = nu >= 18446744073709551616.0 ? UINT64_MAX : (uint64_t)nu;
#endif
        for (int j = 0; j < rn; j++) {
            uint64_t rj = c->mod[c->oR + j];
            const uint64_t *row = c->sr_modr + (size_t)j * (qn + 1);
            u128 acc = 0;
            for (int i = 0; i < qn; i++) acc += (u128)x_qr[(size_t)i * N + k] * row[i];
            acc += (u128)x_qr[(size_t)(qn + j) * N + k] * row[qn];
            uint64_t v = (uint64_t)(acc % rj);
            y_r[(size_t)j * N + k] = add_mod(v, alpha % rj, rj);
        }
    }
}

/* x_q: [qn][N] coefficient form of b + a*s; m: [N] */
void or_hps_decrypt(const or_ctx *c, const uint64_t *x_q, uint64_t *m) {
    size_t N = c->n;
    int qn = c->qn;
    uint64_t t = c->plain;
    int large = c->qMSB + c->sizeQMSB >= 52;
    int hf = c->qMSB >> 1;
    int lazy = large ? (hf + c->tMSB + c->sizeQMSB) < 52 : (c->qMSB + c->tMSB + c->sizeQMSB) < 52;
    for (size_t k = 0; k < N; k++) {
        double dsum = 0.0;
        uint64_t isum = 0;
        double tinv = 1. / (double)t;
        for (int i = 0; i < qn; i++) {
            uint64_t tmp = x_q[(size_t)i * N + k];
            if (!large) {
                dsum = fma((double)tmp, c->dec_frac[i], dsum);
                isum += lazy ? tmp * c->dec_modt[i] : or_mulmod(tmp, c->dec_modt[i], t);
            } else {
                uint64_t hi = tmp >> hf, lo = tmp & ((1ULL << hf) - 1);
                dsum = fma((double)lo, c->dec_frac[i], dsum);
                dsum = fma((double)hi, c->dec_B_frac[i], dsum);
                if (lazy) {
                    isum += lo * c->dec_modt[i];
                    isum += hi * c->dec_B_modt[i];
                } else {
                    isum += or_mulmod(lo, c->dec_modt[i], t);
                    isum += or_mulmod(hi, c->dec_B_modt[i], t);
                }
            }
        }
        dsum += (double)isum;
        uint64_t quot = (uint64_t)(dsum * tinv);
        dsum -= (double)(t * quot);
        m[k] = (uint64_t)llround(dsum);
    }
}

/* Delta*m in RNS as ceil(Q*m/t) mod q_i = [m * (-Q mod t)]_t * t^-1 mod q_i */
void or_add_delta_m(const or_ctx *c, const uint64_t *m, uint64_t *out_q) {
    size_t N = c->n;
    for (int i = 0; i < c->qn; i++) {
        uint64_t qi = c->mod[c->oQ + i];
        for (size_t k = 0; k < N; k++) {
            uint64_t mq = or_mulmod(m[k], c->negQ_mod_t, c->plain);
            out_q[(size_t)i * N + k] = or_mulmod(mq, c->tinv_mod_q[i], qi);
        }
    }
}
