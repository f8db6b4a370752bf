/*
 * Oracle: hybrid key switching in the auxiliary basis T (the reference's "23"
 * method: digit decomposition over blocks of K ciphertext primes, inner
 * product in T, per-block conversion back to PQ, modDown by P).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   modUpPQtoT_23_kernel          include/BasisConv.cuh:11-59   (key generation)
 *   modUpQjtoT_23_kernel          include/BasisConv.cuh:86-135
 *   external_product_T_kernel     include/ExternalProduct.cuh:11-53
 *   modUpTtoPQl_23_kernel         include/BasisConv.cuh:159-213
 *   modDownPQltoQl_23_*_kernel    include/BasisConv.cuh:233-298 (host :300-317)
 * with the float overflow estimate `vv += (float)x / (float)m`, `(int)vv`
 * and the 128-bit wrap-around of `acc - vv_floor * (M mod m)` restated
 * exactly.
 *
 * Pinned by (tests/test_oracle_ks.py), Python big integers: Q_j->T and PQ->T
 * equal the CRT value of the block residues in [0, Q_j) / [0, R_b) reduced mod
 * each t_k, exactly, whenever the exact rational sum of [x_j * hat_j^-1]/m_j
 * lies further than 2^-18 from an integer (the float estimate's error is below
 * 2^-21 there; at K=1 the sum is below 1 and the result is x mod t_k always);
 * the inner product is the mod-t_k dot product; T->PQ is the centered CRT
 * value (in (-T/2, T/2)) reduced mod each PQ prime -- the float estimate
 * cannot err there because the centred input keeps the fractional part near
 * 1/2; modDown equals (x_q - S) * P^-1 mod q exactly, with S = x_p at K=1 and
 * S = sum_i [x_i * phat_i^-1]_{p_i} * phat_i (not reduced mod P, as the
 * reference's multi-prime kernel leaves it) at K>1.  tests/test_oracle_mutations.py
 * shows these checks reject a rounded estimate and an added alpha*P
 * correction.  End to end, decrypt(relin(d2)) recovers s^2 * d2 up to small
 * noise (tests/test_oracle_bfv.py).
 */
#include <stdlib.h>
#include <string.h>

#include "ctx_internal.h"

static uint64_t reduce128(u128 a, uint64_t m) { return (uint64_t)(a % m); }

/* The Q_j -> T and PQ -> T conversions truncate the float overflow estimate
 * (`(int)vv`, include/BasisConv.cuh:32,112). */
#if OR_MUTATION == 2 /* planted mistake (tests only): the estimate rounded */
#include <math.h>
#define TRUNC_VV(v) ((int)lrintf(v))
#else
#define TRUNC_VV(v) ((int)(v))
#endif

/* x_q: [qn][N] coefficient form -> out_t: [dnum][tn][N] */
void or_modup_qj_to_t(const or_ctx *c, const uint64_t *x_q, uint64_t *out_t) {
    size_t N = c->n;
    int pn = c->pn, qn = c->qn, tn = c->tn;
    uint64_t rb[64];
    for (int b = 0; b < c->dnum; b++) {
        for (size_t k = 0; k < N; k++) {
            float vv = 0;
            int cnt = 0;
            for (int j = 0; j < pn && b * pn + j < qn; j++, cnt++) {
                int qi_idx = b * pn + j;
                uint64_t qi = c->mod[c->oQ + qi_idx];
                rb[j] = or_mulmod(x_q[(size_t)qi_idx * N + k], c->qj_hatinv[b * pn + j], qi);
                float rr = (float)rb[j];
                vv += rr / (float)qi;
            }
            int vf = TRUNC_VV(vv);
            for (int i = 0; i < tn; i++) {
                uint64_t t = c->mod[c->oT + i];
                u128 acc = 0;
                for (int j = 0; j < cnt; j++) acc += (u128)rb[j] * c->qj_hat_mod_t[(b * tn + i) * pn + j];
                acc -= (u128)(uint64_t)((uint64_t)vf * c->qj_mod_t[b * tn + i]);
                out_t[((size_t)b * tn + i) * N + k] = reduce128(acc, t);
            }
        }
    }
}

/* x_pq: [pn+qn][N] coefficient form -> out_t: [ri_blocks][tn][N] */
void or_modup_pq_to_t(const or_ctx *c, const uint64_t *x_pq, uint64_t *out_t) {
    size_t N = c->n;
    int pq = c->pn + c->qn, tn = c->tn, g = c->gamma;
    uint64_t rb[64];
    for (int b = 0; b < c->ri_blocks; b++) {
        for (size_t k = 0; k < N; k++) {
            float vv = 0;
            int cnt = 0;
            for (int j = 0; j < g && b * g + j < pq; j++, cnt++) {
                uint64_t ri = c->mod[b * g + j];
                rb[j] = or_mulmod(x_pq[(size_t)(b * g + j) * N + k], c->ri_hatinv[b * g + j], ri);
                vv += (float)rb[j] / (float)ri;
            }
            int vf = TRUNC_VV(vv);
            for (int i = 0; i < tn; i++) {
                uint64_t t = c->mod[c->oT + i];
                u128 acc = 0;
                for (int j = 0; j < cnt; j++) acc += (u128)rb[j] * c->ri_hat_mod_t[(b * tn + i) * g + j];
                acc -= (u128)(uint64_t)((uint64_t)vf * c->ri_mod_t[i * c->ri_blocks + b]);
                out_t[((size_t)b * tn + i) * N + k] = reduce128(acc, t);
            }
        }
    }
}

/* cin_t: [dnum][tn][N] (NTT over T), key_t: [2][key_dnum][key_blocks][tn][N]
 * (NTT; its first dnum digits and ri_blocks blocks) -> out_t: [2][ri_blocks][tn][N] */
void or_inner_product(const or_ctx *c, const uint64_t *cin_t, const uint64_t *key_t,
                      uint64_t *out_t) {
    size_t N = c->n;
    int tn = c->tn, B = c->ri_blocks, D = c->dnum, KD = c->key_dnum, KB = c->key_blocks;
    for (int h = 0; h < 2; h++)
        for (int b = 0; b < B; b++)
            for (int k = 0; k < tn; k++) {
                uint64_t t = c->mod[c->oT + k];
                for (size_t n = 0; n < N; n++) {
                    u128 acc = 0;
                    for (int i = 0; i < D; i++) {
                        uint64_t x = cin_t[((size_t)i * tn + k) * N + n];
                        uint64_t y = key_t[((((size_t)h * KD + i) * KB + b) * tn + k) * N + n];
                        acc += (u128)x * y;
                    }
                    out_t[(((size_t)h * B + b) * tn + k) * N + n] = reduce128(acc, t);
                }
            }
}

/* in_t: [2][ri_blocks][tn][N] coefficient form -> out_pq: [2][pn+qn][N] */
void or_modup_t_to_pq(const or_ctx *c, const uint64_t *in_t, uint64_t *out_pq) {
    size_t N = c->n;
    int pq = c->pn + c->qn, tn = c->tn, g = c->gamma, B = c->ri_blocks;
    uint64_t rb[64];
    for (int h = 0; h < 2; h++)
        for (int b = 0; b < B; b++)
            for (size_t k = 0; k < N; k++) {
                float vv = 0;
                for (int i = 0; i < tn; i++) {
                    uint64_t t = c->mod[c->oT + i];
                    uint64_t ra = in_t[(((size_t)h * B + b) * tn + i) * N + k] + c->halfT_mod[pq + i];
                    rb[i] = or_mulmod(ra, c->t_hatinv[i], t);
                    float rr = (float)rb[i];
                    vv += rr / (float)t;
                }
                int vf = (int)vv;
                for (int j = 0; j < g && b * g + j < pq; j++) {
                    int idx = b * g + j;
                    uint64_t r = c->mod[idx];
                    u128 acc = 0;
                    for (int i = 0; i < tn; i++) acc += (u128)rb[i] * c->t_hat_mod_pq[idx * tn + i];
                    acc -= (u128)(uint64_t)((uint64_t)vf * c->t_mod_pq[idx]);
                    uint64_t v = reduce128(acc, r);
                    v = v + r - c->halfT_mod[idx];
                    if (v >= r) v -= r;
                    out_pq[((size_t)h * pq + idx) * N + k] = v;
                }
            }
}

/* in_pq: [2][pn+qn][N] coefficient form -> out_q: [2][qn][N] */
void or_moddown(const or_ctx *c, const uint64_t *in_pq, uint64_t *out_q) {
    size_t N = c->n;
    int pn = c->pn, qn = c->qn, pq = pn + qn;
    for (int h = 0; h < 2; h++)
        for (int l = 0; l < qn; l++) {
            uint64_t q = c->mod[c->oQ + l];
            for (size_t k = 0; k < N; k++) {
                uint64_t ra;
                if (pn == 1) {
                    ra = in_pq[((size_t)h * pq) * N + k] % q;
                } else {
                    u128 acc = 0;
                    for (int i = 0; i < pn; i++) {
                        uint64_t p = c->mod[c->oP + i];
                        uint64_t rb = or_mulmod(in_pq[((size_t)h * pq + i) * N + k], c->p_hatinv[i], p);
                        acc += (u128)rb * c->p_hat_mod_q[l * pn + i];
                    }
                    ra = reduce128(acc, q);
#if OR_MUTATION == 3 /* planted mistake (tests only): an alpha*P correction the reference does not make */
                    double a = 0;
                    for (int i = 0; i < pn; i++) {
                        uint64_t p = c->mod[c->oP + i];
                        a += (double)or_mulmod(in_pq[((size_t)h * pq + i) * N + k], c->p_hatinv[i], p) / (double)p;
                    }
                    uint64_t pm = 1;
                    for (int i = 0; i < pn; i++) pm = or_mulmod(pm, c->mod[c->oP + i] % q, q);
                    ra = (ra + q - or_mulmod((uint64_t)llround(a) % q, pm, q)) % q;
#endif
                }
                ra = q - ra + in_pq[((size_t)h * pq + pn + l) * N + k];
                out_q[((size_t)h * qn + l) * N + k] = or_mulmod(ra, c->Pinv_mod_q[l], q);
            }
        }
}
