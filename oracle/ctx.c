/*
 * Oracle: BFV context (parameters and every precomputed table).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows BFVContext::BFVContext / preComputeOnCPU / copyMemoryToGPU
 * (include/BFVcontext.cuh:13-101, 130-542, 545-1395) and BaseConverter::init
 * (include/BasisConv_QR.cuh:53-189).  Moduli order differs from the reference's
 * separate pqt / qr tables only in storage: one table [P | Q | R | T | t].
 *
 * Pinned by (tests/test_oracle_ctx.py): Python big-integer recomputation of
 * every table from its closed form (punctured products, inverses, floor and
 * fraction of t*R*(S/s_i)^-1 / s_i, floor(T/2) mod m), and the R basis of the
 * reference's logN=15 chain recorded in tests/golden.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ctx_internal.h"

/* ---------------- small bigint ---------------- */
void or_big_set(or_big *b, uint64_t v) {
    memset(b, 0, sizeof(*b));
    b->w[0] = v;
    b->n = 1;
}
void or_big_mul_u64(or_big *b, uint64_t v) {
    u128 carry = 0;
    for (int i = 0; i < b->n; i++) {
        u128 p = (u128)b->w[i] * v + carry;
        b->w[i] = (uint64_t)p;
        carry = p >> 64;
    }
    if (carry) {
        if (b->n >= OR_MAXW) abort();
        b->w[b->n++] = (uint64_t)carry;
    }
}
uint64_t or_big_divmod_u64(or_big *b, uint64_t d) {
    u128 rem = 0;
    for (int i = b->n - 1; i >= 0; i--) {
        u128 cur = (rem << 64) | b->w[i];
        b->w[i] = (uint64_t)(cur / d);
        rem = cur % d;
    }
    while (b->n > 1 && b->w[b->n - 1] == 0) b->n--;
    return (uint64_t)rem;
}
uint64_t or_big_mod_u64(const or_big *b, uint64_t d) {
    u128 rem = 0;
    for (int i = b->n - 1; i >= 0; i--) rem = ((rem << 64) | b->w[i]) % d;
    return (uint64_t)rem;
}
static int big_cmp(const or_big *a, const or_big *b) {
    int n = a->n > b->n ? a->n : b->n;
    for (int i = n - 1; i >= 0; i--) {
        uint64_t x = i < a->n ? a->w[i] : 0, y = i < b->n ? b->w[i] : 0;
        if (x != y) return x > y ? 1 : -1;
    }
    return 0;
}

static int bits_of(uint64_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

/* product of mods[i] for i in [lo, hi) except skip, reduced mod m */
static uint64_t prod_except_mod(const uint64_t *mods, int lo, int hi, int skip, uint64_t m) {
    uint64_t r = 1 % m;
    for (int i = lo; i < hi; i++)
        if (i != skip) r = or_mulmod(r, mods[i] % m, m);
    return r;
}

static uint64_t *u64a(size_t n) { return (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t)); }
static double *f64a(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static or_ctx *ctx_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn, const uint64_t *t,
                          int tn, int gamma, uint64_t plain, const uint64_t *r_given);

or_ctx *or_ctx_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn,
                      const uint64_t *t, int tn, int gamma, uint64_t plain) {
    return ctx_create(log_n, q, qn, p, pn, t, tn, gamma, plain, NULL);
}

/* The context of ciphertexts at `level` of `full` (the engine's level view):
 * Q_l = q_0..q_level, the HPS basis R_l = the level + 2 moduli that follow
 * q_level in the full context's [Q | R] table (the full context's own moduli,
 * Q_l R_l one contiguous range of it), every Q_l-dependent table recomputed,
 * and the relinearisation key of the full level read with its full shape. */
or_ctx *or_ctx_level(const or_ctx *full, int level) {
    if (level < 0 || level >= full->qn) return NULL;
    or_ctx *c = ctx_create(full->log_n, full->mod + full->oQ, level + 1, full->mod + full->oP, full->pn,
                           full->mod + full->oT, full->tn, full->gamma, full->plain,
                           full->mod + full->oQ + level + 1);
    if (c) {
        c->key_dnum = full->key_dnum;
        c->key_blocks = full->key_blocks;
    }
    return c;
}

static or_ctx *ctx_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn, const uint64_t *t,
                          int tn, int gamma, uint64_t plain, const uint64_t *r_given) {
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->log_n = log_n;
    c->n = (uint64_t)1 << log_n;
    c->qn = qn;
    c->pn = pn;
    c->tn = tn;
    c->rn = qn + 1; /* include/BFVcontext.cuh:57 r_num = q_num + 1 */
    c->gamma = gamma;
    c->dnum = (qn + pn - 1) / pn;
    c->ri_blocks = (pn + qn + gamma - 1) / gamma;
    c->key_dnum = c->dnum;
    c->key_blocks = c->ri_blocks;
    c->plain = plain;
    c->nmod = pn + qn + c->rn + tn + 1;
    c->oP = 0;
    c->oQ = pn;
    c->oR = pn + qn;
    c->oT = pn + qn + c->rn;
    c->oPlain = c->nmod - 1;
    c->mod = u64a(c->nmod);
    c->psi = u64a(c->nmod);
    for (int i = 0; i < pn; i++) c->mod[c->oP + i] = p[i];
    uint64_t qmin = (uint64_t)-1;
    for (int i = 0; i < qn; i++) {
        c->mod[c->oQ + i] = q[i];
        if (q[i] < qmin) qmin = q[i];
    }
    /* include/BFVcontext.cuh:57-58 */
    if (r_given) {
        memcpy(c->mod + c->oR, r_given, (size_t)c->rn * sizeof(uint64_t));
    } else if (or_primes_below(c->n, qmin, c->rn, c->mod + c->oR) != 0) {
        fprintf(stderr, "oracle: not enough R primes\n");
        free(c);
        return NULL;
    }
    for (int i = 0; i < tn; i++) c->mod[c->oT + i] = t[i];
    c->mod[c->oPlain] = plain;
    c->batching = ((plain - 1) % (2 * c->n)) == 0;

    /* psi = g^((m-1)/2N), g the smallest primitive root (include/BFVcontext.cuh:138-213) */
    size_t N = c->n;
    c->tab = u64a((size_t)c->nmod * 4 * N);
    for (int m = 0; m < c->nmod; m++) {
        if (m == c->oPlain && !c->batching) continue;
        c->psi[m] = or_mth_root(2 * N, c->mod[m]);
        uint64_t *base = c->tab + (size_t)m * 4 * N;
        or_fill_tables(c->mod[m], c->psi[m], log_n, base, base + N, base + 2 * N, base + 3 * N);
    }

    const uint64_t *Q = c->mod + c->oQ, *R = c->mod + c->oR, *T = c->mod + c->oT;
    int rn = c->rn;

    /* HPS base converters (include/BasisConv_QR.cuh:53-189) */
    c->q_hatinv = u64a(qn);
    c->q_inv = f64a(qn);
    c->q_hat_mod_r = u64a((size_t)rn * qn);
    c->alphaQ_mod_r = u64a((size_t)(qn + 1) * rn);
    for (int i = 0; i < qn; i++) {
        c->q_hatinv[i] = or_invmod(prod_except_mod(Q, 0, qn, i, Q[i]), Q[i]);
        c->q_inv[i] = 1.0 / (double)Q[i];
    }
    for (int j = 0; j < rn; j++) {
        for (int i = 0; i < qn; i++) c->q_hat_mod_r[j * qn + i] = prod_except_mod(Q, 0, qn, i, R[j]);
        uint64_t Qmod = prod_except_mod(Q, 0, qn, -1, R[j]);
        for (int a = 0; a <= qn; a++) c->alphaQ_mod_r[a * rn + j] = or_mulmod(a, Qmod, R[j]);
    }
    c->r_hatinv = u64a(rn);
    c->r_inv = f64a(rn);
    c->r_hat_mod_q = u64a((size_t)qn * rn);
    c->alphaR_mod_q = u64a((size_t)(rn + 1) * qn);
    for (int i = 0; i < rn; i++) {
        c->r_hatinv[i] = or_invmod(prod_except_mod(R, 0, rn, i, R[i]), R[i]);
        c->r_inv[i] = 1.0 / (double)R[i];
    }
    for (int j = 0; j < qn; j++) {
        for (int i = 0; i < rn; i++) c->r_hat_mod_q[j * rn + i] = prod_except_mod(R, 0, rn, i, Q[j]);
        uint64_t Rmod = prod_except_mod(R, 0, rn, -1, Q[j]);
        for (int a = 0; a <= rn; a++) c->alphaR_mod_q[a * qn + j] = or_mulmod(a, Rmod, Q[j]);
    }

    /* HPS scale-and-round QR -> R (include/BFVcontext.cuh:1268-1388) */
    c->sr_modr = u64a((size_t)rn * (qn + 1));
    c->sr_frac = f64a(qn);
    {
        int sn = qn + rn;
        const uint64_t *S = c->mod + c->oQ; /* Q then R are adjacent */
        or_big tR;
        or_big_set(&tR, plain);
        for (int j = 0; j < rn; j++) or_big_mul_u64(&tR, R[j]);
        for (int i = 0; i < sn; i++) {
            uint64_t shatinv = or_invmod(prod_except_mod(S, 0, sn, i, S[i]), S[i]);
            or_big x = tR;
            or_big_mul_u64(&x, shatinv);
            if (i < qn) c->sr_frac[i] = (double)or_big_mod_u64(&x, S[i]) / (double)S[i];
            or_big_divmod_u64(&x, S[i]);
            if (i < qn) {
                for (int j = 0; j < rn; j++) c->sr_modr[j * (qn + 1) + i] = or_big_mod_u64(&x, R[j]);
            } else {
                int j = i - qn;
                c->sr_modr[j * (qn + 1) + qn] = or_big_mod_u64(&x, R[j]);
            }
        }
    }

    /* HPS decrypt (include/BFVcontext.cuh:1024-1137) */
    {
        uint64_t qmax = 0;
        for (int i = 0; i < qn; i++)
            if (Q[i] > qmax) qmax = Q[i];
        c->qMSB = bits_of(qmax);
        c->sizeQMSB = bits_of((uint64_t)qn);
        c->tMSB = bits_of(plain);
        c->dec_modt = u64a(qn);
        c->dec_frac = f64a(qn);
        c->dec_B_modt = u64a(qn);
        c->dec_B_frac = f64a(qn);
        int hf = c->qMSB >> 1;
        for (int i = 0; i < qn; i++) {
            u128 x = (u128)plain * c->q_hatinv[i];
            c->dec_modt[i] = (uint64_t)((x / Q[i]) % plain);
            c->dec_frac[i] = (double)(uint64_t)(x % Q[i]) / (double)Q[i];
            uint64_t hinvB = (uint64_t)((((u128)c->q_hatinv[i]) << hf) % Q[i]);
            u128 y = (u128)plain * hinvB;
            c->dec_B_modt[i] = (uint64_t)((y / Q[i]) % plain);
            c->dec_B_frac[i] = (double)(uint64_t)(y % Q[i]) / (double)Q[i];
        }
    }

    /* BFV encrypt constants (include/BFVcontext.cuh:881-924) */
    c->negQ_mod_t = plain - prod_except_mod(Q, 0, qn, -1, plain);
    c->tinv_mod_q = u64a(qn);
    for (int i = 0; i < qn; i++) c->tinv_mod_q[i] = or_invmod(plain % Q[i], Q[i]);

    /* key switching tables at full level (include/BFVcontext.cuh:286-490) */
    {
        int pq = pn + qn, B = c->ri_blocks, D = c->dnum;
        const uint64_t *PQ = c->mod; /* P then Q adjacent */
        c->ri_hatinv = u64a((size_t)B * gamma);
        c->ri_hat_mod_t = u64a((size_t)B * tn * gamma);
        c->ri_mod_t = u64a((size_t)tn * B);
        for (int b = 0; b < B; b++) {
            int lo = b * gamma, hi = lo + gamma < pq ? lo + gamma : pq;
            for (int j = lo; j < hi; j++) {
                c->ri_hatinv[b * gamma + (j - lo)] =
                    or_invmod(prod_except_mod(PQ, lo, hi, j, PQ[j]), PQ[j]);
                for (int k = 0; k < tn; k++)
                    c->ri_hat_mod_t[(b * tn + k) * gamma + (j - lo)] =
                        prod_except_mod(PQ, lo, hi, j, T[k]);
            }
            for (int k = 0; k < tn; k++) c->ri_mod_t[k * B + b] = prod_except_mod(PQ, lo, hi, -1, T[k]);
        }
        c->qj_hatinv = u64a((size_t)D * pn);
        c->qj_hat_mod_t = u64a((size_t)D * tn * pn);
        c->qj_mod_t = u64a((size_t)D * tn);
        for (int b = 0; b < D; b++) {
            int lo = b * pn, hi = lo + pn < qn ? lo + pn : qn;
            for (int j = lo; j < hi; j++) {
                c->qj_hatinv[b * pn + (j - lo)] = or_invmod(prod_except_mod(Q, lo, hi, j, Q[j]), Q[j]);
                for (int k = 0; k < tn; k++)
                    c->qj_hat_mod_t[(b * tn + k) * pn + (j - lo)] = prod_except_mod(Q, lo, hi, j, T[k]);
            }
            for (int k = 0; k < tn; k++) c->qj_mod_t[b * tn + k] = prod_except_mod(Q, lo, hi, -1, T[k]);
        }
        c->t_hatinv = u64a(tn);
        c->t_hat_mod_pq = u64a((size_t)pq * tn);
        c->t_mod_pq = u64a(pq);
        for (int k = 0; k < tn; k++) c->t_hatinv[k] = or_invmod(prod_except_mod(T, 0, tn, k, T[k]), T[k]);
        for (int i = 0; i < pq; i++) {
            for (int k = 0; k < tn; k++) c->t_hat_mod_pq[i * tn + k] = prod_except_mod(T, 0, tn, k, PQ[i]);
            c->t_mod_pq[i] = prod_except_mod(T, 0, tn, -1, PQ[i]);
        }
        c->halfT_mod = u64a(pq + tn);
        or_big half;
        or_big_set(&half, 1);
        for (int k = 0; k < tn; k++) or_big_mul_u64(&half, T[k]);
        or_big_divmod_u64(&half, 2);
        for (int i = 0; i < pq + tn; i++) c->halfT_mod[i] = or_big_mod_u64(&half, c->mod[or_pqt_index(c, i)]);
        c->p_hat_mod_q = u64a((size_t)qn * pn);
        c->p_hatinv = u64a(pn);
        c->P_mod_q = u64a(qn);
        c->Pinv_mod_q = u64a(qn);
        const uint64_t *P = c->mod + c->oP;
        for (int l = 0; l < qn; l++)
            for (int i = 0; i < pn; i++) c->p_hat_mod_q[l * pn + i] = prod_except_mod(P, 0, pn, i, Q[l]);
        for (int i = 0; i < pn; i++) c->p_hatinv[i] = or_invmod(prod_except_mod(P, 0, pn, i, P[i]), P[i]);
        for (int i = 0; i < qn; i++) {
            c->P_mod_q[i] = prod_except_mod(P, 0, pn, -1, Q[i]);
            c->Pinv_mod_q[i] = or_invmod(c->P_mod_q[i], Q[i]);
        }

        /* Correctness bound of the T-basis inner product (not checked by the
         * reference, whose assert at include/BFVcontext.cuh:223 multiplies by an
         * empty product): T > 4 * dnum * N * max(Q_j) * max(R_b). */
        or_big lhs, Tb;
        or_big_set(&Tb, 1);
        for (int k = 0; k < tn; k++) or_big_mul_u64(&Tb, T[k]);
        or_big maxqj, maxrb;
        or_big_set(&maxqj, 0);
        or_big_set(&maxrb, 0);
        for (int b = 0; b < D; b++) {
            or_big x;
            or_big_set(&x, 1);
            for (int j = b * pn; j < b * pn + pn && j < qn; j++) or_big_mul_u64(&x, Q[j]);
            if (big_cmp(&x, &maxqj) > 0) maxqj = x;
        }
        for (int b = 0; b < B; b++) {
            or_big x;
            or_big_set(&x, 1);
            for (int j = b * gamma; j < b * gamma + gamma && j < pq; j++) or_big_mul_u64(&x, PQ[j]);
            if (big_cmp(&x, &maxrb) > 0) maxrb = x;
        }
        {
            or_big acc;
            memset(&acc, 0, sizeof(acc));
            acc.n = maxqj.n + maxrb.n + 1;
            for (int i = 0; i < maxqj.n; i++) {
                u128 carry = 0;
                for (int j = 0; j < maxrb.n; j++) {
                    u128 cur = (u128)maxqj.w[i] * maxrb.w[j] + acc.w[i + j] + carry;
                    acc.w[i + j] = (uint64_t)cur;
                    carry = cur >> 64;
                }
                acc.w[i + maxrb.n] += (uint64_t)carry;
            }
            while (acc.n > 1 && acc.w[acc.n - 1] == 0) acc.n--;
            lhs = acc;
        }
        or_big_mul_u64(&lhs, 4 * (uint64_t)D * N);
        if (big_cmp(&Tb, &lhs) <= 0) {
            fprintf(stderr, "oracle: T basis too small for dnum=%d gamma=%d\n", D, gamma);
            or_ctx_destroy(c);
            return NULL;
        }
    }
    return c;
}

void or_ctx_destroy(or_ctx *c) {
    if (!c) return;
    void *ptrs[] = {c->mod, c->psi, c->tab, c->q_hatinv, c->q_hat_mod_r, c->alphaQ_mod_r, c->q_inv,
                    c->r_hatinv, c->r_hat_mod_q, c->alphaR_mod_q, c->r_inv, c->sr_modr, c->sr_frac,
                    c->dec_modt, c->dec_frac, c->dec_B_modt, c->dec_B_frac, c->tinv_mod_q,
                    c->ri_hatinv, c->ri_hat_mod_t, c->ri_mod_t, c->qj_hatinv, c->qj_hat_mod_t,
                    c->qj_mod_t, c->t_hatinv, c->t_hat_mod_pq, c->t_mod_pq, c->halfT_mod,
                    c->p_hat_mod_q, c->p_hatinv, c->P_mod_q, c->Pinv_mod_q};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(c);
}

/* out: log_n, n, qn, pn, tn, rn, gamma, dnum, ri_blocks, batching, nmod,
 *      oP, oQ, oR, oT, oPlain */
int or_ctx_info(const or_ctx *c, int *out) {
    int v[16] = {c->log_n, (int)c->n, c->qn, c->pn, c->tn, c->rn, c->gamma, c->dnum,
                 c->ri_blocks, c->batching, c->nmod, c->oP, c->oQ, c->oR, c->oT, c->oPlain};
    memcpy(out, v, sizeof(v));
    return 16;
}
int or_ctx_mod_count(const or_ctx *c) { return c->nmod; }
void or_ctx_moduli(const or_ctx *c, uint64_t *out) { memcpy(out, c->mod, c->nmod * sizeof(uint64_t)); }
void or_ctx_psi(const or_ctx *c, uint64_t *out) { memcpy(out, c->psi, c->nmod * sizeof(uint64_t)); }
const uint64_t *or_ctx_table(const or_ctx *c, int m, int which) { return or_tab(c, m, which); }

static int put(uint64_t *out, int cap, const uint64_t *src, int n) {
    if (n > cap) return -n;
    memcpy(out, src, n * sizeof(uint64_t));
    return n;
}
int or_ctx_export(const or_ctx *c, const char *name, uint64_t *out, int cap) {
    int qn = c->qn, rn = c->rn, tn = c->tn, pq = c->pn + c->qn;
    if (!strcmp(name, "q_hatinv")) return put(out, cap, c->q_hatinv, qn);
    if (!strcmp(name, "q_hat_mod_r")) return put(out, cap, c->q_hat_mod_r, rn * qn);
    if (!strcmp(name, "alphaQ_mod_r")) return put(out, cap, c->alphaQ_mod_r, (qn + 1) * rn);
    if (!strcmp(name, "r_hatinv")) return put(out, cap, c->r_hatinv, rn);
    if (!strcmp(name, "r_hat_mod_q")) return put(out, cap, c->r_hat_mod_q, qn * rn);
    if (!strcmp(name, "alphaR_mod_q")) return put(out, cap, c->alphaR_mod_q, (rn + 1) * qn);
    if (!strcmp(name, "sr_modr")) return put(out, cap, c->sr_modr, rn * (qn + 1));
    if (!strcmp(name, "dec_modt")) return put(out, cap, c->dec_modt, qn);
    if (!strcmp(name, "dec_B_modt")) return put(out, cap, c->dec_B_modt, qn);
    if (!strcmp(name, "tinv_mod_q")) return put(out, cap, c->tinv_mod_q, qn);
    if (!strcmp(name, "negQ_mod_t")) return put(out, cap, &c->negQ_mod_t, 1);
    if (!strcmp(name, "ri_hatinv")) return put(out, cap, c->ri_hatinv, c->ri_blocks * c->gamma);
    if (!strcmp(name, "ri_hat_mod_t")) return put(out, cap, c->ri_hat_mod_t, c->ri_blocks * tn * c->gamma);
    if (!strcmp(name, "ri_mod_t")) return put(out, cap, c->ri_mod_t, tn * c->ri_blocks);
    if (!strcmp(name, "qj_hatinv")) return put(out, cap, c->qj_hatinv, c->dnum * c->pn);
    if (!strcmp(name, "qj_hat_mod_t")) return put(out, cap, c->qj_hat_mod_t, c->dnum * tn * c->pn);
    if (!strcmp(name, "qj_mod_t")) return put(out, cap, c->qj_mod_t, c->dnum * tn);
    if (!strcmp(name, "t_hatinv")) return put(out, cap, c->t_hatinv, tn);
    if (!strcmp(name, "t_hat_mod_pq")) return put(out, cap, c->t_hat_mod_pq, pq * tn);
    if (!strcmp(name, "t_mod_pq")) return put(out, cap, c->t_mod_pq, pq);
    if (!strcmp(name, "halfT_mod")) return put(out, cap, c->halfT_mod, pq + tn);
    if (!strcmp(name, "p_hat_mod_q")) return put(out, cap, c->p_hat_mod_q, qn * c->pn);
    if (!strcmp(name, "p_hatinv")) return put(out, cap, c->p_hatinv, c->pn);
    if (!strcmp(name, "P_mod_q")) return put(out, cap, c->P_mod_q, qn);
    if (!strcmp(name, "Pinv_mod_q")) return put(out, cap, c->Pinv_mod_q, qn);
    if (!strcmp(name, "dec_msb")) {
        uint64_t v[3] = {(uint64_t)c->qMSB, (uint64_t)c->sizeQMSB, (uint64_t)c->tMSB};
        return put(out, cap, v, 3);
    }
    return 0;
}
int or_ctx_export_f64(const or_ctx *c, const char *name, double *out, int cap) {
    const double *src = NULL;
    int n = 0;
    if (!strcmp(name, "q_inv")) src = c->q_inv, n = c->qn;
    else if (!strcmp(name, "r_inv")) src = c->r_inv, n = c->rn;
    else if (!strcmp(name, "sr_frac")) src = c->sr_frac, n = c->qn;
    else if (!strcmp(name, "dec_frac")) src = c->dec_frac, n = c->qn;
    else if (!strcmp(name, "dec_B_frac")) src = c->dec_B_frac, n = c->qn;
    if (!src) return 0;
    if (n > cap) return -n;
    memcpy(out, src, n * sizeof(double));
    return n;
}

void or_ntt_fwd_range(const or_ctx *c, uint64_t *a, int mod0, int limbs, int poly_num) {
    for (int y = 0; y < poly_num; y++)
        for (int i = 0; i < limbs; i++)
            or_ntt_fwd(a + ((size_t)y * limbs + i) * c->n, c->log_n, c->mod[mod0 + i],
                       or_tab(c, mod0 + i, 0));
}
void or_ntt_inv_range(const or_ctx *c, uint64_t *a, int mod0, int limbs, int poly_num) {
    for (int y = 0; y < poly_num; y++)
        for (int i = 0; i < limbs; i++)
            or_ntt_inv(a + ((size_t)y * limbs + i) * c->n, c->log_n, c->mod[mod0 + i],
                       or_tab(c, mod0 + i, 1));
}
