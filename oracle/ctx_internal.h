/*
 * Oracle context layout (internal).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef KBFV_ORACLE_CTX_INTERNAL_H
#define KBFV_ORACLE_CTX_INTERNAL_H

#include "oracle.h"

#define OR_MAXW 160 /* bigint words: 10240 bits */

/* Planted mistakes for the tests that show the pins reject them
 * (tests/test_oracle_mutations.py builds the sources with -DOR_MUTATION=n);
 * 0 in every real build. */
#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif

typedef struct {
    int n;
    uint64_t w[OR_MAXW];
} or_big;

struct or_ctx {
    int log_n;
    uint64_t n;
    int qn, pn, tn, rn, gamma;
    int dnum;       /* digits: ceil(qn / pn) */
    int ri_blocks;  /* ceil((pn + qn) / gamma) */
    /* shape of the relinearisation key read by the inner product,
     * [2][key_dnum][key_blocks][tn][N]: the context's own dnum / ri_blocks, or
     * those of the full level for a level context (or_ctx_level) */
    int key_dnum, key_blocks;
    uint64_t plain; /* t */
    int batching;   /* plain has a 2N-th root of unity */
    int nmod;       /* pn + qn + rn + tn + 1 */
    int oP, oQ, oR, oT, oPlain;
    uint64_t *mod;   /* [nmod] */
    uint64_t *psi;   /* [nmod] */
    uint64_t *tab;   /* [nmod][4][N]: psi, psiinv, psi_shoup, psiinv_shoup */

    /* HPS Q->R and R->Q base converters (include/BasisConv_QR.cuh:53-189) */
    uint64_t *q_hatinv;     /* [qn]  (Q/q_i)^-1 mod q_i */
    uint64_t *q_hat_mod_r;  /* [rn][qn] (Q/q_i) mod r_j */
    uint64_t *alphaQ_mod_r; /* [qn+1][rn] a*Q mod r_j */
    double *q_inv;          /* [qn] 1.0/q_i */
    uint64_t *r_hatinv;     /* [rn] */
    uint64_t *r_hat_mod_q;  /* [qn][rn] */
    uint64_t *alphaR_mod_q; /* [rn+1][qn] */
    double *r_inv;          /* [rn] */

    /* HPS scale-and-round QR -> R (include/BFVcontext.cuh:1268-1388) */
    uint64_t *sr_modr; /* [rn][qn+1] */
    double *sr_frac;   /* [qn] */

    /* HPS decrypt (include/BFVcontext.cuh:1024-1137) */
    int qMSB, sizeQMSB, tMSB;
    uint64_t *dec_modt;   /* [qn] */
    double *dec_frac;     /* [qn] */
    uint64_t *dec_B_modt; /* [qn] */
    double *dec_B_frac;   /* [qn] */

    /* BFV encrypt (include/BFVcontext.cuh:881-924) */
    uint64_t negQ_mod_t;
    uint64_t *tinv_mod_q; /* [qn] */

    /* key switching, full level (include/BFVcontext.cuh:286-490) */
    uint64_t *ri_hatinv;   /* [ri_blocks][gamma] */
    uint64_t *ri_hat_mod_t;/* [ri_blocks][tn][gamma] */
    uint64_t *ri_mod_t;    /* [tn][ri_blocks] */
    uint64_t *qj_hatinv;   /* [dnum][pn] */
    uint64_t *qj_hat_mod_t;/* [dnum][tn][pn] */
    uint64_t *qj_mod_t;    /* [dnum][tn] */
    uint64_t *t_hatinv;    /* [tn] */
    uint64_t *t_hat_mod_pq;/* [pn+qn][tn] */
    uint64_t *t_mod_pq;    /* [pn+qn] */
    uint64_t *halfT_mod;   /* [pn+qn+tn] indexed like pqt: P, Q, T */
    uint64_t *p_hat_mod_q; /* [qn][pn] */
    uint64_t *p_hatinv;    /* [pn] */
    uint64_t *P_mod_q;     /* [qn] */
    uint64_t *Pinv_mod_q;  /* [qn] */
};

/* index of the i-th modulus of the reference's "pqt" order [P | Q | T] */
static inline int or_pqt_index(const or_ctx *c, int i) {
    return i < c->pn + c->qn ? i : c->oT + (i - c->pn - c->qn);
}
static inline const uint64_t *or_tab(const or_ctx *c, int m, int which) {
    return c->tab + ((size_t)m * 4 + which) * c->n;
}

void or_big_set(or_big *b, uint64_t v);
void or_big_mul_u64(or_big *b, uint64_t v);
uint64_t or_big_divmod_u64(or_big *b, uint64_t d); /* b /= d, returns b % d */
uint64_t or_big_mod_u64(const or_big *b, uint64_t d);

#endif
