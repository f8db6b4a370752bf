/*
 * Oracle: negacyclic NTT / INTT over one 64-bit prime.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows the transform the reference computes with its two-kernel radix-8
 * implementation (include/ntt_60bit.cuh:289-823, dispatch at
 * include/BFVcontext.cuh:1513-1552): Cooley-Tukey forward with the bit-reversed
 * psi table (stage with m groups uses psi[m + group], include/ntt_60bit.cuh:315
 * ntt8 / :414 kernel1 tw_idx = 1), Gentleman-Sande inverse with the bit-reversed
 * psi^-1 table, n^-1 folded into psiinv[1] for the difference half of the last
 * stage and applied by a separate multiply to the sum half
 * (include/ntt_60bit.cuh:812-821).  Outputs are canonical in [0, q): index j of
 * the forward transform holds a(psi^(2*brv(j)+1)).
 *
 * Pinned by (tests/test_oracle_ntt.py): brute-force evaluation of
 * a(psi^(2 brv(j)+1)) at sampled j, INTT(NTT(a)) == a, and pointwise product
 * == schoolbook negacyclic product (the reference's refPolyMul128,
 * include/Utils.cuh:236).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void or_ntt_fwd(uint64_t *a, int log_n, uint64_t q, const uint64_t *psi_tab) {
    size_t n = (size_t)1 << log_n;
    for (size_t m = 1; m < n; m <<= 1) {
        size_t t = n / (2 * m);
        for (size_t i = 0; i < m; i++) {
            uint64_t w = psi_tab[m + i];
            size_t j0 = 2 * i * t;
            for (size_t j = j0; j < j0 + t; j++) {
                uint64_t u = a[j];
                uint64_t v = or_mulmod(a[j + t], w, q);
                uint64_t s = u + v;
                if (s >= q) s -= q;
                uint64_t d = u >= v ? u - v : u + q - v;
                a[j] = s;
                a[j + t] = d;
            }
        }
    }
}

void or_ntt_inv(uint64_t *a, int log_n, uint64_t q, const uint64_t *psiinv_tab) {
    size_t n = (size_t)1 << log_n;
    uint64_t ninv = or_invmod(n % q, q);
    for (size_t m = n >> 1; m >= 1; m >>= 1) {
        size_t t = n / (2 * m);
        for (size_t i = 0; i < m; i++) {
            uint64_t w = psiinv_tab[m + i];
            size_t j0 = 2 * i * t;
            for (size_t j = j0; j < j0 + t; j++) {
                uint64_t u = a[j], v = a[j + t];
                uint64_t s = u + v;
                if (s >= q) s -= q;
                uint64_t d = u >= v ? u - v : u + q - v;
                if (m == 1) s = or_mulmod(s, ninv, q); /* sum half of the last stage */
                a[j] = s;
                a[j + t] = or_mulmod(d, w, q);
            }
        }
        if (m == 1) break;
    }
}

/* include/Utils.cuh:236 refPolyMul128: c = a*b in Z_q[X]/(X^n+1) */
void or_negacyclic_mul_schoolbook(const uint64_t *a, const uint64_t *b, uint64_t *c, int n,
                                  uint64_t q) {
    uint64_t *acc = (uint64_t *)calloc(2 * (size_t)n, sizeof(uint64_t));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            acc[i + j] = (uint64_t)(((u128)a[i] * b[j] + acc[i + j]) % q);
        }
    for (int i = 0; i < n; i++) {
        uint64_t lo = acc[i], hi = acc[i + n];
        c[i] = lo >= hi ? lo - hi : lo + q - hi;
    }
    free(acc);
}
