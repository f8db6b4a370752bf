/*
 * CPU oracle for the kunpBFV hot path: batched BFV HomMult + relinearisation.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path
 * (kunpbfv_amd/) never links or calls it.
 *
 * Plain C11 restatement of the reference's algorithms (QiiPatri/kunpBFV,
 * header-only CUDA C++).  Every function cites the reference file:line it
 * follows.  Integer arithmetic is done with unsigned __int128 and canonical
 * reductions, so any correct implementation of an integer step gives the same
 * bits.  The reference's floating-point estimates (HPS alpha/nu, the float
 * overflow counts of the key-switch conversions, HPS decrypt) are restated
 * operation for operation, with the multiply-add contractions written as
 * explicit fma() (the reference is built by nvcc with --fmad=true, which
 * contracts `acc += x * y` into one fused multiply-add).
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed in the header comment of each .c file and in tests/.
 */
#ifndef KBFV_ORACLE_H
#define KBFV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 u128;

/* ---------------- scalar number theory (nt.c) ---------------- */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q);
uint64_t or_invmod(uint64_t a, uint64_t q);
uint64_t or_shoup(uint64_t w, uint64_t q);
int or_is_prime(uint64_t n);
uint64_t or_primitive_root(uint64_t q);
uint64_t or_mth_root(uint64_t m, uint64_t q);
int or_primes_below(uint64_t ntt_size, uint64_t upper_bound, int count, uint64_t *out);
int or_primes_descending(uint64_t below, uint64_t step, int count, uint64_t *out);
uint64_t or_bitrev(uint64_t x, int bits);
void or_fill_tables(uint64_t q, uint64_t psi, int log_n, uint64_t *psi_tab,
                    uint64_t *psiinv_tab, uint64_t *psi_shoup, uint64_t *psiinv_shoup);
uint64_t or_barrett_single(uint64_t lo, uint64_t hi, uint64_t q);

/* ---------------- negacyclic NTT (ntt.c) ---------------- */
void or_ntt_fwd(uint64_t *a, int log_n, uint64_t q, const uint64_t *psi_tab);
void or_ntt_inv(uint64_t *a, int log_n, uint64_t q, const uint64_t *psiinv_tab);
void or_negacyclic_mul_schoolbook(const uint64_t *a, const uint64_t *b, uint64_t *c,
                                  int n, uint64_t q);

/* ---------------- context (ctx.c) ---------------- */
typedef struct or_ctx or_ctx;

/* Moduli are stored in one table in the order [P | Q | R | T] followed by the
 * plain modulus t (index or_ctx_mod_count()-1). */
/* the context of ciphertexts at `level` of a full context (its own, owned by the caller) */
or_ctx *or_ctx_level(const or_ctx *full, int level);
or_ctx *or_ctx_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn,
                      const uint64_t *t, int tn, int gamma, uint64_t plain);
void or_ctx_destroy(or_ctx *c);
int or_ctx_info(const or_ctx *c, int *out /* 16 ints */);
int or_ctx_mod_count(const or_ctx *c);
void or_ctx_moduli(const or_ctx *c, uint64_t *out);
void or_ctx_psi(const or_ctx *c, uint64_t *out);
const uint64_t *or_ctx_table(const or_ctx *c, int mod_idx, int which);
int or_ctx_export(const or_ctx *c, const char *name, uint64_t *out, int cap);
int or_ctx_export_f64(const or_ctx *c, const char *name, double *out, int cap);

/* limb-range NTT on [poly_num][limbs][N] with modulus indices mod0.. */
void or_ntt_fwd_range(const or_ctx *c, uint64_t *a, int mod0, int limbs, int poly_num);
void or_ntt_inv_range(const or_ctx *c, uint64_t *a, int mod0, int limbs, int poly_num);

/* ---------------- HPS tensoring (hps.c) ---------------- */
void or_bconv_q_to_r(const or_ctx *c, const uint64_t *x_q, uint64_t *y_r);
void or_bconv_r_to_q(const or_ctx *c, const uint64_t *x_r, uint64_t *y_q);
void or_tensor(const or_ctx *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out, int limbs,
               int mod0);
void or_scale_round(const or_ctx *c, const uint64_t *x_qr, uint64_t *y_r);
void or_hps_decrypt(const or_ctx *c, const uint64_t *x_q, uint64_t *m);
void or_add_delta_m(const or_ctx *c, const uint64_t *m, uint64_t *out_q);

/* ---------------- key switching, KLSS "23" method (ks.c) ---------------- */
void or_modup_qj_to_t(const or_ctx *c, const uint64_t *x_q, uint64_t *out_t);
void or_modup_pq_to_t(const or_ctx *c, const uint64_t *x_pq, uint64_t *out_t);
void or_inner_product(const or_ctx *c, const uint64_t *cin_t, const uint64_t *key_t,
                      uint64_t *out_t);
void or_modup_t_to_pq(const or_ctx *c, const uint64_t *in_t, uint64_t *out_pq);
void or_moddown(const or_ctx *c, const uint64_t *in_pq, uint64_t *out_q);

/* ---------------- XSalsa20 and samplers (xsalsa20.c) ---------------- */
void or_salsa20_core(const uint32_t in[16], uint32_t out[16], int add_input);
void or_hsalsa20(const uint8_t key[32], const uint8_t nonce16[16], uint8_t subkey[32]);
void or_xsalsa20_stream(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0,
                        uint8_t *out, size_t len);
void or_sample_uniform(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                   int mod0, int limbs, uint64_t *out);
void or_sample_ternary(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                   int mod0, int limbs, uint64_t *out);
void or_sample_cbd(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                   int mod0, int limbs, uint64_t *out);
void or_sample_zo(const or_ctx *c, const uint8_t seed[32], uint64_t tag, uint64_t entry,
                   int mod0, int limbs, uint64_t *out);

/* sampler tags: (purpose << 32) | index */
#define OR_TAG(purpose, index) (((uint64_t)(purpose) << 32) | (uint64_t)(index))
enum {
    OR_TAG_SK = 1,
    OR_TAG_PK_A = 2,
    OR_TAG_PK_E = 3,
    OR_TAG_RLK_A = 4,
    OR_TAG_RLK_E = 5,
    OR_TAG_ENC_V = 6,
    OR_TAG_ENC_E0 = 7,
    OR_TAG_ENC_E1 = 8,
    OR_TAG_RGSW_A = 9,
    OR_TAG_RGSW_E = 10,
    OR_TAG_GAL_A = 11,
    OR_TAG_GAL_E = 12,
    OR_TAG_SWK_A = 13,
    OR_TAG_SWK_E = 14
};

/* ---------------- BFV scheme (bfv.c) ---------------- */
void or_keygen_sk(const or_ctx *c, const uint8_t seed[32], uint64_t *sk_pq);
void or_keygen_pk(const or_ctx *c, const uint8_t seed[32], const uint64_t *sk_pq, uint64_t *pk);
void or_keygen_rlk(const or_ctx *c, const uint8_t seed[32], const uint64_t *sk_pq,
                   uint64_t *rlk_t);
/* key-switching key from src_q ([qn][N], NTT over Q) to the secret, layout of
 * the relinearisation key (the reference's addSWKey_23 is an unfilled stub,
 * include/KeySwitch_23.cuh:6-16; this is addMultKey_23 with s^2 -> src) */
void or_keygen_swk(const or_ctx *c, const uint8_t seed[32], const uint64_t *sk_pq, const uint64_t *src_q,
                   uint64_t *key_t);
void or_encrypt(const or_ctx *c, const uint8_t seed[32], uint64_t entry, const uint64_t *pk,
                const uint64_t *m, uint64_t *ct);
void or_decrypt(const or_ctx *c, const uint64_t *sk_pq, const uint64_t *ct, uint64_t *m);
/* modulus switching from_level -> to_level (include/BFVScheme.cuh:469-487):
 * in [2][from_level + 1][N], out [2][to_level + 1][N], NTT form */
void or_mod_switch(const or_ctx *c, const uint64_t *in, int from_level, uint64_t *out, int to_level);
void or_mult_relin(const or_ctx *c, const uint64_t *rlk_t, const uint64_t *ct1,
                   const uint64_t *ct2, uint64_t *out);
void or_mult_tensor(const or_ctx *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *d012);
void or_relin(const or_ctx *c, const uint64_t *rlk_t, const uint64_t *d2_coeff, uint64_t *ks_out);
void or_add(const or_ctx *c, const uint64_t *a, const uint64_t *b, uint64_t *out, int limbs,
            int mod0, int poly_num);
/* ciphertext arithmetic over Q, NTT form (include/BFVScheme.cuh:192-326):
 * out = -ct; out = c * ct (+ add); out = ct + (0, NTT(Delta m)) */
void or_negate(const or_ctx *c, const uint64_t *ct, uint64_t *out);
void or_mul_scalar(const or_ctx *c, const uint64_t *ct, uint64_t k, const uint64_t *add, uint64_t *out);
void or_add_plain(const or_ctx *c, const uint64_t *ct, const uint64_t *m, uint64_t *out);
void or_encode(const or_ctx *c, const uint64_t *slots, uint64_t *m);
void or_decode(const or_ctx *c, const uint64_t *m, uint64_t *slots);
/* SEAL slot order (BFVContext::preComputeIndex, include/BFVcontext.cuh:102-129):
 * slot i of row 0 sits at NTT index brv((5^i mod 2N - 1) / 2), slot i of row 1
 * at brv((2N - 5^i mod 2N - 1) / 2); requires batching (t = 1 mod 2N). */
void or_seal_index_map(const or_ctx *c, uint64_t *map);
void or_encode_seal(const or_ctx *c, const uint64_t *slots, uint64_t *m);
void or_decode_seal(const or_ctx *c, const uint64_t *m, uint64_t *slots);

/* Galois automorphisms X -> X^g (g odd, mod 2N) and slot rotation.
 * or_galois_elt: 5^steps mod 2N (the reference's rotGroups, include/BFVcontext.cuh:788-796);
 * steps < 0 gives the conjugation element 2N - 1.
 * or_automorphism: coefficient form, polys x limbs starting at mod0, canonical output
 *   (include/poly_arithmetic.cuh:496-512 sk_and_poly_LeftRot_kernel).
 * or_keygen_galois: key-switching key in T, layout of the relinearisation key, that
 *   switches from sigma_g^-1(s) to s (include/KeySwitch_23.cuh:120-166).
 * or_rotate: ct (NTT over Q) -> sigma_g(KS(a) + (0, b)), NTT over Q
 *   (include/KeySwitch_23.cuh:544-578). */
uint64_t or_galois_elt(const or_ctx *c, long steps);
void or_automorphism(const or_ctx *c, uint64_t g, const uint64_t *in, uint64_t *out, int mod0,
                     int limbs, int polys);
void or_keygen_galois(const or_ctx *c, const uint8_t seed[32], const uint64_t *sk_pq, uint64_t g,
                      uint64_t *key_t);
void or_rotate(const or_ctx *c, const uint64_t *key_t, uint64_t g, const uint64_t *ct, uint64_t *out);

/* RGSW x RLWE external product, gadget base 2^base_bits over each Q limb */
int or_rgsw_rows(const or_ctx *c, int base_bits);
void or_rgsw_encrypt(const or_ctx *c, const uint8_t seed[32], const uint64_t *sk_pq,
                     const uint64_t *m_coeff, int base_bits, uint64_t *rgsw);
void or_external_product(const or_ctx *c, const uint64_t *rgsw, const uint64_t *ct, int base_bits,
                         uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
