// Internal launchers shared by the translation units (host side).
#pragma once
#include "kbfv_internal.h"

// rns.hip
int kb_poly_binop(const kbfv_context *c, int op, const uint64_t *a, const uint64_t *b, uint64_t *o,
                  int mod0, int limbs, int polys, hipStream_t st);
// noscale: x is N times the canonical input (kb_intt_noscale output, Q -> R
// only); the n^-1 rides in the first constant
int kb_bconv(const kbfv_context *c, bool q_to_r, const uint64_t *x, long in_stride, uint64_t *y,
             long out_stride, int polys, hipStream_t st, bool noscale = false);
int kb_tensor_qr(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, const uint64_t *r1,
                 const uint64_t *r2, uint64_t *out, int batch, hipStream_t st);
int kb_tensor(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out, int mod0,
              int limbs, int batch, hipStream_t st);
int kb_scale_round(const kbfv_context *c, const uint64_t *x, uint64_t *y, int polys, hipStream_t st);
int kb_hps_decrypt(const kbfv_context *c, const uint64_t *x, uint64_t *m, int batch, hipStream_t st);
int kb_delta_m(const kbfv_context *c, const uint64_t *m, uint64_t *out, int batch, hipStream_t st);
int kb_reduce_plain(const kbfv_context *c, uint64_t *v, long count, hipStream_t st);

// rns_mfma.hip: the same base conversion and scale-and-round on the int8 matrix
// cores; return 1 (nothing launched) when the bases are too large for them
bool kb_rns_gemm_fits(int k_rows, int out_n);
int kb_bconv_mfma(const kbfv_context *c, bool q_to_r, const uint64_t *x, long in_stride, uint64_t *y,
                  long out_stride, int polys, hipStream_t st, bool noscale = false);
int kb_scale_round_mfma(const kbfv_context *c, const uint64_t *x, uint64_t *y, int polys, hipStream_t st);

// ks.hip
int kb_modup_qj_to_t(const kbfv_context *c, const uint64_t *x, long xs, uint64_t *out, int batch,
                     hipStream_t st);
int kb_modup_pq_to_t(const kbfv_context *c, const uint64_t *x, uint64_t *out, int batch, hipStream_t st);
int kb_inner_product(const kbfv_context *c, const uint64_t *cin, const uint64_t *key, uint64_t *out,
                     int batch, hipStream_t st);
// ks_mfma.hip: the inner product on the int8 matrix cores (returns 1 when the
// key shape does not fit it)
int kb_inner_product_mfma(const kbfv_context *c, const uint64_t *cin, const uint64_t *key, uint64_t *out,
                          int batch, hipStream_t st);
int kb_modup_t_to_pq(const kbfv_context *c, const uint64_t *in, uint64_t *out, int batch, hipStream_t st);
int kb_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
               long add_stride, int batch, hipStream_t st);
// kb_modup_t_to_pq then kb_moddown in one kernel, without the PQ intermediate,
// for the shapes kb_modup_moddown_fused accepts; other shapes run the two
// kernels through pq_tmp ([batch][2][pn+qn][N], unused and may be null when fused)
bool kb_modup_moddown_fused(const kbfv_context *c);
int kb_modup_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
                     long add_stride, int batch, uint64_t *pq_tmp, hipStream_t st);
// the key-switch tail: inner-product output ip [batch][2][ri_blocks][tn][N]
// (NTT over T, left as it is) -> INTT over T into ip_coeff (same shape) ->
// modUp T->PQ and modDown into out [batch][2][qn][N] (+ add).  For the fused
// shapes the INTT leaves out n^-1 and the modUp folds it into its constants.
int kb_ks_tail(const kbfv_context *c, const uint64_t *ip, uint64_t *ip_coeff, uint64_t *out, const uint64_t *add,
               long add_stride, int batch, uint64_t *pq_tmp, hipStream_t st);
// digit decomposition of the Q limbs into T followed by the forward NTT over T
int kb_digits_to_t_ntt(const kbfv_context *c, const uint64_t *x_coeff, long x_stride, uint64_t *out,
                       int batch, uint64_t *tmp, hipStream_t st);

// rng.hip
struct kb_stream_key {
    uint32_t sub[8];
    uint32_t n8[2];
};
kb_stream_key kb_stream_setup(const uint8_t seed[32], uint64_t tag);
// kind: 0 uniform, 1 ternary, 2 zero-one, 3 CBD.  out [batch][limbs][N]; batch entry b reads
// stream entry entry0 + b (XSalsa20 nonce bytes 16..23).
// ostride: elements between batch entries (<= 0: limbs*N, contiguous)
int kb_sample(const kbfv_context *c, int kind, const uint8_t seed[32], uint64_t tag, uint64_t entry0,
              int mod0, int limbs, uint64_t *out, long ostride, int batch, hipStream_t st);
int kb_xsalsa20(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0, uint8_t *out, size_t len,
                hipStream_t st);

// extprod.hip
// forward NTT of the gadget digit polynomials of rows r0..r0+rc-1 with the
// digit split fused into the first pass (digits must stay below 8 q_min)
int kb_digit_ntt(const kbfv_context *c, const uint64_t *coef, uint64_t *dig, int batch, int bits, int nd, int r0,
                 int rc, hipStream_t st);
int kb_digit_decompose(const kbfv_context *c, const uint64_t *coef, uint64_t *dig, int batch, int base_bits,
                       int nd, int r0, int rc, hipStream_t st);
// out [batch][2][qn][N] += sum over rc rows of dig [batch][rc][qn][N] * rgsw_rows [rc][2][qn][N]
int kb_gadget_mac(const kbfv_context *c, const uint64_t *dig, const uint64_t *rgsw_rows, uint64_t *out,
                  int batch, int rc, hipStream_t st);
int kb_rgsw_add_gadget(const kbfv_context *c, uint64_t *rgsw, const uint64_t *m_ntt, int base_bits, int nd,
                       hipStream_t st);
