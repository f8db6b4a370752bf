/*
 * kunpbfv -- C ABI of the MI355X-native BFV engine (libkunpbfv.so).
 *
 * Every entry point takes plain pointers, sizes and a hipStream_t passed as
 * void*.  Device buffers are owned by the caller; the library owns only its
 * context tables.  No entry point allocates device memory except context
 * creation; multi-kernel operations take a caller workspace whose size is
 * reported by the matching *_workspace_bytes() call.  Status: 0 on success,
 * negative on error (kbfv_last_error() has the message).  No exception crosses
 * this boundary.
 *
 * Data layout in device memory (uint64 residues, canonical in [0, q)):
 *   polynomial over a basis     [limbs][N], limb i at +i*N
 *   ciphertext                  [2][qn][N]   = (a | b), NTT domain, decrypts as b + a*s
 *   batch of ciphertexts        [batch][2][qn][N]
 *   secret key                  [pn+qn][N]   NTT domain over P then Q
 *   public key                  [2][qn][N]   (a | b = e - a*s), NTT domain
 *   relinearisation key         [2][dnum][ri_blocks][tn][N], NTT domain over T
 *   plaintext (encoded)         [N] coefficients mod t
 * Modulus indices (kbfv_context_info) run over one table [P | Q | R | T | t].
 *
 * Reference interfaces replaced (QiiPatri/kunpBFV, header-only CUDA C++; it has
 * no C ABI, so each entry point names the C++ member or kernel launcher whose
 * job it takes over):
 */
#ifndef KUNPBFV_H
#define KUNPBFV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KBFV_OK 0
#define KBFV_EINVAL (-1)
#define KBFV_EHIP (-2)
#define KBFV_ENOMEM (-3)
#define KBFV_ESTATE (-4)

typedef struct kbfv_context kbfv_context;

/* Message of the last failed call on this thread. */
const char *kbfv_last_error(void);

/* ---------------------------------------------------------------- parameters */

/* 60-bit chain: the largest primes below 2^60 that are 1 mod 2^17, taken in
 * descending order as T (tn), then P (pn), then Q (qn).  With tn = 4 the T part
 * equals the reference's tVec (include/Params.cuh:129-140). */
int kbfv_chain_60bit(int qn, int pn, int tn, uint64_t *q, uint64_t *p, uint64_t *t);

/* Primes below `upper` stepping by 2N (replaces get_primes_below,
 * include/uintarith_bfv.h:983); used for the R basis. */
int kbfv_primes_below(uint64_t ntt_size, uint64_t upper, int count, uint64_t *out);

/* 32 bytes from the operating system's entropy source (getrandom), for key
 * and encryption seeds (replaces the reference's std::random_device draws,
 * include/Utils.cuh:133-134, include/RNG.cuh:27-28).  Every seeded entry point
 * below is deterministic in its seed; this is where fresh seeds come from. */
int kbfv_random_seed(uint8_t out[32]);

/* Default parameter set of a ring degree (replaces BFVContext::getPrimeBFV,
 * include/Params.cuh:7-148): fills at most 64 entries of q/p/t. */
int kbfv_default_params(int log_n, uint64_t *q, int *qn, uint64_t *p, int *pn, uint64_t *t,
                        int *tn, int *gamma, uint64_t *plain);

/* ---------------------------------------------------------------- context */

/* Replaces BFVContext::BFVContext + preComputeOnCPU + copyMemoryToGPU
 * (include/BFVcontext.cuh:13-101, 130-542, 545-1395).  Tables are per context
 * (kernel arguments, no __constant__ globals), so contexts of different N
 * coexist.  Fails with KBFV_EINVAL if the T basis cannot hold the key-switch
 * inner product (T <= 4 * dnum * N * max Q_j * max R_b). */
int kbfv_context_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn,
                        const uint64_t *t, int tn, int gamma, uint64_t plain, kbfv_context **out);
int kbfv_context_destroy(kbfv_context *ctx);
/* out[16] = log_n, n, qn, pn, tn, rn, gamma, dnum, ri_blocks, batching, nmod,
 *           oP, oQ, oR, oT, oPlain */
int kbfv_context_info(const kbfv_context *ctx, int *out);
int kbfv_context_moduli(const kbfv_context *ctx, uint64_t *out);
/* host copy of a named precomputed table (parity tests); returns count */
int kbfv_context_export(const kbfv_context *ctx, const char *name, uint64_t *out, int cap);
int kbfv_context_export_f64(const kbfv_context *ctx, const char *name, double *out, int cap);

/* ---------------------------------------------------------------- NTT */

/* Replaces BFVContext::ToNTTInplace / FromNTTInplace and their _for_QR,
 * _for_externalProduct and _for_BFV twins (include/BFVcontext.cuh:1513-1677).
 * dir = +1 forward, -1 inverse.  Element (b, y, l) lives at
 *   src + b*s_batch + y*s_poly + l*s_limb  (in uint64 units; same for dst)
 * and limb l uses modulus mod0 + l.  src may equal dst. */
int kbfv_ntt(const kbfv_context *ctx, int dir, const uint64_t *src, uint64_t *dst, int mod0,
             int limbs, int polys, int batch, int64_t s_batch, int64_t s_poly, int64_t s_limb,
             int64_t d_batch, int64_t d_poly, int64_t d_limb, void *stream);

/* NTT plans and the autotuner (no counterpart in the reference, whose NTT has
 * one fixed radix-8 shape, include/ntt_60bit.cuh:414-823).  A transform runs as
 * a column pass of 2^c-point sub-transforms and a row pass of 2^(logN-c)-point
 * ones.  Plan word: bits 0-3 = c, bits 4-7 / 8-11 = log2 of the column / row
 * pass tile (elements per workgroup, 11..13), bit 12 / 13 = XCD-ordered grid
 * for the column / row pass, bits 14-16 = chunk k (0: each pass over the whole
 * launch; 2..6: both passes run per chunk of at most 2^(k+23) bytes, so the
 * second pass finds the first pass's output in the last-level cache); 0 = the
 * built-in plan.  kbfv_ntt_plan_candidates lists the unchunked shapes; any of
 * them with a chunk k is valid too.  Every plan computes the same bits.  A context keeps one plan per direction and bucket of
 * limbs*polys*batch (<64, <512, <4096, more); kbfv_ntt uses it. */
#define KBFV_NTT_BUCKETS 4
int kbfv_ntt_plan_valid(int log_n, uint32_t plan);                  /* 0 if valid; no device */
int kbfv_ntt_plan_candidates(int log_n, uint32_t *out, int cap);    /* returns the count */
int kbfv_ntt_plan_bucket(int64_t transforms);                       /* bucket of a transform count */
int kbfv_ntt_plan_get(const kbfv_context *ctx, int dir, int bucket, uint32_t *plan);
int kbfv_ntt_plan_set(kbfv_context *ctx, int dir, int bucket, uint32_t plan);
/* Times every pass variant (split, tile, grid order) of direction dir over
 * `polys` contiguous [polys][limbs][N] polynomials in `scratch` (overwritten,
 * polys*limbs*N words) on `stream`, `reps` launches each, and stores the
 * fastest plan for the bucket of limbs*polys.  best/best_ms may be null.  Run
 * it before other work on the context: it writes the plan table other calls read. */
int kbfv_ntt_autotune(kbfv_context *ctx, int dir, int mod0, int limbs, int polys, uint64_t *scratch,
                      int reps, void *stream, uint32_t *best, float *best_ms);

/* Ragged batches: polynomials of different RNS lengths (ciphertexts at uneven
 * levels, keys over P+Q next to ciphertexts over Q) transformed in one launch
 * per pass.  Replaces a host loop of ToNTTInplace / FromNTTInplace calls, one
 * per polynomial and level (include/BFVcontext.cuh:1513-1552).  Descriptor i is
 * six int64 words {src, dst, s_limb, d_limb, mod0, limbs}: limb l (modulus
 * mod0 + l) at src + l*s_limb, written to dst + l*d_limb (src may equal dst;
 * limbs may be 0).  kbfv_ntt_desc_upload validates `count` host descriptors
 * against the context and copies them with their limb offsets (count + 1
 * int32) to caller device buffers, returning the total limb count;
 * kbfv_ntt_desc then reads only device memory, so it can be captured in a HIP
 * graph and replayed while the descriptors stay valid; pass it the count and
 * the total the upload returned (limbs past the last descriptor are skipped).
 * Built-in plan. */
int kbfv_ntt_desc_upload(const kbfv_context *ctx, const int64_t *host_desc, int count, int64_t *dev_desc,
                         int32_t *dev_limb_start, int *total_limbs, void *stream);
int kbfv_ntt_desc(const kbfv_context *ctx, int dir, const int64_t *dev_desc, const int32_t *dev_limb_start,
                  int count, int total_limbs, void *stream);
/* Element-wise add (op 0), sub (1) or mul (2) over a ragged batch, one launch
 * (replaces per-level loops of poly_add/sub/mul batch kernels,
 * include/poly_arithmetic.cuh:26-360).  Descriptor i is eight int64 words
 * {a, b, out, s_a, s_b, s_out, mod0, limbs}: limb l of each operand at
 * base + l*stride, modulus mod0 + l, canonical residues (out may alias a or
 * b).  Upload validates and copies as kbfv_ntt_desc_upload does; the launch
 * reads only device memory (graph-capturable). */
int kbfv_poly_binop_desc_upload(const kbfv_context *ctx, const int64_t *host_desc, int count, int64_t *dev_desc,
                                int32_t *dev_limb_start, int *total_limbs, void *stream);
int kbfv_poly_binop_desc(const kbfv_context *ctx, int op, const int64_t *dev_desc, const int32_t *dev_limb_start,
                         int count, int total_limbs, void *stream);

/* ---------------------------------------------------------------- RNS kernels */

/* element-wise over `polys` polynomials [polys][limbs][N] (replaces
 * poly_add/sub/barrett batch kernels, include/poly_arithmetic.cuh:26-360) */
int kbfv_poly_add(const kbfv_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                  int mod0, int limbs, int polys, void *stream);
int kbfv_poly_sub(const kbfv_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                  int mod0, int limbs, int polys, void *stream);
int kbfv_poly_mul(const kbfv_context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                  int mod0, int limbs, int polys, void *stream);

/* HPS fast base conversion Q -> R / R -> Q, coefficient form
 * (replaces BaseConverter::bConv_HPS, include/BasisConv_QR.cuh:191-208).
 * in [polys][in_limbs][N] at in_stride elements per poly, out likewise. */
int kbfv_bconv_q_to_r(const kbfv_context *ctx, const uint64_t *x, int64_t in_stride, uint64_t *y,
                      int64_t out_stride, int polys, void *stream);
int kbfv_bconv_r_to_q(const kbfv_context *ctx, const uint64_t *x, int64_t in_stride, uint64_t *y,
                      int64_t out_stride, int polys, void *stream);

/* NTT-domain 2x2 tensor over `limbs` limbs starting at mod0
 * (replaces tensor_prod_2x2_rns_poly, include/poly_arithmetic.cuh:1180):
 * ct1, ct2 [2][limbs][N] -> out [3][limbs][N] = (a1a2, a1b2+a2b1, b1b2). */
int kbfv_tensor(const kbfv_context *ctx, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out,
                int mod0, int limbs, int batch, void *stream);

/* HPS scale-and-round t/Q over QR -> R (replaces scaleAndRound_HPS_QR_R,
 * include/poly_arithmetic.cuh:1263-1317). x [polys][qn+rn][N] -> y [polys][rn][N]. */
int kbfv_scale_round(const kbfv_context *ctx, const uint64_t *x, uint64_t *y, int polys,
                     void *stream);

/* HPS decryption scale-and-round (replaces hps_decrypt_scale_and_round,
 * include/BFVcontext.cuh:1680-1717): x [batch][qn][N] coefficient form of
 * b + a*s -> m [batch][N]. */
int kbfv_hps_decrypt(const kbfv_context *ctx, const uint64_t *x, uint64_t *m, int batch,
                     void *stream);

/* ---------------------------------------------------------------- key switching */

/* Replaces BFVContext::modUpQjtoT_23 (include/BasisConv.cuh:137):
 * x [batch][qn][N] coeff -> out [batch][dnum][tn][N] coeff. */
int kbfv_modup_qj_to_t(const kbfv_context *ctx, const uint64_t *x, uint64_t *out, int batch,
                       void *stream);
/* Replaces BFVContext::modUpPQtoT_23 (include/BasisConv.cuh:61):
 * x [batch][pn+qn][N] -> out [batch][ri_blocks][tn][N]. */
int kbfv_modup_pq_to_t(const kbfv_context *ctx, const uint64_t *x, uint64_t *out, int batch,
                       void *stream);
/* Replaces BFVContext::external_product_T / _swk_reuse
 * (include/ExternalProduct.cuh:85-177): key-stationary inner product in T,
 * cin [batch][dnum][tn][N] (NTT) x key [2][dnum][ri_blocks][tn][N] (NTT)
 * -> out [batch][2][ri_blocks][tn][N].  The key is read once per call. */
int kbfv_inner_product(const kbfv_context *ctx, const uint64_t *cin, const uint64_t *key,
                       uint64_t *out, int batch, void *stream);
/* Replaces BFVContext::modUpTtoPQl_23 (include/BasisConv.cuh:215):
 * in [batch][2][ri_blocks][tn][N] coeff -> out [batch][2][pn+qn][N]. */
int kbfv_modup_t_to_pq(const kbfv_context *ctx, const uint64_t *in, uint64_t *out, int batch,
                       void *stream);
/* Replaces BFVContext::modDownPQltoQl_23 (include/BasisConv.cuh:300):
 * in [batch][2][pn+qn][N] -> out [batch][2][qn][N]; when add != NULL the
 * [batch][2][qn][N] polynomials `add` (at add_stride per batch entry) are added. */
int kbfv_moddown(const kbfv_context *ctx, const uint64_t *in, uint64_t *out, const uint64_t *add,
                 int64_t add_stride, int batch, void *stream);

/* ---------------------------------------------------------------- randomness */

/* XSalsa20 keystream (replaces RNG::generateRandom_device, include/RNG.cuh:18) */
int kbfv_xsalsa20(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0, uint8_t *out,
                  size_t len, void *stream);
/* Samplers (replace Sampler::*, include/Sampler_impl.cuh:13-107).  kind:
 * 0 uniform mod q, 1 ternary, 2 zero-one (1/2, 1/4, 1/4), 3 CBD k=21.
 * Stream of (seed, tag); out [limbs][N] for moduli mod0.. */
int kbfv_sample(const kbfv_context *ctx, int kind, const uint8_t seed[32], uint64_t tag, int mod0,
                int limbs, uint64_t *out, void *stream);

/* ---------------------------------------------------------------- BFV scheme */

/* SecretKey ctor (include/SecretKey.cuh:20-37): ternary s over PQ, NTT form */
int kbfv_keygen_sk(const kbfv_context *ctx, const uint8_t seed[32], uint64_t *sk, void *stream);
/* BFVScheme::addEncKey (include/BFVScheme.cuh:57-85) */
int kbfv_keygen_pk(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *sk,
                   uint64_t *pk, void *stream);
/* BFVScheme::addMultKey_23 (include/KeySwitch_23.cuh:21-66) */
size_t kbfv_keygen_rlk_workspace_bytes(const kbfv_context *ctx);
int kbfv_keygen_rlk(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *sk,
                    uint64_t *rlk, void *workspace, size_t ws_bytes, void *stream);
/* BFVScheme::addSWKey_23 (include/KeySwitch_23.cuh:6-16, a stub there that
 * returns an unfilled key): key-switching key from src to s, where src
 * [qn][N] is any polynomial over Q in NTT form (src = s^2 gives a
 * relinearisation key, sampled under other stream tags than
 * kbfv_keygen_rlk's).  Layout and workspace as kbfv_keygen_rlk. */
int kbfv_keygen_swk(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *sk,
                    const uint64_t *src_q, uint64_t *key, void *workspace, size_t ws_bytes, void *stream);
/* BFVScheme::encryptMsg (include/BFVScheme.cuh:118-146); entry b of the batch
 * draws its randomness with tag index b.  m [batch][N] encoded plaintexts. */
size_t kbfv_encrypt_workspace_bytes(const kbfv_context *ctx, int batch);
int kbfv_encrypt(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *pk,
                 const uint64_t *m, uint64_t *ct, int batch, void *workspace, size_t ws_bytes,
                 void *stream);
/* BFVScheme::decryptMsg (include/BFVScheme.cuh:148-166) -> m [batch][N] */
size_t kbfv_decrypt_workspace_bytes(const kbfv_context *ctx, int batch);
int kbfv_decrypt(const kbfv_context *ctx, const uint64_t *sk, const uint64_t *ct, uint64_t *m,
                 int batch, void *workspace, size_t ws_bytes, void *stream);
/* BFVScheme::add / addAndEqual (include/BFVScheme.cuh:169-190) */
int kbfv_add(const kbfv_context *ctx, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out,
             int batch, void *stream);

/* HomMult + relinearisation (replaces BFVScheme::multAndEqual_23 / mult_23,
 * include/KeySwitch_23.cuh:192-310, and compute_t_QInv_c0c1c2 :703-926).
 * Batched form: every kernel covers the whole batch and the relinearisation key
 * is read once per call.  ct1, ct2, out: [batch][2][qn][N]; out may alias ct1. */
size_t kbfv_mult_workspace_bytes(const kbfv_context *ctx, int batch);
int kbfv_mult_relin(const kbfv_context *ctx, const uint64_t *rlk, const uint64_t *ct1,
                    const uint64_t *ct2, uint64_t *out, int batch, void *workspace,
                    size_t ws_bytes, void *stream);

/* Levels.  A ciphertext at level l (0 <= l <= qn - 1) lives modulo
 * Q_l = q_0 ... q_l and is stored compactly as [2][l + 1][N] (NTT form), the
 * reference's Ciphertext::l (include/Ciphertext.cuh:5-30) without its unused
 * limbs.  kbfv_context_level returns a view of the context at level l, owned
 * by the context (never destroyed by the caller; the full level returns the
 * context itself).  On a view, ciphertext arguments are level-l batches
 * [batch][2][l + 1][N] and these entry points run at that level: kbfv_ntt,
 * the element-wise and HPS primitives, kbfv_keygen_pk (a level-l public key
 * from the full-level secret key), kbfv_encrypt (with such a key),
 * kbfv_decrypt, kbfv_add, kbfv_negate, kbfv_mul_scalar, kbfv_add_plain and
 * kbfv_mult_relin / kbfv_mult_relin_noalloc with the full-level relinearisation
 * key.  Key generation, rotation and the external product need the full-level
 * context.  HPS at level l uses R_l = the l + 2 moduli that follow q_l in the
 * context's [Q | R] modulus table (so Q_l R_l stays one contiguous range); key
 * switching reads the full-level key on its first ceil((l + 1) / pn) digits and
 * ceil((pn + l + 1) / gamma) blocks, as the reference's key switch does with
 * its swk_blockNum (include/KeySwitch_23.cuh:255-297). */
int kbfv_context_level(const kbfv_context *ctx, int level, const kbfv_context **view);

/* Modulus switching (replaces BFVScheme::rescaleToAndEqual,
 * include/BFVScheme.cuh:469-487): level from_level -> to_level, dropping the
 * top limbs one at a time, x_i <- (x_i - x_l) q_l^-1 mod q_i in coefficient
 * form.  in [batch][2][from_level + 1][N], out [batch][2][to_level + 1][N],
 * NTT form both; out must not overlap in. */
size_t kbfv_mod_switch_workspace_bytes(const kbfv_context *ctx, int from_level, int batch);
int kbfv_mod_switch(const kbfv_context *ctx, const uint64_t *in, int from_level, uint64_t *out, int to_level,
                    int batch, void *workspace, size_t ws_bytes, void *stream);

/* Batched HomMult + relinearisation over ciphertexts of uneven RNS length:
 * entry b is a pair at level levels[b] (host array, 1 <= level <= qn - 1), the
 * entries packed back to back in ct1, ct2 and out (entry b at the sum of
 * 2 (levels[i] + 1) N words over i < b).  Each run of consecutive entries of
 * one level is one batched pipeline at that level (sort by level for the
 * largest batches); the result of every entry equals kbfv_mult_relin on its
 * level's view.  The workspace may be any size that holds one full-level
 * ciphertext (kbfv_mult_workspace_bytes(ctx, 1)); runs are chunked to fit. */
int kbfv_mult_relin_ragged(const kbfv_context *ctx, const uint64_t *rlk, const uint64_t *ct1,
                           const uint64_t *ct2, uint64_t *out, const int *levels, int batch, void *workspace,
                           size_t ws_bytes, void *stream);

/* Allocation-free HomMult + relinearisation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and bits as kbfv_mult_relin, but the intermediates live
 * in a fixed device array of KBFV_NOALLOC_SCRATCH_BYTES per GPU (part of the
 * library's code object, held until the process ends; its address is resolved
 * per device at context creation), and the call allocates nothing.  The batch
 * runs in chunks of kbfv_mult_noalloc_max_batch(ctx) ciphertexts; a shape
 * whose single ciphertext needs more scratch than the array holds (N = 2^16
 * with 28 Q primes needs about 302 MB) is rejected with KBFV_EINVAL.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams or host threads) race and must be serialised by the
 * caller.  Nothing in the array needs to start at zero. */
#define KBFV_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
size_t kbfv_mult_noalloc_max_batch(const kbfv_context *ctx);
int kbfv_mult_relin_noalloc(const kbfv_context *ctx, const uint64_t *rlk, const uint64_t *ct1,
                            const uint64_t *ct2, uint64_t *out, int batch, void *stream);

/* BFVScheme::compute_t_QInv_c0c1c2 (include/KeySwitch_23.cuh:703-926), batched:
 * the HPS-scaled tensor of each pair, out3 [batch][3][qn][N] = (a1a2 = axax_mul,
 * a1b2 + a2b1 = axbx1_mul, b1b2 = bxbx_mul), NTT form over Q; without the
 * reference's secret-key shortcut (:823-832).  Workspace:
 * kbfv_mult_workspace_bytes. */
int kbfv_mult_tensor(const kbfv_context *ctx, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out3,
                     int batch, void *workspace, size_t ws_bytes, void *stream);

/* The reference's split-IP seam, batched (BFVScheme::multAndEqual_beforeIP_23 /
 * multAndEqual_afterIP_23, include/KeySwitch_23.cuh:312-372).  before_ip runs
 * the tensor half of HomMult and writes the inner product's input ip_in
 * [batch][dnum][tn][N] (digits of d2 over T, NTT form) and the linear terms
 * d01 [batch][2][qn][N] (d1 = a1b2 + a2b1, d0 = b1b2, coefficient form);
 * kbfv_inner_product (key-stationary over the batch, the reference's
 * external_product_T_swk_reuse, include/ExternalProduct.cuh:118-177) maps
 * ip_in to ip_out [batch][2][ri_blocks][tn][N]; after_ip finishes the
 * relinearised product into out [batch][2][qn][N] (ip_out is not modified).
 * The three calls give kbfv_mult_relin's result bit for bit.  Workspace:
 * kbfv_mult_workspace_bytes. */
int kbfv_mult_before_ip(const kbfv_context *ctx, const uint64_t *ct1, const uint64_t *ct2, uint64_t *ip_in,
                        uint64_t *d01, int batch, void *workspace, size_t ws_bytes, void *stream);
int kbfv_mult_after_ip(const kbfv_context *ctx, const uint64_t *ip_out, const uint64_t *d01, uint64_t *out,
                       int batch, void *workspace, size_t ws_bytes, void *stream);

/* HIP-graph form of kbfv_mult_relin: the batched pipeline for fixed buffers is
 * captured once and each later kbfv_graph_launch replays every kernel with a
 * single submission.  No counterpart in the reference, which launches every
 * kernel from the host on each call (include/KeySwitch_23.cuh:244-310).  The
 * buffers given at creation are the ones every replay reads and writes; they
 * must outlive the graph. */
typedef struct kbfv_graph kbfv_graph;
int kbfv_mult_relin_graph_create(const kbfv_context *ctx, const uint64_t *rlk, const uint64_t *ct1,
                                 const uint64_t *ct2, uint64_t *out, int batch, void *workspace,
                                 size_t ws_bytes, kbfv_graph **graph);
int kbfv_graph_launch(kbfv_graph *graph, void *stream);
void kbfv_graph_destroy(kbfv_graph *graph);

/* Ciphertext arithmetic over [batch][2][qn][N], NTT form (BFVScheme::negate,
 * multConstAndEqual(uint64), multConstAndAddCipherEqual, addConstAndEqual,
 * include/BFVScheme.cuh:192-326).  out may alias the inputs.
 * kbfv_mul_scalar: out = k * ct (+ add when add is not null).
 * kbfv_add_plain:  out = ct + (0, NTT(Delta m)), m: [batch][N] encoded plaintexts. */
int kbfv_negate(const kbfv_context *ctx, const uint64_t *ct, uint64_t *out, int batch, void *stream);
int kbfv_mul_scalar(const kbfv_context *ctx, const uint64_t *ct, uint64_t k, const uint64_t *add,
                    uint64_t *out, int batch, void *stream);
size_t kbfv_add_plain_workspace_bytes(const kbfv_context *ctx, int batch);
int kbfv_add_plain(const kbfv_context *ctx, const uint64_t *ct, const uint64_t *m, uint64_t *out, int batch,
                   void *workspace, size_t ws_bytes, void *stream);

/* Batch encoder (BFVContext::encode / decode, include/Encoder.cuh:23-113):
 * INTT / NTT over t when t = 1 mod 2N, identity (coefficient encoding) otherwise.
 * In place on [batch][N]. */
int kbfv_encode(const kbfv_context *ctx, uint64_t *vals, int batch, void *stream);
int kbfv_decode(const kbfv_context *ctx, uint64_t *vals, int batch, void *stream);
/* The same in SEAL's slot order (BFVContext::preComputeIndex with
 * encode_BFV_kernel / decode_BFV_kernel, include/BFVcontext.cuh:102-129,
 * include/Encoder.cuh:6-21, which the reference computes but never calls):
 * two rows of N/2 slots, so the Galois element 5^r rotates each row left by r
 * and 2N - 1 swaps the rows.  In place on vals [batch][N]; tmp: [batch][N]
 * scratch.  Needs batching (t = 1 mod 2N). */
int kbfv_encode_seal(const kbfv_context *ctx, uint64_t *vals, uint64_t *tmp, int batch, void *stream);
int kbfv_decode_seal(const kbfv_context *ctx, uint64_t *vals, uint64_t *tmp, int batch, void *stream);

/* ---------------------------------------------------------------- Galois automorphisms */

/* Galois element of a left rotation by `steps`: 5^steps mod 2N (the reference's
 * rotGroups, include/BFVcontext.cuh:788-796); steps < 0 gives the conjugation
 * element 2N - 1.  Returns 0 for a null context. */
uint64_t kbfv_galois_elt(const kbfv_context *ctx, int64_t steps);
/* Coefficient-form automorphism X -> X^g on [polys][limbs][N] limbs from mod0
 * (sk_and_poly_LeftRot_kernel, include/poly_arithmetic.cuh:496-512; outputs
 * canonical).  in and out must differ. */
int kbfv_automorphism(const kbfv_context *ctx, uint64_t g, const uint64_t *in, uint64_t *out, int mod0,
                      int limbs, int polys, void *stream);
/* Galois key for g (BFVScheme::addLeftRotKey_23 / addConjKey_23,
 * include/KeySwitch_23.cuh:70-166): switches from sigma_g^-1(s) to s, in the
 * relinearisation key's layout [2][dnum][ri_blocks][tn][N], NTT over T. */
size_t kbfv_keygen_galois_workspace_bytes(const kbfv_context *ctx);
int kbfv_keygen_galois(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *sk, uint64_t g,
                       uint64_t *key, void *workspace, size_t ws_bytes, void *stream);
/* Rotation / conjugation by g (BFVScheme::leftRotateAndEqual_23, conjugate_23,
 * include/KeySwitch_23.cuh:466-578): out = sigma_g(KS(a) + (0, b)), batched.
 * ct, out: [batch][2][qn][N] NTT form; out may alias ct. */
size_t kbfv_rotate_workspace_bytes(const kbfv_context *ctx, int batch);
int kbfv_rotate(const kbfv_context *ctx, const uint64_t *gkey, uint64_t g, const uint64_t *ct, uint64_t *out,
                int batch, void *workspace, size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------- RGSW x RLWE */

/* No counterpart in the reference (its "external product" is the key-switch
 * inner product).  Gadget: base 2^base_bits digits of each Q residue. */
int kbfv_rgsw_rows(const kbfv_context *ctx, int base_bits);
size_t kbfv_rgsw_encrypt_workspace_bytes(const kbfv_context *ctx);
int kbfv_rgsw_encrypt(const kbfv_context *ctx, const uint8_t seed[32], const uint64_t *sk,
                      const uint64_t *m_coeff, int base_bits, uint64_t *rgsw, void *workspace,
                      size_t ws_bytes, void *stream);
size_t kbfv_external_product_workspace_bytes(const kbfv_context *ctx, int batch, int base_bits);
int kbfv_external_product(const kbfv_context *ctx, const uint64_t *rgsw, const uint64_t *ct,
                          uint64_t *out, int batch, int base_bits, void *workspace,
                          size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------- wire format */

/* Serialization of ciphertexts, plaintexts and keys.  No counterpart in the
 * reference, which has no serialization; the format (64-byte header, then the
 * object's words as in device memory) is specified in csrc/wire.cpp.
 * aux: the Galois element of a Galois key, the gadget base bits of an RGSW
 * ciphertext, otherwise 0.  Shapes per object: ciphertext and public key
 * [2][qn][N]; plaintext [N] (mod t); secret key [pn+qn][N]; relinearisation and
 * Galois keys [2][dnum][ri_blocks][tn][N]; RGSW [rows][2][qn][N]. */
#define KBFV_WIRE_CIPHERTEXT 1
#define KBFV_WIRE_PLAINTEXT 2
#define KBFV_WIRE_SECRET_KEY 3
#define KBFV_WIRE_PUBLIC_KEY 4
#define KBFV_WIRE_RELIN_KEY 5
#define KBFV_WIRE_GALOIS_KEY 6
#define KBFV_WIRE_RGSW 7
/* bytes of a blob holding `count` objects of `kind` (0 on bad arguments) */
size_t kbfv_wire_bytes(const kbfv_context *ctx, int kind, int count, uint64_t aux);
/* device objects -> blob in host memory; synchronizes the stream */
int kbfv_wire_save(const kbfv_context *ctx, int kind, const uint64_t *dev, int count, uint64_t aux,
                   void *buf, size_t cap, void *stream);
/* blob -> device objects after checking magic, version, kind, count, shape,
 * parameter fingerprint, checksum and every residue's range; synchronizes the
 * stream (the blob may be freed on return) */
int kbfv_wire_load(const kbfv_context *ctx, const void *buf, size_t len, int kind, uint64_t *dev,
                   int count, void *stream);
/* header fields of a blob without a context (kind, count, aux may be null) */
int kbfv_wire_peek(const void *buf, size_t len, int *kind, int *count, uint64_t *aux);
/* the payload checksum of the format */
uint64_t kbfv_wire_checksum(const uint64_t *words, size_t n);

#ifdef __cplusplus
}
#endif
#endif
