// BFV scheme pipelines and the extern "C" entry points of libkunpbfv.
//
// Replaces BFVScheme (include/BFVScheme.cuh:7-166), the relinearisation-key
// generation and HomMult of include/KeySwitch_23.cuh (:21-66, :192-310,
// :703-926) and the encoder (include/Encoder.cuh:23-113).
//
// HomMult + relinearisation, per batch chunk (every launch covers the chunk):
//   INTT_Q(ct1, ct2)  -> bConv Q->R -> NTT_R  (Q part stays the NTT-domain input)
//   tensor over QR    -> INTT_QR -> scale-and-round t/Q -> bConv R->Q
//   a1a2: digits -> NTT_T (fused) -> key-stationary inner product -> INTT_T
//         -> T->PQ -> modDown (+ a1b2+a2b1, b1b2) -> NTT_Q -> out
// No device allocation happens here: intermediates live in the caller's
// workspace (sizes from *_workspace_bytes), so a captured graph can replay it.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "kbfv_internal.h"
#include "kbfv_ops.h"

namespace {

constexpr int BS = 256;
inline unsigned blocks_for(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }
inline hipStream_t S(void *s) { return (hipStream_t)s; }

// sampler tags: (purpose << 32) | index, the same layout the oracle uses
constexpr uint64_t TAG(uint64_t purpose) { return purpose << 32; }
enum : uint64_t {
    TAG_SK = 1,
    TAG_PK_A = 2,
    TAG_PK_E = 3,
    TAG_RLK_A = 4,
    TAG_RLK_E = 5,
    TAG_ENC_V = 6,
    TAG_ENC_E0 = 7,
    TAG_ENC_E1 = 8,
    TAG_RGSW_A = 9,
    TAG_RGSW_E = 10,
    TAG_GAL_A = 11,
    TAG_GAL_E = 12,
    TAG_SWK_A = 13,
    TAG_SWK_E = 14
};

// bump allocator over the caller's workspace (256-byte aligned slices); with a
// null base it only measures.
struct Bump {
    uint8_t *base;
    size_t used = 0;
    explicit Bump(void *b) : base((uint8_t *)b) {}
    uint64_t *take(size_t elems) {
        used = (used + 255) & ~size_t(255);
        uint64_t *p = base ? (uint64_t *)(base + used) : nullptr;
        used += elems * sizeof(uint64_t);
        return p;
    }
};

// ---------------------------------------------------------------- small kernels
// out = x + sign * a*s over [polys][limbs][N]; s: [limbs][N] broadcast over polys
template <bool SUB>
__global__ void k_mac_bcast(kb_dev_tables t, const uint64_t *__restrict__ x, const uint64_t *__restrict__ a,
                            const uint64_t *__restrict__ s, uint64_t *__restrict__ out, int mod0, int limbs,
                            int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long L = (long)limbs * n;
    const long r = i % L;
    const int m = mod0 + (int)(r / n);
    const uint64_t q = t.mod[m];
    const uint64_t p = kb_mulmod(a[i], s[r], q, t.mu[2 * m], t.mu[2 * m + 1]);
    out[i] = SUB ? kb_submod(x[i], p, q) : kb_addmod(x[i], p, q);
}

// ct.a = v*pk.a + e0, ct.b = v*pk.b + u1   (v, e0, u1: [batch][qn][N]; pk broadcast)
__global__ void k_enc_combine(kb_dev_tables t, const uint64_t *__restrict__ v,
                              const uint64_t *__restrict__ e0, const uint64_t *__restrict__ u1,
                              const uint64_t *__restrict__ pk, uint64_t *__restrict__ ct, int qn, int oQ,
                              int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long L = (long)qn * n;
    const long bt = i / L, r = i % L;
    const int m = oQ + (int)(r / n);
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const uint64_t vv = v[i];
    ct[bt * 2 * L + r] = kb_addmod(kb_mulmod(vv, pk[r], q, mlo, mhi), e0[i], q);
    ct[bt * 2 * L + L + r] = kb_addmod(kb_mulmod(vv, pk[L + r], q, mlo, mhi), u1[i], q);
}

// x = b + a*s over Q (ct [batch][2][qn][N], s_q [qn][N] broadcast)
__global__ void k_dec_combine(kb_dev_tables t, const uint64_t *__restrict__ ct,
                              const uint64_t *__restrict__ s_q, uint64_t *__restrict__ x, int qn, int oQ,
                              int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long L = (long)qn * n;
    const long bt = i / L, r = i % L;
    const int m = oQ + (int)(r / n);
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    x[i] = kb_addmod(ct[bt * 2 * L + L + r], kb_mulmod(ct[bt * 2 * L + r], s_q[r], q, mlo, mhi), q);
}

// key-switching key body for digit d (over PQ):
//   b = e + P*src*[limbs of digit d] - a*keypoly
// relinearisation: src = s^2, keypoly = s; Galois g: src = s, keypoly = sigma_g^-1(s)
__global__ void k_ksk_combine(kb_dev_tables t, const uint64_t *__restrict__ e,
                              const uint64_t *__restrict__ a, const uint64_t *__restrict__ key_pq,
                              const uint64_t *__restrict__ src_q, uint64_t *__restrict__ b, int pn, int qn,
                              int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int pq = pn + qn;
    const long L = (long)pq * n;
    const long d = i / L, r = i % L;
    const int l = (int)(r / n);  // modulus index (P then Q are the first moduli)
    const uint64_t q = t.mod[l], mlo = t.mu[2 * l], mhi = t.mu[2 * l + 1];
    uint64_t v = e[i];
    const int ql = l - pn;
    if (ql >= 0 && ql >= (int)d * pn && ql < (int)d * pn + pn) {
        const uint64_t ps = kb_mulmod(src_q[(long)ql * n + (r % n)], t.P_mod_q[ql], q, mlo, mhi);
        v = kb_addmod(v, ps, q);
    }
    b[i] = kb_submod(v, kb_mulmod(a[i], key_pq[r], q, mlo, mhi), q);
}

// ciphertext negation over [batch][2][qn][N]
__global__ void k_negate(kb_dev_tables t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int oQ,
                         int qn, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t q = t.mod[oQ + (int)((i / n) % qn)];
    const uint64_t x = in[i];
    out[i] = x ? q - x : 0;
}

// out = k * in (+ add) over [batch][2][qn][N]; k reduced per limb by Barrett
template <bool ADD>
__global__ void k_mul_scalar(kb_dev_tables t, const uint64_t *__restrict__ in, uint64_t kval,
                             const uint64_t *__restrict__ add, uint64_t *__restrict__ out, int oQ, int qn, int n,
                             long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int m = oQ + (int)((i / n) % qn);
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const uint64_t kq = kb_barrett128(kbfv_u128{kval, 0}, q, mlo, mhi);
    uint64_t v = kb_mulmod(in[i], kq, q, mlo, mhi);
    if (ADD) v = kb_addmod(v, add[i], q);
    out[i] = v;
}

// slot permutation over [batch][N]: SCATTER out[map[i]] = in[i], else out[i] = in[map[i]]
template <bool SCATTER>
__global__ void k_slot_permute(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                               const uint64_t *__restrict__ map, int log_n, long total) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const long n = 1L << log_n;
    const long base = idx & ~(n - 1);
    const long i = idx & (n - 1);
    if (SCATTER)
        out[base + (long)map[i]] = in[idx];
    else
        out[idx] = in[base + (long)map[i]];
}

// coefficient-form automorphism X -> X^g (gather form, g^-1 = ginv mod 2N):
// out_j = in_i for i = j ginv mod 2N < N, else -in_(i - N)
__global__ void k_automorphism(kb_dev_tables t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                               uint64_t ginv, int mod0, int limbs, int log_n, long total) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const long n = 1L << log_n;
    const long row = idx >> log_n;
    const uint64_t j = (uint64_t)(idx & (n - 1));
    const uint64_t i = (j * ginv) & (2 * (uint64_t)n - 1);
    const uint64_t q = t.mod[mod0 + (int)(row % limbs)];
    const uint64_t *src = in + row * n;
    uint64_t v;
    if (i < (uint64_t)n) {
        v = src[i];
    } else {
        const uint64_t x = src[i - n];
        v = x ? q - x : 0;
    }
    out[idx] = v;
}

// signed lift of a plaintext polynomial m [N] (values above t/2 negative) into Q
__global__ void k_lift_plain(kb_dev_tables t, const uint64_t *__restrict__ m, uint64_t *__restrict__ out,
                             int oQ, uint64_t plain, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n), l = (int)(i / n);
    const uint64_t q = t.mod[oQ + l];
    const uint64_t v = m[k];
    out[i] = v > plain / 2 ? q - (plain - v) : v;
}

// rgsw [rows][2][qn][N]: component 1 = e - component 0 * s
__global__ void k_rgsw_body(kb_dev_tables t, const uint64_t *__restrict__ e, const uint64_t *__restrict__ s_q,
                            uint64_t *__restrict__ rgsw, int qn, int oQ, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long L = (long)qn * n;
    const long r = i / L, off = i % L;
    const int m = oQ + (int)(off / n);
    const uint64_t q = t.mod[m];
    const uint64_t a = rgsw[r * 2 * L + off];
    rgsw[r * 2 * L + L + off] = kb_submod(e[i], kb_mulmod(a, s_q[off], q, t.mu[2 * m], t.mu[2 * m + 1]), q);
}

__global__ void k_fill_zero(uint64_t *__restrict__ p, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) p[i] = 0;
}

int mac_bcast(const kbfv_context *c, bool sub, const uint64_t *x, const uint64_t *a, const uint64_t *s,
              uint64_t *out, int mod0, int limbs, int polys, hipStream_t st) {
    const long total = (long)polys * limbs * c->n;
    if (sub)
        hipLaunchKernelGGL(k_mac_bcast<true>, dim3(blocks_for(total, BS)), dim3(BS), 0, st, c->t, x, a, s, out,
                           mod0, limbs, c->n, total);
    else
        hipLaunchKernelGGL(k_mac_bcast<false>, dim3(blocks_for(total, BS)), dim3(BS), 0, st, c->t, x, a, s, out,
                           mod0, limbs, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int fill_zero(uint64_t *p, long total, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(total, BS)), dim3(BS), 0, st, p, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

// ---------------------------------------------------------------- HomMult layout
struct MultWs {
    uint64_t *X1, *X2;      // coefficient-form inputs [cb][2][qn][N]      (region A)
    uint64_t *SR;           // scale-and-round out   [cb][3][rn][N]        (region A, later)
    uint64_t *R1, *R2;      // R-extension           [cb][2][rn][N]        (region B)
    uint64_t *DQ;           // t/Q-scaled tensor     [cb][3][qn][N]        (region B, later)
    uint64_t *D3;           // QR tensor             [cb][3][qn+rn][N]     (region C)
    uint64_t *CT;           // digits over T         [cb][dnum][tn][N]     (region C, later)
    uint64_t *MD;           // modDown out           [cb][2][qn][N]        (region C, later)
    uint64_t *IP;           // inner product         [cb][2][B][tn][N]     (region E)
    uint64_t *PQ;           // T -> PQ, only for shapes without the fused modUp/modDown (region F)
};

size_t mult_layout(const kbfv_context *c, int cb, void *base, MultWs *w) {
    const size_t N = c->n, qn = c->qn, rn = c->rn;
    const size_t A = std::max(4 * qn, 3 * rn) * N;
    const size_t Bq = std::max(4 * rn, 3 * qn) * N;
    const size_t Cc = std::max(std::max(3 * (qn + rn), (size_t)c->dnum * c->tn), 2 * qn) * N;
    const size_t E = 2 * (size_t)c->ri_blocks * c->tn * N;
    const size_t F = kb_modup_moddown_fused(c) ? 0 : 2 * (size_t)(c->pn + c->qn) * N;
    Bump b(base);
    uint64_t *a = b.take(A * cb), *bb = b.take(Bq * cb), *cc = b.take(Cc * cb);
    uint64_t *e = b.take(E * cb), *f = F ? b.take(F * cb) : nullptr;
    if (w) {
        w->X1 = a;
        w->X2 = a ? a + 2 * qn * N * cb : nullptr;
        w->SR = a;
        w->R1 = bb;
        w->R2 = bb ? bb + 2 * rn * N * cb : nullptr;
        w->DQ = bb;
        w->D3 = cc;
        w->CT = cc;
        w->MD = cc;
        w->IP = e;
        w->PQ = f;
    }
    return b.used;
}

int mult_chunk(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1, const uint64_t *ct2,
               uint64_t *out, int cb, const MultWs &w, hipStream_t st) {
    const long N = c->n, qn = c->qn, rn = c->rn, sn = qn + rn;
    const long CTS = 2 * qn * N;  // ciphertext stride
    // 1. INTT of both inputs (out of place; the NTT-domain inputs stay for the
    // tensor), left N times too large: the base extension folds n^-1 in
    KB_TRY(kb_intt_noscale(c, ct1, w.X1, c->oQ, qn, 2, cb, CTS, qn * N, N, CTS, qn * N, N, st));
    KB_TRY(kb_intt_noscale(c, ct2, w.X2, c->oQ, qn, 2, cb, CTS, qn * N, N, CTS, qn * N, N, st));
    // 2. HPS base extension Q -> R, 3. NTT over R
    // (X2 follows X1 and R2 follows R1: one launch covers both ciphertexts)
    KB_TRY(kb_bconv(c, true, w.X1, qn * N, w.R1, rn * N, 4 * cb, st, true));
    KB_TRY(kb_ntt(c, +1, w.R1, w.R1, c->oR, rn, 4 * cb, 1, 0, rn * N, N, 0, rn * N, N, st));
    // 4. tensor over QR, 5. INTT over QR
    KB_TRY(kb_tensor_qr(c, ct1, ct2, w.R1, w.R2, w.D3, cb, st));
    KB_TRY(kb_ntt(c, -1, w.D3, w.D3, c->oQ, sn, 3 * cb, 1, 0, sn * N, N, 0, sn * N, N, st));
    // 6. scale by t/Q and round into R, 7. back to Q
    KB_TRY(kb_scale_round(c, w.D3, w.SR, 3 * cb, st));
    KB_TRY(kb_bconv(c, false, w.SR, rn * N, w.DQ, qn * N, 3 * cb, st));
    // 8. relinearise a1a2: digits -> NTT over T
    KB_TRY(kb_digits_to_t_ntt(c, w.DQ, 3 * qn * N, w.CT, cb, nullptr, st));
    // 9. key-stationary inner product; 10. INTT over T, 11. T -> PQ and
    // 12. modDown with the a1b2+a2b1 and b1b2 terms added
    KB_TRY(kb_inner_product(c, w.CT, rlk, w.IP, cb, st));
    KB_TRY(kb_ks_tail(c, w.IP, w.IP, w.MD, w.DQ + qn * N, 3 * qn * N, cb, w.PQ, st));
    // 13. back to the NTT domain, into the output ciphertexts
    return kb_ntt(c, +1, w.MD, out, c->oQ, qn, 2, cb, CTS, qn * N, N, CTS, qn * N, N, st);
}

}  // namespace

// ---------------------------------------------------------------- thin wrappers
extern "C" int kbfv_ntt(const kbfv_context *c, int dir, const uint64_t *src, uint64_t *dst, int mod0,
                        int limbs, int polys, int batch, int64_t sb, int64_t sp, int64_t sl, int64_t db,
                        int64_t dp, int64_t dl, void *st) {
    KB_REQUIRE(dir == 1 || dir == -1, "dir must be +1 or -1");
    return kb_ntt(c, dir, src, dst, mod0, limbs, polys, batch, sb, sp, sl, db, dp, dl, S(st));
}
extern "C" int kbfv_poly_add(const kbfv_context *c, const uint64_t *a, const uint64_t *b, uint64_t *o,
                             int mod0, int limbs, int polys, void *st) {
    return kb_poly_binop(c, 0, a, b, o, mod0, limbs, polys, S(st));
}
extern "C" int kbfv_poly_sub(const kbfv_context *c, const uint64_t *a, const uint64_t *b, uint64_t *o,
                             int mod0, int limbs, int polys, void *st) {
    return kb_poly_binop(c, 1, a, b, o, mod0, limbs, polys, S(st));
}
extern "C" int kbfv_poly_mul(const kbfv_context *c, const uint64_t *a, const uint64_t *b, uint64_t *o,
                             int mod0, int limbs, int polys, void *st) {
    return kb_poly_binop(c, 2, a, b, o, mod0, limbs, polys, S(st));
}
extern "C" int kbfv_bconv_q_to_r(const kbfv_context *c, const uint64_t *x, int64_t is, uint64_t *y,
                                 int64_t os, int polys, void *st) {
    return kb_bconv(c, true, x, is, y, os, polys, S(st));
}
extern "C" int kbfv_bconv_r_to_q(const kbfv_context *c, const uint64_t *x, int64_t is, uint64_t *y,
                                 int64_t os, int polys, void *st) {
    return kb_bconv(c, false, x, is, y, os, polys, S(st));
}
extern "C" int kbfv_tensor(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out,
                           int mod0, int limbs, int batch, void *st) {
    return kb_tensor(c, ct1, ct2, out, mod0, limbs, batch, S(st));
}
extern "C" int kbfv_scale_round(const kbfv_context *c, const uint64_t *x, uint64_t *y, int polys, void *st) {
    return kb_scale_round(c, x, y, polys, S(st));
}
extern "C" int kbfv_hps_decrypt(const kbfv_context *c, const uint64_t *x, uint64_t *m, int batch, void *st) {
    return kb_hps_decrypt(c, x, m, batch, S(st));
}
extern "C" int kbfv_modup_qj_to_t(const kbfv_context *c, const uint64_t *x, uint64_t *out, int batch,
                                  void *st) {
    KB_REQUIRE(c, "null context");
    return kb_modup_qj_to_t(c, x, (long)c->qn * c->n, out, batch, S(st));
}
extern "C" int kbfv_modup_pq_to_t(const kbfv_context *c, const uint64_t *x, uint64_t *out, int batch,
                                  void *st) {
    return kb_modup_pq_to_t(c, x, out, batch, S(st));
}
extern "C" int kbfv_inner_product(const kbfv_context *c, const uint64_t *cin, const uint64_t *key,
                                  uint64_t *out, int batch, void *st) {
    return kb_inner_product(c, cin, key, out, batch, S(st));
}
extern "C" int kbfv_modup_t_to_pq(const kbfv_context *c, const uint64_t *in, uint64_t *out, int batch,
                                  void *st) {
    return kb_modup_t_to_pq(c, in, out, batch, S(st));
}
extern "C" int kbfv_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
                            int64_t add_stride, int batch, void *st) {
    return kb_moddown(c, in, out, add, add_stride, batch, S(st));
}
extern "C" int kbfv_xsalsa20(const uint8_t key[32], const uint8_t nonce[24], uint64_t block0, uint8_t *out,
                             size_t len, void *st) {
    return kb_xsalsa20(key, nonce, block0, out, len, S(st));
}
extern "C" int kbfv_sample(const kbfv_context *c, int kind, const uint8_t seed[32], uint64_t tag, int mod0,
                           int limbs, uint64_t *out, void *st) {
    return kb_sample(c, kind, seed, tag, 0, mod0, limbs, out, 0, 1, S(st));
}

// ---------------------------------------------------------------- keys
extern "C" int kbfv_keygen_sk(const kbfv_context *c, const uint8_t seed[32], uint64_t *sk, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && seed && sk, "null argument");
    const int pq = c->pn + c->qn;
    KB_TRY(kb_sample(c, 1, seed, TAG(TAG_SK), 0, c->oP, pq, sk, 0, 1, S(st)));
    return kb_ntt_inplace(c, +1, sk, c->oP, pq, 1, S(st));
}

extern "C" int kbfv_keygen_pk(const kbfv_context *c, const uint8_t seed[32], const uint64_t *sk, uint64_t *pk,
                              void *st) {
    KB_REQUIRE(c && seed && sk && pk, "null argument");
    const long L = (long)c->qn * c->n;
    uint64_t *a = pk, *b = pk + L;
    const uint64_t *s_q = sk + (long)c->pn * c->n;
    KB_TRY(kb_sample(c, 0, seed, TAG(TAG_PK_A), 0, c->oQ, c->qn, a, 0, 1, S(st)));
    KB_TRY(kb_sample(c, 3, seed, TAG(TAG_PK_E), 0, c->oQ, c->qn, b, 0, 1, S(st)));
    KB_TRY(kb_ntt_inplace(c, +1, b, c->oQ, c->qn, 1, S(st)));
    return mac_bcast(c, true, b, a, s_q, b, c->oQ, c->qn, 1, S(st));  // b = e - a*s
}

static size_t rlk_layout(const kbfv_context *c, void *base, uint64_t **s2, uint64_t **e, uint64_t **a,
                         uint64_t **b) {
    const size_t N = c->n, pq = c->pn + c->qn, D = c->dnum;
    Bump bp(base);
    uint64_t *p0 = bp.take(c->qn * N), *p1 = bp.take(D * pq * N), *p2 = bp.take(D * pq * N),
             *p3 = bp.take(D * pq * N);
    if (s2) *s2 = p0, *e = p1, *a = p2, *b = p3;
    return bp.used;
}

extern "C" size_t kbfv_keygen_rlk_workspace_bytes(const kbfv_context *c) {
    return c ? rlk_layout(c, nullptr, nullptr, nullptr, nullptr, nullptr) : 0;
}

// key-switching key from src (NTT over Q) to s, sampled under (tag_a, tag_e):
// digit d holds a_d uniform and b_d = e_d + P*src on the limbs of digit d - a_d*s,
// converted to T block by block and NTT over T (include/KeySwitch_23.cuh:21-66)
static int keygen_ksk(const kbfv_context *c, const uint8_t seed[32], uint64_t tag_a, uint64_t tag_e,
                      const uint64_t *src_q, const uint64_t *sk, uint64_t *key, uint64_t *e, uint64_t *a,
                      uint64_t *b, hipStream_t s) {
    const long N = c->n;
    const int pn = c->pn, qn = c->qn, pq = pn + qn, D = c->dnum, B = c->ri_blocks, tn = c->tn;
    KB_TRY(kb_sample(c, 3, seed, tag_e, 0, c->oP, pq, e, 0, D, s));
    KB_TRY(kb_ntt_inplace(c, +1, e, c->oP, pq, D, s));
    KB_TRY(kb_sample(c, 0, seed, tag_a, 0, c->oP, pq, a, 0, D, s));
    const long total = (long)D * pq * N;
    hipLaunchKernelGGL(k_ksk_combine, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, e, a, sk, src_q, b, pn,
                       qn, c->n, total);
    KB_LAUNCH_CHECK();
    KB_TRY(kb_ntt_inplace(c, -1, a, c->oP, pq, D, s));
    KB_TRY(kb_ntt_inplace(c, -1, b, c->oP, pq, D, s));
    const long half = (long)D * B * tn * N;
    KB_TRY(kb_modup_pq_to_t(c, a, key, D, s));
    KB_TRY(kb_modup_pq_to_t(c, b, key + half, D, s));
    return kb_ntt_inplace(c, +1, key, c->oT, tn, 2 * D * B, s);
}

extern "C" int kbfv_keygen_rlk(const kbfv_context *c, const uint8_t seed[32], const uint64_t *sk, uint64_t *rlk,
                               void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && seed && sk && rlk && ws, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_keygen_rlk_workspace_bytes(c), "workspace too small");
    uint64_t *s2, *e, *a, *b;
    rlk_layout(c, ws, &s2, &e, &a, &b);
    const uint64_t *s_q = sk + (long)c->pn * c->n;
    KB_TRY(kb_poly_binop(c, 2, s_q, s_q, s2, c->oQ, c->qn, 1, S(st)));
    return keygen_ksk(c, seed, TAG(TAG_RLK_A), TAG(TAG_RLK_E), s2, sk, rlk, e, a, b, S(st));
}

extern "C" int kbfv_keygen_swk(const kbfv_context *c, const uint8_t seed[32], const uint64_t *sk,
                               const uint64_t *src_q, uint64_t *key, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && seed && sk && src_q && key && ws, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_keygen_rlk_workspace_bytes(c), "workspace too small");
    uint64_t *s2, *e, *a, *b;
    rlk_layout(c, ws, &s2, &e, &a, &b);
    return keygen_ksk(c, seed, TAG(TAG_SWK_A), TAG(TAG_SWK_E), src_q, sk, key, e, a, b, S(st));
}

// ---------------------------------------------------------------- encrypt / decrypt
static size_t enc_layout(const kbfv_context *c, int batch, void *base, uint64_t **v, uint64_t **e0,
                         uint64_t **u1, uint64_t **dm) {
    const size_t L = (size_t)c->qn * c->n * batch;
    Bump bp(base);
    uint64_t *p0 = bp.take(L), *p1 = bp.take(L), *p2 = bp.take(L), *p3 = bp.take(L);
    if (v) *v = p0, *e0 = p1, *u1 = p2, *dm = p3;
    return bp.used;
}
extern "C" size_t kbfv_encrypt_workspace_bytes(const kbfv_context *c, int batch) {
    return (c && batch > 0) ? enc_layout(c, batch, nullptr, nullptr, nullptr, nullptr, nullptr) : 0;
}

extern "C" int kbfv_encrypt(const kbfv_context *c, const uint8_t seed[32], const uint64_t *pk, const uint64_t *m,
                            uint64_t *ct, int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && seed && pk && m && ct && ws && batch > 0, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_encrypt_workspace_bytes(c, batch), "workspace too small");
    uint64_t *v, *e0, *u1, *dm;
    enc_layout(c, batch, ws, &v, &e0, &u1, &dm);
    const int qn = c->qn;
    hipStream_t s = S(st);
    KB_TRY(kb_sample(c, 2, seed, TAG(TAG_ENC_V), 0, c->oQ, qn, v, 0, batch, s));
    KB_TRY(kb_sample(c, 3, seed, TAG(TAG_ENC_E0), 0, c->oQ, qn, e0, 0, batch, s));
    KB_TRY(kb_sample(c, 3, seed, TAG(TAG_ENC_E1), 0, c->oQ, qn, u1, 0, batch, s));
    KB_TRY(kb_delta_m(c, m, dm, batch, s));
    KB_TRY(kb_poly_binop(c, 0, u1, dm, u1, c->oQ, qn, batch, s));  // e1 + Delta*m
    KB_TRY(kb_ntt_inplace(c, +1, v, c->oQ, qn, batch, s));
    KB_TRY(kb_ntt_inplace(c, +1, e0, c->oQ, qn, batch, s));
    KB_TRY(kb_ntt_inplace(c, +1, u1, c->oQ, qn, batch, s));
    const long total = (long)batch * qn * c->n;
    hipLaunchKernelGGL(k_enc_combine, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, v, e0, u1, pk, ct, qn,
                       c->oQ, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

extern "C" size_t kbfv_decrypt_workspace_bytes(const kbfv_context *c, int batch) {
    return (c && batch > 0) ? (size_t)c->qn * c->n * batch * sizeof(uint64_t) : 0;
}

extern "C" int kbfv_decrypt(const kbfv_context *c, const uint64_t *sk, const uint64_t *ct, uint64_t *m, int batch,
                            void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && sk && ct && m && ws && batch > 0, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_decrypt_workspace_bytes(c, batch), "workspace too small");
    uint64_t *x = (uint64_t *)ws;
    const int qn = c->qn;
    hipStream_t s = S(st);
    const long total = (long)batch * qn * c->n;
    hipLaunchKernelGGL(k_dec_combine, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, ct,
                       sk + (long)c->pn * c->n, x, qn, c->oQ, c->n, total);
    KB_LAUNCH_CHECK();
    KB_TRY(kb_ntt_inplace(c, -1, x, c->oQ, qn, batch, s));
    return kb_hps_decrypt(c, x, m, batch, s);
}

extern "C" int kbfv_add(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out, int batch,
                        void *st) {
    KB_REQUIRE(c && ct1 && ct2 && out && batch > 0, "null argument");
    // [batch][2][qn][N] == [2*batch][qn][N]
    return kb_poly_binop(c, 0, ct1, ct2, out, c->oQ, c->qn, 2 * batch, S(st));
}

// ---------------------------------------------------------------- HomMult + relinearisation
extern "C" size_t kbfv_mult_workspace_bytes(const kbfv_context *c, int batch) {
    return (c && batch > 0) ? mult_layout(c, batch, nullptr, nullptr) : 0;
}

namespace {
// batched HomMult + relinearisation over a workspace of ws_bytes, in chunks of
// as many ciphertexts as it holds
int mult_relin_ws(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1, const uint64_t *ct2,
                  uint64_t *out, int batch, void *ws, size_t ws_bytes, hipStream_t st) {
    KB_REQUIRE(c->rn <= 64, "too many limbs");
    // largest chunk that fits the workspace
    int cb = batch;
    while (cb > 1 && mult_layout(c, cb, nullptr, nullptr) > ws_bytes) cb = (cb + 1) / 2;
    KB_REQUIRE(mult_layout(c, cb, nullptr, nullptr) <= ws_bytes, "workspace too small for one ciphertext");
    const long CTS = 2L * c->qn * c->n;
    for (int b0 = 0; b0 < batch; b0 += cb) {
        const int n = std::min(cb, batch - b0);
        // laid out for this chunk's own size: the pipeline treats the second
        // input's polynomials (X2, R2) as following the first's directly
        MultWs w;
        mult_layout(c, n, ws, &w);
        KB_TRY(mult_chunk(c, rlk, ct1 + b0 * CTS, ct2 + b0 * CTS, out + b0 * CTS, n, w, st));
    }
    return KBFV_OK;
}
}  // namespace

// Scratch of the allocation-free HomMult (kbfv_mult_relin_noalloc): one copy
// per GPU, loaded with this file's code object and held until the process
// ends.  In host code `scratch` is only a host-side stand-in for the symbol;
// its device address on each GPU comes from hipGetSymbolAddress, resolved
// once per device by kb_noalloc_setup (called at context creation with that
// device current) and kept in noalloc_ptr.
static __device__ __align__(256) uint8_t scratch[KBFV_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
void *noalloc_ptr[kMaxDevices] = {};
}  // namespace

int kb_noalloc_setup(int device) {
    KB_REQUIRE(device >= 0 && device < kMaxDevices, "device index %d out of range", device);
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return KBFV_OK;
    int cur = -1;
    KB_HIP(hipGetDevice(&cur));
    KB_REQUIRE(cur == device, "scratch set-up needs device %d current (current: %d)", device, cur);
    void *ptr = nullptr;
    KB_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    KB_REQUIRE(ptr != nullptr, "no device address for the scratch array");
    noalloc_ptr[device] = ptr;
    return KBFV_OK;
}

extern "C" int kbfv_mult_relin(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1,
                               const uint64_t *ct2, uint64_t *out, int batch, void *ws, size_t ws_bytes,
                               void *st) {
    KB_REQUIRE(c && rlk && ct1 && ct2 && out && ws && batch > 0, "null argument");
    return mult_relin_ws(c, rlk, ct1, ct2, out, batch, ws, ws_bytes, S(st));
}

// ---------------------------------------------------------------- levels
namespace {
// one modulus-switching step over [polys][stride][N] coefficient-form limbs:
// x_i <- (x_i - x_l) q_l^-1 mod q_i for i < l (limb l is dropped), x_l the
// canonical residue mod q_l
__global__ void k_drop_limb(kb_dev_tables t, uint64_t *__restrict__ x, int l, int qn, int oQ, long stride, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const long poly = blockIdx.z;
    uint64_t *px = x + poly * stride;
    const int m = oQ + i;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const uint64_t xl = kb_barrett128(kbfv_u128{px[(long)l * n + k], 0}, q, mlo, mhi);
    const uint64_t d = kb_submod(px[(long)i * n + k], xl, q);
    const kbfv_tw w = t.ql_inv_mod_qi[(long)l * qn + i];
    px[(long)i * n + k] = kb_shoup(d, w.w, w.ws, q);
}
}  // namespace

extern "C" size_t kbfv_mod_switch_workspace_bytes(const kbfv_context *c, int from_level, int batch) {
    if (!c || batch <= 0 || from_level < 0 || from_level >= kb_root(c)->qn) return 0;
    return (size_t)batch * 2 * (from_level + 1) * c->n * sizeof(uint64_t);
}

extern "C" int kbfv_mod_switch(const kbfv_context *c, const uint64_t *in, int from_level, uint64_t *out,
                               int to_level, int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && in && out && ws && batch > 0, "null argument");
    const kbfv_context *r = kb_root(c);
    KB_REQUIRE(from_level < r->qn && to_level >= 0 && to_level <= from_level, "levels %d -> %d out of range",
               from_level, to_level);
    KB_REQUIRE(ws_bytes >= kbfv_mod_switch_workspace_bytes(r, from_level, batch), "workspace too small");
    KB_REQUIRE(2L * batch <= 65535, "batch too large for one launch");
    const long N = r->n, lf = from_level + 1, lt = to_level + 1;
    uint64_t *x = (uint64_t *)ws;
    // coefficient form of every limb, then drop limbs from_level .. to_level + 1
    KB_TRY(kb_ntt(r, -1, in, x, r->oQ, (int)lf, 2 * batch, 1, 0, lf * N, N, 0, lf * N, N, S(st)));
    for (int l = from_level; l > to_level; --l) {
        hipLaunchKernelGGL(k_drop_limb, dim3(r->n / BS, l, 2 * batch), dim3(BS), 0, S(st), r->t, x, l, r->qn, r->oQ,
                           lf * N, r->n);
        KB_LAUNCH_CHECK();
    }
    return kb_ntt(r, +1, x, out, r->oQ, (int)lt, 2 * batch, 1, 0, lf * N, N, 0, lt * N, N, S(st));
}

extern "C" int kbfv_mult_relin_ragged(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1,
                                      const uint64_t *ct2, uint64_t *out, const int *levels, int batch, void *ws,
                                      size_t ws_bytes, void *st) {
    KB_REQUIRE(c && rlk && ct1 && ct2 && out && levels && ws && batch > 0, "null argument");
    const kbfv_context *r = kb_root(c);
    for (int b = 0; b < batch; ++b)
        KB_REQUIRE(levels[b] >= 1 && levels[b] < r->qn, "entry %d: level %d out of range [1, %d]", b, levels[b],
                   r->qn - 1);
    // runs of equal level are contiguous compact batches: one batched pipeline each
    long off = 0;
    for (int b0 = 0; b0 < batch;) {
        int b1 = b0 + 1;
        while (b1 < batch && levels[b1] == levels[b0]) ++b1;
        const kbfv_context *v = levels[b0] == r->qn - 1 ? r : r->views[levels[b0]];
        KB_TRY(mult_relin_ws(v, rlk, ct1 + off, ct2 + off, out + off, b1 - b0, ws, ws_bytes, S(st)));
        off += (long)(b1 - b0) * 2 * (levels[b0] + 1) * r->n;
        b0 = b1;
    }
    return KBFV_OK;
}

extern "C" size_t kbfv_mult_noalloc_max_batch(const kbfv_context *c) {
    if (!c) return 0;
    size_t cb = 0;
    while (cb < (1u << 20) && mult_layout(c, (int)cb + 1, nullptr, nullptr) <= KBFV_NOALLOC_SCRATCH_BYTES) ++cb;
    return cb;
}

extern "C" int kbfv_mult_relin_noalloc(const kbfv_context *c, const uint64_t *rlk, const uint64_t *ct1,
                                       const uint64_t *ct2, uint64_t *out, int batch, void *st) {
    KB_REQUIRE(c && rlk && ct1 && ct2 && out && batch > 0, "null argument");
    const size_t one = mult_layout(c, 1, nullptr, nullptr);
    KB_REQUIRE(one <= KBFV_NOALLOC_SCRATCH_BYTES,
               "HomMult of one ciphertext at N=%d with %d Q primes needs %zu bytes of scratch; the "
               "allocation-free entry point holds %zu (use kbfv_mult_relin with a workspace)",
               c->n, c->qn, one, (size_t)KBFV_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    {
        std::lock_guard<std::mutex> lk(noalloc_mu);
        ws = (c->device >= 0 && c->device < kMaxDevices) ? noalloc_ptr[c->device] : nullptr;
    }
    KB_REQUIRE(ws != nullptr, "scratch of device %d was not set up at context creation", c->device);
    return mult_relin_ws(c, rlk, ct1, ct2, out, batch, ws, KBFV_NOALLOC_SCRATCH_BYTES, S(st));
}

// compute_t_QInv_c0c1c2 (include/KeySwitch_23.cuh:703-926), batched: steps 1-7
// of mult_chunk, then the three scaled tensor polynomials back to the NTT domain
extern "C" int kbfv_mult_tensor(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2, uint64_t *out3,
                                int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && ct1 && ct2 && out3 && ws && batch > 0, "null argument");
    int cb = batch;
    while (cb > 1 && mult_layout(c, cb, nullptr, nullptr) > ws_bytes) cb = (cb + 1) / 2;
    KB_REQUIRE(mult_layout(c, cb, nullptr, nullptr) <= ws_bytes, "workspace too small for one ciphertext");
    const long N = c->n, qn = c->qn, rn = c->rn, sn = qn + rn, CTS = 2 * qn * N, D3S = 3 * qn * N;
    hipStream_t s = S(st);
    for (int b0 = 0; b0 < batch; b0 += cb) {
        const int n = std::min(cb, batch - b0);
        MultWs w;
        mult_layout(c, n, ws, &w);
        const uint64_t *a = ct1 + b0 * CTS, *b = ct2 + b0 * CTS;
        KB_TRY(kb_intt_noscale(c, a, w.X1, c->oQ, qn, 2, n, CTS, qn * N, N, CTS, qn * N, N, s));
        KB_TRY(kb_intt_noscale(c, b, w.X2, c->oQ, qn, 2, n, CTS, qn * N, N, CTS, qn * N, N, s));
        KB_TRY(kb_bconv(c, true, w.X1, qn * N, w.R1, rn * N, 4 * n, s, true));
        KB_TRY(kb_ntt(c, +1, w.R1, w.R1, c->oR, rn, 4 * n, 1, 0, rn * N, N, 0, rn * N, N, s));
        KB_TRY(kb_tensor_qr(c, a, b, w.R1, w.R2, w.D3, n, s));
        KB_TRY(kb_ntt(c, -1, w.D3, w.D3, c->oQ, sn, 3 * n, 1, 0, sn * N, N, 0, sn * N, N, s));
        KB_TRY(kb_scale_round(c, w.D3, w.SR, 3 * n, s));
        KB_TRY(kb_bconv(c, false, w.SR, rn * N, w.DQ, qn * N, 3 * n, s));
        KB_TRY(kb_ntt(c, +1, w.DQ, out3 + b0 * D3S, c->oQ, qn, 3 * n, 1, 0, qn * N, N, 0, qn * N, N, s));
    }
    return KBFV_OK;
}

// ---------------------------------------------------------------- split-IP seam
// The reference's multAndEqual_beforeIP_23 / multAndEqual_afterIP_23
// (include/KeySwitch_23.cuh:312-372), batched, around kbfv_inner_product.
extern "C" int kbfv_mult_before_ip(const kbfv_context *c, const uint64_t *ct1, const uint64_t *ct2,
                                   uint64_t *ip_in, uint64_t *d01, int batch, void *ws, size_t ws_bytes,
                                   void *st) {
    KB_REQUIRE(c && ct1 && ct2 && ip_in && d01 && ws && batch > 0, "null argument");
    int cb = batch;
    while (cb > 1 && mult_layout(c, cb, nullptr, nullptr) > ws_bytes) cb = (cb + 1) / 2;
    KB_REQUIRE(mult_layout(c, cb, nullptr, nullptr) <= ws_bytes, "workspace too small for one ciphertext");
    const long N = c->n, qn = c->qn, rn = c->rn, sn = qn + rn, CTS = 2 * qn * N;
    const long IPS = (long)c->dnum * c->tn * N;
    hipStream_t s = S(st);
    for (int b0 = 0; b0 < batch; b0 += cb) {
        const int n = std::min(cb, batch - b0);
        MultWs w;  // per chunk size, as in kbfv_mult_relin
        mult_layout(c, n, ws, &w);
        const uint64_t *a = ct1 + b0 * CTS, *b = ct2 + b0 * CTS;
        // steps 1-8 of the fused pipeline (mult_chunk), the digits going to ip_in
        KB_TRY(kb_intt_noscale(c, a, w.X1, c->oQ, qn, 2, n, CTS, qn * N, N, CTS, qn * N, N, s));
        KB_TRY(kb_intt_noscale(c, b, w.X2, c->oQ, qn, 2, n, CTS, qn * N, N, CTS, qn * N, N, s));
        KB_TRY(kb_bconv(c, true, w.X1, qn * N, w.R1, rn * N, 4 * n, s, true));
        KB_TRY(kb_ntt(c, +1, w.R1, w.R1, c->oR, rn, 4 * n, 1, 0, rn * N, N, 0, rn * N, N, s));
        KB_TRY(kb_tensor_qr(c, a, b, w.R1, w.R2, w.D3, n, s));
        KB_TRY(kb_ntt(c, -1, w.D3, w.D3, c->oQ, sn, 3 * n, 1, 0, sn * N, N, 0, sn * N, N, s));
        KB_TRY(kb_scale_round(c, w.D3, w.SR, 3 * n, s));
        KB_TRY(kb_bconv(c, false, w.SR, rn * N, w.DQ, qn * N, 3 * n, s));
        KB_TRY(kb_digits_to_t_ntt(c, w.DQ, 3 * qn * N, ip_in + b0 * IPS, n, nullptr, s));
        // d1 (a1b2 + a2b1) and d0 (b1b2), coefficient form: polys 1 and 2 of each triple
        KB_HIP(hipMemcpy2DAsync(d01 + b0 * CTS, CTS * 8, w.DQ + qn * N, 3 * qn * N * 8, CTS * 8, n,
                                hipMemcpyDeviceToDevice, s));
    }
    return KBFV_OK;
}

extern "C" int kbfv_mult_after_ip(const kbfv_context *c, const uint64_t *ip_out, const uint64_t *d01,
                                  uint64_t *out, int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && ip_out && d01 && out && ws && batch > 0, "null argument");
    int cb = batch;
    while (cb > 1 && mult_layout(c, cb, nullptr, nullptr) > ws_bytes) cb = (cb + 1) / 2;
    KB_REQUIRE(mult_layout(c, cb, nullptr, nullptr) <= ws_bytes, "workspace too small for one ciphertext");
    MultWs w;
    mult_layout(c, cb, ws, &w);
    const long N = c->n, qn = c->qn, CTS = 2 * qn * N;
    const long IPO = 2L * c->ri_blocks * c->tn * N;
    hipStream_t s = S(st);
    for (int b0 = 0; b0 < batch; b0 += cb) {
        const int n = std::min(cb, batch - b0);
        // steps 10-13 of the fused pipeline; the caller's inner-product output is left as it is
        KB_TRY(kb_ks_tail(c, ip_out + b0 * IPO, w.IP, w.MD, d01 + b0 * CTS, CTS, n, w.PQ, s));
        KB_TRY(kb_ntt(c, +1, w.MD, out + b0 * CTS, c->oQ, qn, 2, n, CTS, qn * N, N, CTS, qn * N, N, s));
    }
    return KBFV_OK;
}

// ---------------------------------------------------------------- Galois automorphisms
// (BFVScheme::addLeftRotKey_23 / addConjKey_23 / leftRotateAndEqual_23 /
// conjugate_23, include/KeySwitch_23.cuh:70-166, 466-584)
static uint64_t inv_mod_2n(uint64_t g, uint64_t m2) {
    uint64_t x = g;  // g * g = 1 mod 8; each Newton step doubles the correct bits
    for (int i = 0; i < 6; ++i) x *= 2 - g * x;
    return x & (m2 - 1);
}

static int launch_automorphism(const kbfv_context *c, uint64_t g, const uint64_t *in, uint64_t *out, int mod0,
                               int limbs, int polys, hipStream_t s) {
    const uint64_t m2 = 2 * (uint64_t)c->n;
    KB_REQUIRE((g & 1) && g < m2, "Galois element must be odd and below 2N");
    KB_REQUIRE(in != out, "automorphism cannot run in place");
    KB_REQUIRE(mod0 >= 0 && limbs > 0 && mod0 + limbs <= c->nmod && polys > 0, "bad limb range");
    const long total = (long)polys * limbs * c->n;
    hipLaunchKernelGGL(k_automorphism, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, in, out,
                       inv_mod_2n(g, m2), mod0, limbs, c->log_n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

extern "C" uint64_t kbfv_galois_elt(const kbfv_context *c, int64_t steps) {
    if (!c) return 0;
    const uint64_t m2 = 2 * (uint64_t)c->n;
    if (steps < 0) return m2 - 1;  // conjugation X -> X^-1
    uint64_t g = 1;
    for (int64_t i = 0; i < steps % (c->n / 2); ++i) g = g * 5 % m2;  // 5 has order N/2 mod 2N
    return g;
}

extern "C" int kbfv_automorphism(const kbfv_context *c, uint64_t g, const uint64_t *in, uint64_t *out,
                                 int mod0, int limbs, int polys, void *st) {
    KB_REQUIRE(c && in && out, "null argument");
    return launch_automorphism(c, g, in, out, mod0, limbs, polys, S(st));
}

static size_t gal_layout(const kbfv_context *c, void *base, uint64_t **kc, uint64_t **kp, uint64_t **e,
                         uint64_t **a, uint64_t **b) {
    const size_t N = c->n, pq = c->pn + c->qn, D = c->dnum;
    Bump bp(base);
    uint64_t *p0 = bp.take(pq * N), *p1 = bp.take(pq * N), *p2 = bp.take(D * pq * N), *p3 = bp.take(D * pq * N),
             *p4 = bp.take(D * pq * N);
    if (kc) *kc = p0, *kp = p1, *e = p2, *a = p3, *b = p4;
    return bp.used;
}

extern "C" size_t kbfv_keygen_galois_workspace_bytes(const kbfv_context *c) {
    return c ? gal_layout(c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) : 0;
}

extern "C" int kbfv_keygen_galois(const kbfv_context *c, const uint8_t seed[32], const uint64_t *sk, uint64_t g,
                                  uint64_t *key, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && seed && sk && key && ws, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_keygen_galois_workspace_bytes(c), "workspace too small");
    const uint64_t m2 = 2 * (uint64_t)c->n;
    KB_REQUIRE((g & 1) && g < m2, "Galois element must be odd and below 2N");
    uint64_t *kc, *kp, *e, *a, *b;
    gal_layout(c, ws, &kc, &kp, &e, &a, &b);
    const long N = c->n;
    const int pn = c->pn, qn = c->qn, pq = pn + qn, D = c->dnum, B = c->ri_blocks, tn = c->tn;
    hipStream_t s = S(st);
    // key polynomial sigma_g^-1(s) over P u Q, NTT form
    KB_TRY(kb_ntt(c, -1, sk, kc, c->oP, pq, 1, 1, 0, (long)pq * N, N, 0, (long)pq * N, N, s));
    KB_TRY(launch_automorphism(c, inv_mod_2n(g, m2), kc, kp, c->oP, pq, 1, s));
    KB_TRY(kb_ntt_inplace(c, +1, kp, c->oP, pq, 1, s));
    KB_TRY(kb_sample(c, 3, seed, TAG(TAG_GAL_E) | g, 0, c->oP, pq, e, 0, D, s));
    KB_TRY(kb_ntt_inplace(c, +1, e, c->oP, pq, D, s));
    KB_TRY(kb_sample(c, 0, seed, TAG(TAG_GAL_A) | g, 0, c->oP, pq, a, 0, D, s));
    const long total = (long)D * pq * N;
    hipLaunchKernelGGL(k_ksk_combine, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, e, a, kp,
                       sk + (long)pn * N, b, pn, qn, c->n, total);
    KB_LAUNCH_CHECK();
    KB_TRY(kb_ntt_inplace(c, -1, a, c->oP, pq, D, s));
    KB_TRY(kb_ntt_inplace(c, -1, b, c->oP, pq, D, s));
    const long half = (long)D * B * tn * N;
    KB_TRY(kb_modup_pq_to_t(c, a, key, D, s));
    KB_TRY(kb_modup_pq_to_t(c, b, key + half, D, s));
    return kb_ntt_inplace(c, +1, key, c->oT, tn, 2 * D * B, s);
}

struct RotWs {
    uint64_t *X, *CT, *IP, *MD, *PQ;
};
static size_t rot_layout(const kbfv_context *c, int cb, void *base, RotWs *w) {
    const size_t N = c->n, qn = c->qn, pq = c->pn + c->qn;
    Bump bp(base);
    uint64_t *x = bp.take(2 * qn * N * cb), *ct = bp.take((size_t)c->dnum * c->tn * N * cb),
             *ip = bp.take(2 * (size_t)c->ri_blocks * c->tn * N * cb), *md = bp.take(2 * qn * N * cb);
    uint64_t *pqb = kb_modup_moddown_fused(c) ? nullptr : bp.take(2 * pq * N * cb);
    if (w) *w = RotWs{x, ct, ip, md, pqb};
    return bp.used;
}

extern "C" size_t kbfv_rotate_workspace_bytes(const kbfv_context *c, int batch) {
    return (c && batch > 0) ? rot_layout(c, batch, nullptr, nullptr) : 0;
}

// ct -> sigma_g(KS(a) + (0, b)): key switching from sigma_g^-1(s) to s on the
// coefficient-form a, then the automorphism, then back to the NTT domain
extern "C" int kbfv_rotate(const kbfv_context *c, const uint64_t *gkey, uint64_t g, const uint64_t *ct,
                           uint64_t *out, int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && gkey && ct && out && ws && batch > 0, "null argument");
    int cb = batch;
    while (cb > 1 && rot_layout(c, cb, nullptr, nullptr) > ws_bytes) cb = (cb + 1) / 2;
    KB_REQUIRE(rot_layout(c, cb, nullptr, nullptr) <= ws_bytes, "workspace too small for one ciphertext");
    RotWs w;
    rot_layout(c, cb, ws, &w);
    const long N = c->n, qn = c->qn, CTS = 2 * qn * N;
    hipStream_t s = S(st);
    for (int b0 = 0; b0 < batch; b0 += cb) {
        const int nb = std::min(cb, batch - b0);
        KB_TRY(kb_ntt(c, -1, ct + b0 * CTS, w.X, c->oQ, qn, 2, nb, CTS, qn * N, N, CTS, qn * N, N, s));
        KB_TRY(kb_digits_to_t_ntt(c, w.X, CTS, w.CT, nb, nullptr, s));
        KB_TRY(kb_inner_product(c, w.CT, gkey, w.IP, nb, s));
        KB_TRY(kb_ks_tail(c, w.IP, w.IP, w.MD, nullptr, 0, nb, w.PQ, s));
        for (int b = 0; b < nb; ++b)  // KS(a) + (0, b): b joins the second polynomial only
            KB_TRY(kb_poly_binop(c, 0, w.MD + b * CTS + qn * N, w.X + b * CTS + qn * N, w.MD + b * CTS + qn * N,
                                 c->oQ, qn, 1, s));
        KB_TRY(launch_automorphism(c, g, w.MD, w.X, c->oQ, qn, 2 * nb, s));
        KB_TRY(kb_ntt(c, +1, w.X, out + b0 * CTS, c->oQ, qn, 2, nb, CTS, qn * N, N, CTS, qn * N, N, s));
    }
    return KBFV_OK;
}

// ---------------------------------------------------------------- encoder
extern "C" int kbfv_encode(const kbfv_context *c, uint64_t *vals, int batch, void *st) {
    KB_REQUIRE(c && vals && batch > 0, "null argument");
    if (!c->batching) return KBFV_OK;  // coefficient encoding
    return kb_ntt_inplace(c, -1, vals, c->oPlain, 1, batch, S(st));
}
// ---------------------------------------------------------------- ciphertext arithmetic
// (BFVScheme::negate / multConstAndEqual / multConstAndAddCipherEqual /
// addConstAndEqual, include/BFVScheme.cuh:192-326)
extern "C" int kbfv_negate(const kbfv_context *c, const uint64_t *ct, uint64_t *out, int batch, void *st) {
    KB_REQUIRE(c && ct && out && batch > 0, "null argument");
    const long total = (long)batch * 2 * c->qn * c->n;
    hipLaunchKernelGGL(k_negate, dim3(blocks_for(total, BS)), dim3(BS), 0, S(st), c->t, ct, out, c->oQ, c->qn,
                       c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
extern "C" int kbfv_mul_scalar(const kbfv_context *c, const uint64_t *ct, uint64_t k, const uint64_t *add,
                               uint64_t *out, int batch, void *st) {
    KB_REQUIRE(c && ct && out && batch > 0, "null argument");
    const long total = (long)batch * 2 * c->qn * c->n;
    if (add)
        hipLaunchKernelGGL(k_mul_scalar<true>, dim3(blocks_for(total, BS)), dim3(BS), 0, S(st), c->t, ct, k, add,
                           out, c->oQ, c->qn, c->n, total);
    else
        hipLaunchKernelGGL(k_mul_scalar<false>, dim3(blocks_for(total, BS)), dim3(BS), 0, S(st), c->t, ct, k,
                           nullptr, out, c->oQ, c->qn, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
extern "C" size_t kbfv_add_plain_workspace_bytes(const kbfv_context *c, int batch) {
    return (c && batch > 0) ? (size_t)batch * c->qn * c->n * sizeof(uint64_t) : 0;
}
extern "C" int kbfv_add_plain(const kbfv_context *c, const uint64_t *ct, const uint64_t *m, uint64_t *out,
                              int batch, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && ct && m && out && ws && batch > 0, "null argument");
    KB_REQUIRE(ws_bytes >= kbfv_add_plain_workspace_bytes(c, batch), "workspace too small");
    const long N = c->n, L = (long)c->qn * N;
    uint64_t *dm = (uint64_t *)ws;
    hipStream_t s = S(st);
    KB_TRY(kb_delta_m(c, m, dm, batch, s));
    KB_TRY(kb_ntt_inplace(c, +1, dm, c->oQ, c->qn, batch, s));
    for (int b = 0; b < batch; ++b) {  // a stays, b += NTT(Delta m)
        if (out != ct) KB_HIP(hipMemcpyAsync(out + 2 * L * b, ct + 2 * L * b, L * 8, hipMemcpyDeviceToDevice, s));
        KB_TRY(kb_poly_binop(c, 0, ct + 2 * L * b + L, dm + L * b, out + 2 * L * b + L, c->oQ, c->qn, 1, s));
    }
    return KBFV_OK;
}

// SEAL slot order (BFVContext::preComputeIndex, include/BFVcontext.cuh:102-129;
// encode_BFV_kernel / decode_BFV_kernel, include/Encoder.cuh:6-21)
extern "C" int kbfv_encode_seal(const kbfv_context *c, uint64_t *vals, uint64_t *tmp, int batch, void *st) {
    KB_REQUIRE(c && vals && tmp && batch > 0 && vals != tmp, "null or aliased argument");
    KB_REQUIRE(c->batching && c->t.seal_map, "the SEAL slot order needs t = 1 mod 2N");
    const long total = (long)batch * c->n;
    hipLaunchKernelGGL(k_slot_permute<true>, dim3(blocks_for(total, BS)), dim3(BS), 0, S(st), vals, tmp,
                       c->t.seal_map, c->log_n, total);
    KB_LAUNCH_CHECK();
    const long N = c->n;
    return kb_ntt(c, -1, tmp, vals, c->oPlain, 1, 1, batch, N, 0, 0, N, 0, 0, S(st));
}
extern "C" int kbfv_decode_seal(const kbfv_context *c, uint64_t *vals, uint64_t *tmp, int batch, void *st) {
    KB_REQUIRE(c && vals && tmp && batch > 0 && vals != tmp, "null or aliased argument");
    KB_REQUIRE(c->batching && c->t.seal_map, "the SEAL slot order needs t = 1 mod 2N");
    const long total = (long)batch * c->n, N = c->n;
    KB_TRY(kb_reduce_plain(c, vals, total, S(st)));
    KB_TRY(kb_ntt(c, +1, vals, tmp, c->oPlain, 1, 1, batch, N, 0, 0, N, 0, 0, S(st)));
    hipLaunchKernelGGL(k_slot_permute<false>, dim3(blocks_for(total, BS)), dim3(BS), 0, S(st), tmp, vals,
                       c->t.seal_map, c->log_n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
extern "C" int kbfv_decode(const kbfv_context *c, uint64_t *vals, int batch, void *st) {
    KB_REQUIRE(c && vals && batch > 0, "null argument");
    KB_TRY(kb_reduce_plain(c, vals, (long)batch * c->n, S(st)));
    if (!c->batching) return KBFV_OK;
    return kb_ntt_inplace(c, +1, vals, c->oPlain, 1, batch, S(st));
}

// ---------------------------------------------------------------- RGSW x RLWE
static int ndigits(const kbfv_context *c, int base_bits) {
    int maxb = 0;
    for (int i = 0; i < c->qn; ++i) {
        uint64_t q = c->mod[c->oQ + i];
        int b = 0;
        while (q) b++, q >>= 1;
        maxb = std::max(maxb, b);
    }
    return (maxb + base_bits - 1) / base_bits;
}

extern "C" int kbfv_rgsw_rows(const kbfv_context *c, int base_bits) {
    KB_REQUIRE(c && base_bits >= 1 && base_bits <= 62, "bad gadget base");
    return 2 * c->qn * ndigits(c, base_bits);
}

static size_t rgsw_enc_layout(const kbfv_context *c, int rows, void *base, uint64_t **e, uint64_t **mq) {
    const size_t L = (size_t)c->qn * c->n;
    Bump bp(base);
    uint64_t *p0 = bp.take(L * rows), *p1 = bp.take(L);
    if (e) *e = p0, *mq = p1;
    return bp.used;
}

extern "C" size_t kbfv_rgsw_encrypt_workspace_bytes(const kbfv_context *c) {
    // sized for the finest gadget (1-bit digits)
    return c ? rgsw_enc_layout(c, 2 * c->qn * 64, nullptr, nullptr, nullptr) : 0;
}

extern "C" int kbfv_rgsw_encrypt(const kbfv_context *c, const uint8_t seed[32], const uint64_t *sk,
                                 const uint64_t *m_coeff, int base_bits, uint64_t *rgsw, void *ws, size_t ws_bytes,
                                 void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && seed && sk && m_coeff && rgsw && ws, "null argument");
    const int rows = kbfv_rgsw_rows(c, base_bits);
    KB_REQUIRE(rows > 0, "bad gadget base");
    KB_REQUIRE(ws_bytes >= rgsw_enc_layout(c, rows, nullptr, nullptr, nullptr), "workspace too small");
    const int nd = ndigits(c, base_bits), qn = c->qn;
    const long N = c->n, L = (long)qn * N;
    uint64_t *e, *mq;
    rgsw_enc_layout(c, rows, ws, &e, &mq);
    hipStream_t s = S(st);
    // row r: a_r uniform into component 0, b_r = e_r - a_r*s into component 1
    KB_TRY(kb_sample(c, 0, seed, TAG(TAG_RGSW_A), 0, c->oQ, qn, rgsw, 2 * L, rows, s));
    KB_TRY(kb_sample(c, 3, seed, TAG(TAG_RGSW_E), 0, c->oQ, qn, e, L, rows, s));
    KB_TRY(kb_ntt_inplace(c, +1, e, c->oQ, qn, rows, s));
    const long total = (long)rows * L;
    hipLaunchKernelGGL(k_rgsw_body, dim3(blocks_for(total, BS)), dim3(BS), 0, s, c->t, e,
                       sk + (long)c->pn * N, rgsw, qn, c->oQ, c->n, total);
    KB_LAUNCH_CHECK();
    // gadget term m * 2^(w*j) on limb i of component h
    hipLaunchKernelGGL(k_lift_plain, dim3(blocks_for(L, BS)), dim3(BS), 0, s, c->t, m_coeff, mq, c->oQ, c->plain,
                       c->n, L);
    KB_LAUNCH_CHECK();
    KB_TRY(kb_ntt_inplace(c, +1, mq, c->oQ, qn, 1, s));
    return kb_rgsw_add_gadget(c, rgsw, mq, base_bits, nd, s);
}

extern "C" size_t kbfv_external_product_workspace_bytes(const kbfv_context *c, int batch, int base_bits) {
    if (!c || batch <= 0) return 0;
    const size_t L = (size_t)c->qn * c->n;
    Bump bp(nullptr);
    bp.take(2 * L * batch);  // coefficient form of the input
    bp.take(8 * L * batch);  // digit chunk (up to 8 rows)
    (void)base_bits;
    return bp.used;
}

extern "C" int kbfv_external_product(const kbfv_context *c, const uint64_t *rgsw, const uint64_t *ct, uint64_t *out,
                                     int batch, int base_bits, void *ws, size_t ws_bytes, void *st) {
    KB_REQUIRE(c && !c->parent, "this operation needs the full-level context");
    KB_REQUIRE(c && rgsw && ct && out && ws && batch > 0, "null argument");
    const int rows = kbfv_rgsw_rows(c, base_bits);
    KB_REQUIRE(rows > 0, "bad gadget base");
    KB_REQUIRE(ws_bytes >= kbfv_external_product_workspace_bytes(c, batch, base_bits), "workspace too small");
    const int nd = ndigits(c, base_bits), qn = c->qn;
    const long N = c->n, L = (long)qn * N;
    Bump bp(ws);
    uint64_t *coef = bp.take(2 * L * batch), *dig = bp.take(8 * L * batch);
    hipStream_t s = S(st);
    KB_TRY(kb_ntt(c, -1, ct, coef, c->oQ, qn, 2, batch, 2 * L, L, N, 2 * L, L, N, s));
    KB_TRY(fill_zero(out, 2 * L * batch, s));
    // digits below 8 q_min are valid lazy NTT inputs: split them inside the
    // NTT's first pass instead of writing them to every limb first
    uint64_t qmin = ~0ULL;
    for (int i = 0; i < qn; ++i) qmin = std::min(qmin, c->mod[c->oQ + i]);
    const bool fused = base_bits <= 62 && ((1ULL << base_bits) >> 3) <= qmin;
    for (int r0 = 0; r0 < rows; r0 += 8) {
        const int rc = std::min(8, rows - r0);
        if (fused) {
            KB_TRY(kb_digit_ntt(c, coef, dig, batch, base_bits, nd, r0, rc, s));
        } else {
            KB_TRY(kb_digit_decompose(c, coef, dig, batch, base_bits, nd, r0, rc, s));
            KB_TRY(kb_ntt(c, +1, dig, dig, c->oQ, qn, rc, batch, (long)rc * L, L, N, (long)rc * L, L, N, s));
        }
        KB_TRY(kb_gadget_mac(c, dig, rgsw + (long)r0 * 2 * L, out, batch, rc, s));
    }
    return KBFV_OK;
}
