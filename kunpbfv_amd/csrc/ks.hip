// Hybrid key switching in the auxiliary basis T (the reference's "23" method).
//
// Replaces modUpPQtoT_23_kernel (include/BasisConv.cuh:11-59),
// modUpQjtoT_23_kernel (:86-135), external_product_T_kernel and its
// _swk_reuse variant (include/ExternalProduct.cuh:11-177), modUpTtoPQl_23_kernel
// (include/BasisConv.cuh:159-213) and the modDownPQltoQl_23 kernels (:233-298).
// The float overflow estimates (`vv += (float)x / (float)m`, truncation, and
// the 128-bit wrap of `acc - vv * (M mod m)`) are kept operation for operation;
// the surrounding arithmetic is exact with canonical outputs.
//
// MI355X design: with one special prime (K = 1) the digit decomposition of a
// limb is the limb itself, so "modUp Qj -> T then NTT over T" is one NTT that
// reads the Q limb directly (kb_digits_to_t_ntt).  The inner product is
// key-stationary: each thread holds its key column (2 x dnum values) in
// registers and walks the whole batch, so the relinearisation key is read from
// HBM once per call instead of once per ciphertext.
#include "kbfv_internal.h"
#include "kbfv_ops.h"

namespace {

constexpr int KS_BS = 256;
constexpr int KS_MAXG = 16;  // max gamma / tn held in registers

inline unsigned blocks_for(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }

// x: [batch][qn][N] -> out: [batch][dnum][tn][N]
__global__ void k_modup_qj_to_t(kb_dev_tables t, const uint64_t *__restrict__ x, long xs,
                                uint64_t *__restrict__ out, int pn, int qn, int tn, int dnum, int oQ,
                                int oT, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n);
    long rest = i / n;
    const int b = (int)(rest % dnum);
    const long bt = rest / dnum;
    const uint64_t *xp = x + bt * xs + k;
    uint64_t rb[KS_MAXG];
    float vv = 0;
    int cnt = 0;
    for (int j = 0; j < pn && b * pn + j < qn; ++j, ++cnt) {
        const int qi = b * pn + j;
        const uint64_t q = t.mod[oQ + qi];
        const kbfv_tw h = t.qj_hatinv[b * pn + j];
        rb[j] = kb_shoup(xp[(long)qi * n], h.w, h.ws, q);
        const float rr = (float)rb[j];
        vv += rr / (float)q;
    }
    const int vf = (int)vv;
    uint64_t *op = out + (bt * dnum + b) * (long)tn * n + k;
    for (int kk = 0; kk < tn; ++kk) {
        const int m = oT + kk;
        kbfv_u128 acc{0, 0};
        for (int j = 0; j < cnt; ++j) kb_mac128(acc, rb[j], t.qj_hat_mod_t[(b * tn + kk) * pn + j]);
        kb_sub128(acc, (uint64_t)vf * t.qj_mod_t[b * tn + kk]);
        op[(long)kk * n] = kb_barrett128(acc, t.mod[m], t.mu[2 * m], t.mu[2 * m + 1]);
    }
}

// x: [batch][pn+qn][N] -> out: [batch][ri_blocks][tn][N]
__global__ void k_modup_pq_to_t(kb_dev_tables t, const uint64_t *__restrict__ x,
                                uint64_t *__restrict__ out, int pq, int tn, int gamma, int B, int oT,
                                int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n);
    long rest = i / n;
    const int b = (int)(rest % B);
    const long bt = rest / B;
    const uint64_t *xp = x + bt * (long)pq * n + k;
    uint64_t rb[KS_MAXG];
    float vv = 0;
    int cnt = 0;
    for (int j = 0; j < gamma && b * gamma + j < pq; ++j, ++cnt) {
        const int idx = b * gamma + j;
        const uint64_t r = t.mod[idx];  // P then Q are the first moduli
        const kbfv_tw h = t.ri_hatinv[idx];
        rb[j] = kb_shoup(xp[(long)idx * n], h.w, h.ws, r);
        vv += (float)rb[j] / (float)r;
    }
    const int vf = (int)vv;
    uint64_t *op = out + (bt * B + b) * (long)tn * n + k;
    for (int kk = 0; kk < tn; ++kk) {
        const int m = oT + kk;
        kbfv_u128 acc{0, 0};
        for (int j = 0; j < cnt; ++j) kb_mac128(acc, rb[j], t.ri_hat_mod_t[(b * tn + kk) * gamma + j]);
        kb_sub128(acc, (uint64_t)vf * t.ri_mod_t[kk * B + b]);
        op[(long)kk * n] = kb_barrett128(acc, t.mod[m], t.mu[2 * m], t.mu[2 * m + 1]);
    }
}

// Key-stationary inner product.  cin [batch][D][tn][N], key [2][KD][KB][tn][N]
// (read on its first D digits and B blocks: a ciphertext below the full level
// uses a prefix of the full-level key), out [batch][2][B][tn][N].  A workgroup owns one T limb, 64 coefficients and
// up to IP_WAVES key blocks (one wave per block b, both key halves in
// registers).  It walks the whole batch: per ciphertext the D x 64 digit tile
// is staged once in LDS and read by every wave, so the key is read from HBM
// once per call and each digit tile ceil(B / IP_WAVES) times.
constexpr int IP_WAVES = 8;
template <int MAXD>
__global__ void __launch_bounds__(64 * IP_WAVES)
    k_inner_product(kb_dev_tables t, const uint64_t *__restrict__ cin, const uint64_t *__restrict__ key,
                    uint64_t *__restrict__ out, int D, int B, int KD, int KB, int tn, int oT, int n,
                    int batch) {
    __shared__ uint64_t cs[MAXD][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int bgroups = (B + IP_WAVES - 1) / IP_WAVES;
    const int tiles = n / 64;
    int bid = blockIdx.x;
    const int bg = bid % bgroups;
    bid /= bgroups;
    const int tile = bid % tiles;
    const int kk = bid / tiles;
    const int b = bg * IP_WAVES + wave;
    const bool active = b < B;
    const int k = tile * 64 + lane;
    const int m = oT + kk;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const long N = n;
    const long kstride = (long)KB * tn * N;  // digit stride in the key
    uint64_t ka[MAXD], kbv[MAXD];
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
        ka[i] = 0;
        kbv[i] = 0;
        if (active && i < D) {
            const long off = (long)i * kstride + ((long)b * tn + kk) * N + k;
            ka[i] = key[off];
            kbv[i] = key[(long)KD * kstride + off];
        }
    }
    // each thread stages at most ceil(MAXD*64 / (64*IP_WAVES)) digit values per ciphertext
    constexpr int PER = (MAXD + IP_WAVES - 1) / IP_WAVES;
    uint64_t pre[PER];
    auto fetch = [&](int bt) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = wave + r * IP_WAVES;
            if (i < D) pre[r] = cin[(((long)bt * D + i) * tn + kk) * N + tile * 64 + lane];
        }
    };
    fetch(0);
    for (int bt = 0; bt < batch; ++bt) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = wave + r * IP_WAVES;
            if (i < D) cs[i][lane] = pre[r];
        }
        __syncthreads();
        if (bt + 1 < batch) fetch(bt + 1);
        if (active) {
            kbfv_u128 a0{0, 0}, a1{0, 0};
#pragma unroll
            for (int i = 0; i < MAXD; ++i) {
                if (i < D) {
                    const uint64_t cv = cs[i][lane];
                    kb_mac128(a0, cv, ka[i]);
                    kb_mac128(a1, cv, kbv[i]);
                }
            }
            uint64_t *op = out + (long)bt * 2 * B * tn * N + ((long)b * tn + kk) * N + k;
            op[0] = kb_barrett128(a0, q, mlo, mhi);
            op[(long)B * tn * N] = kb_barrett128(a1, q, mlo, mhi);
        }
        __syncthreads();
    }
}

// fallback for dnum > 64: one thread per output, key re-read per ciphertext
__global__ void k_inner_product_generic(kb_dev_tables t, const uint64_t *__restrict__ cin,
                                        const uint64_t *__restrict__ key, uint64_t *__restrict__ out,
                                        int D, int B, int KD, int KB, int tn, int oT, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long N = n;
    const int k = (int)(i % n);
    long rest = i / n;
    const int kk = (int)(rest % tn);
    rest /= tn;
    const int b = (int)(rest % B);
    rest /= B;
    const int h = (int)(rest % 2);
    const long bt = rest / 2;
    const int m = oT + kk;
    kbfv_u128 acc{0, 0};
    for (int d = 0; d < D; ++d)
        kb_mac128(acc, cin[((bt * D + d) * tn + kk) * N + k],
                  key[(((long)h * KD + d) * KB + b) * tn * N + (long)kk * N + k]);
    out[i] = kb_barrett128(acc, t.mod[m], t.mu[2 * m], t.mu[2 * m + 1]);
}

// in [batch][2][B][tn][N] coefficient form -> out [batch][2][pq][N]
// grid (n / KS_BS, B, batch * 2): coefficient, block b, ciphertext-poly (bt*2 + h)
__global__ void k_modup_t_to_pq(kb_dev_tables t, const uint64_t *__restrict__ in,
                                uint64_t *__restrict__ out, int pq, int tn, int gamma, int B, int oT,
                                int n) {
    const int k = blockIdx.x * KS_BS + threadIdx.x;
    const int b = blockIdx.y;
    const long rest = blockIdx.z;
    const uint64_t *ip = in + (rest * B + b) * (long)tn * n + k;
    uint64_t rb[KS_MAXG];
    float vv = 0;
    for (int kk = 0; kk < tn; ++kk) {
        const uint64_t tm = t.mod[oT + kk];
        const uint64_t ra = ip[(long)kk * n] + t.halfT_t[kk];
        const kbfv_tw h = t.t_hatinv[kk];
        rb[kk] = kb_shoup(ra, h.w, h.ws, tm);
        const float rr = (float)rb[kk];
        vv += rr / (float)tm;
    }
    const int vf = (int)vv;
    uint64_t *op = out + rest * (long)pq * n + k;
    for (int j = 0; j < gamma && b * gamma + j < pq; ++j) {
        const int idx = b * gamma + j;
        const uint64_t r = t.mod[idx];
        kbfv_u128 acc{0, 0};
        for (int kk = 0; kk < tn; ++kk) kb_mac128(acc, rb[kk], t.t_hat_mod_pq[idx * tn + kk]);
        kb_sub128(acc, (uint64_t)vf * t.t_mod_pq[idx]);
        uint64_t v = kb_barrett128(acc, r, t.mu[2 * idx], t.mu[2 * idx + 1]);
        v = kb_csub(v + r - t.halfT_mod[idx], r);
        op[(long)idx * n] = v;
    }
}

// in [batch][2][pq][N] -> out [batch][2][qn][N] (+ add [batch](2 polys at add_stride))
// grid (n / KS_BS, qn, batch * 2): coefficient, limb l, ciphertext-poly (bt*2 + h)
__global__ void k_moddown(kb_dev_tables t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                          const uint64_t *__restrict__ add, long add_stride, int pn, int qn, int oQ,
                          int n) {
    const int k = blockIdx.x * KS_BS + threadIdx.x;
    const int l = blockIdx.y;
    const long rest = blockIdx.z;
    const int h = (int)(rest & 1);
    const long bt = rest >> 1;
    const int pq = pn + qn;
    const uint64_t *ip = in + rest * (long)pq * n + k;
    const int m = oQ + l;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    uint64_t ra;
    if (pn == 1) {
        ra = kb_barrett128(kbfv_u128{ip[0], 0}, q, mlo, mhi);
    } else {
        kbfv_u128 acc{0, 0};
        for (int j = 0; j < pn; ++j) {
            const kbfv_tw hi = t.p_hatinv[j];
            const uint64_t rb = kb_shoup(ip[(long)j * n], hi.w, hi.ws, t.mod[j]);
            kb_mac128(acc, rb, t.p_hat_mod_q[l * pn + j]);
        }
        ra = kb_barrett128(acc, q, mlo, mhi);
    }
    ra = q - ra + ip[(long)(pn + l) * n];
    const kbfv_tw pi = t.Pinv_mod_q[l];
    uint64_t v = kb_shoup(ra, pi.w, pi.ws, q);
    if (add) v = kb_addmod(v, add[bt * add_stride + ((long)h * qn + l) * n + k], q);
    out[(rest * qn + l) * (long)n + k] = v;
}

// modUp T -> PQ and modDown PQ -> Q fused (the two kernels above, composed):
// in [batch][2][ri_blocks][tn][N] coeff -> out [batch][2][qn][N] (+ add), with
// one thread per (coefficient, ciphertext-poly).  The PQ limbs come out of the
// ri blocks in index order, P limbs first, so the thread keeps the PN special
// residues in registers and finishes each Q limb's modDown as soon as its
// modUp value exists: the [2][pn+qn] PQ polynomials never reach memory.  The
// arithmetic of every limb is that of k_modup_t_to_pq then k_moddown.  TN, G
// and PN (tn, gamma, pn) are compile-time so that the per-thread arrays stay
// in registers.  NOSC: the input is N times the INTT (lazy, below 4t); the
// first step [(x + T/2) (T/t)^-1]_t is a Shoup product exact for any 64-bit
// operand, so it takes ([N T/2]_t, [(T/t)^-1 n^-1]_t) instead and gives the
// same canonical value.  grid (n / KS_BS, batch * 2)
template <int TN, int G, int PN, bool NOSC>
__global__ void k_modup_moddown(kb_dev_tables t, const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                const uint64_t *__restrict__ add, long add_stride, int qn, int B, int oQ,
                                int oT, int n) {
    const int k = blockIdx.x * KS_BS + threadIdx.x;
    const long rest = blockIdx.y;
    const int h = (int)(rest & 1);
    const long bt = rest >> 1;
    const int pq = PN + qn;
    uint64_t xp[PN];  // the P residues of this coefficient
    // the T values of ri block b + 1 are loaded while block b is converted
    const uint64_t *ip0 = in + rest * B * (long)TN * n + k;
    uint64_t nxt[TN];
#pragma unroll
    for (int kk = 0; kk < TN; ++kk) nxt[kk] = ip0[(long)kk * n];
    for (int b = 0; b < B; ++b) {
        uint64_t cur[TN];
#pragma unroll
        for (int kk = 0; kk < TN; ++kk) cur[kk] = nxt[kk];
        if (b + 1 < B) {
#pragma unroll
            for (int kk = 0; kk < TN; ++kk) nxt[kk] = ip0[((long)(b + 1) * TN + kk) * n];
        }
        uint64_t rb[TN];
        float vv = 0;
#pragma unroll
        for (int kk = 0; kk < TN; ++kk) {
            const uint64_t tm = t.mod[oT + kk];
            const uint64_t ra = cur[kk] + (NOSC ? t.halfT_n[kk] : t.halfT_t[kk]);
            const kbfv_tw hw = NOSC ? t.t_hatinv_n[kk] : t.t_hatinv[kk];
            rb[kk] = kb_shoup(ra, hw.w, hw.ws, tm);
            const float rr = (float)rb[kk];
            vv += rr / (float)tm;
        }
        const int vf = (int)vv;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int idx = b * G + j;
            if (idx >= pq) break;
            const uint64_t r = t.mod[idx];
            kbfv_u128 acc{0, 0};
#pragma unroll
            for (int kk = 0; kk < TN; ++kk) kb_mac128(acc, rb[kk], t.t_hat_mod_pq[idx * TN + kk]);
            kb_sub128(acc, (uint64_t)vf * t.t_mod_pq[idx]);
            uint64_t v = kb_barrett128(acc, r, t.mu[2 * idx], t.mu[2 * idx + 1]);
            v = kb_csub(v + r - t.halfT_mod[idx], r);
            if (idx < PN) {
#pragma unroll
                for (int jj = 0; jj < PN; ++jj)
                    if (jj == idx) xp[jj] = v;
                continue;
            }
            // modDown of Q limb l from its modUp value v and the P residues
            const int l = idx - PN;
            const int m = oQ + l;
            const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
            uint64_t ra;
            if constexpr (PN == 1) {
                ra = kb_barrett128(kbfv_u128{xp[0], 0}, q, mlo, mhi);
            } else {
                kbfv_u128 a2{0, 0};
#pragma unroll
                for (int jj = 0; jj < PN; ++jj) {
                    const kbfv_tw hi = t.p_hatinv[jj];
                    const uint64_t rp = kb_shoup(xp[jj], hi.w, hi.ws, t.mod[jj]);
                    kb_mac128(a2, rp, t.p_hat_mod_q[l * PN + jj]);
                }
                ra = kb_barrett128(a2, q, mlo, mhi);
            }
            ra = q - ra + v;
            const kbfv_tw pi = t.Pinv_mod_q[l];
            uint64_t o = kb_shoup(ra, pi.w, pi.ws, q);
            if (add) o = kb_addmod(o, add[bt * add_stride + ((long)h * qn + l) * n + k], q);
            out[(rest * qn + l) * (long)n + k] = o;
        }
    }
}

}  // namespace

bool kb_modup_moddown_fused(const kbfv_context *c) {
    return c->pn == 1 && ((c->tn == 4 && c->gamma == 2) || (c->tn == 3 && c->gamma == 1));
}

static int modup_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
                         long add_stride, int batch, uint64_t *pq_tmp, bool nosc, hipStream_t st) {
    KB_REQUIRE(c && in && out && batch > 0, "bad modUp/modDown arguments");
    KB_REQUIRE(2L * batch <= 65535, "batch too large for one modUp/modDown launch");
    // PN special residues first: the P limbs must all lie before any Q limb
#define KB_MUMD(TN_, G_, PN_, NOSC_)                                                                       \
    if (c->tn == TN_ && c->gamma == G_ && c->pn == PN_ && nosc == NOSC_) {                                 \
        hipLaunchKernelGGL((k_modup_moddown<TN_, G_, PN_, NOSC_>), dim3(c->n / KS_BS, 2 * batch), dim3(KS_BS), \
                           0, st, c->t, in, out, add, add_stride, c->qn, c->ri_blocks, c->oQ, c->oT, c->n); \
        KB_LAUNCH_CHECK();                                                                                 \
        return KBFV_OK;                                                                                    \
    }
    KB_MUMD(4, 2, 1, false)
    KB_MUMD(3, 1, 1, false)
    KB_MUMD(4, 2, 1, true)
    KB_MUMD(3, 1, 1, true)
#undef KB_MUMD
    KB_REQUIRE(!nosc, "unscaled modUp input needs a fused shape");
    // other shapes: the two kernels through the PQ buffer
    KB_REQUIRE(pq_tmp, "modUp/modDown of this shape needs a PQ buffer");
    const int rc = kb_modup_t_to_pq(c, in, pq_tmp, batch, st);
    if (rc != KBFV_OK) return rc;
    return kb_moddown(c, pq_tmp, out, add, add_stride, batch, st);
}

int kb_modup_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
                     long add_stride, int batch, uint64_t *pq_tmp, hipStream_t st) {
    return modup_moddown(c, in, out, add, add_stride, batch, pq_tmp, false, st);
}

int kb_ks_tail(const kbfv_context *c, const uint64_t *ip, uint64_t *ip_coeff, uint64_t *out, const uint64_t *add,
               long add_stride, int batch, uint64_t *pq_tmp, hipStream_t st) {
    KB_REQUIRE(c && ip && ip_coeff && out && batch > 0, "bad key-switch arguments");
    const long N = c->n, TS = (long)c->tn * N;
    const int polys = 2 * c->ri_blocks * batch;
    const bool nosc = kb_modup_moddown_fused(c);
    const int rc = nosc ? kb_intt_noscale(c, ip, ip_coeff, c->oT, c->tn, polys, 1, 0, TS, N, 0, TS, N, st)
                        : kb_ntt(c, -1, ip, ip_coeff, c->oT, c->tn, polys, 1, 0, TS, N, 0, TS, N, st);
    if (rc) return rc;
    return modup_moddown(c, ip_coeff, out, add, add_stride, batch, pq_tmp, nosc, st);
}

int kb_modup_qj_to_t(const kbfv_context *c, const uint64_t *x, long xs, uint64_t *out, int batch,
                     hipStream_t st) {
    KB_REQUIRE(c && x && out && batch > 0, "bad modUp arguments");
    KB_REQUIRE(c->pn <= KS_MAXG, "too many special primes");
    const long total = (long)batch * c->dnum * c->n;
    hipLaunchKernelGGL(k_modup_qj_to_t, dim3(blocks_for(total, KS_BS)), dim3(KS_BS), 0, st, c->t, x, xs, out,
                       c->pn, c->qn, c->tn, c->dnum, c->oQ, c->oT, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_modup_pq_to_t(const kbfv_context *c, const uint64_t *x, uint64_t *out, int batch, hipStream_t st) {
    KB_REQUIRE(c && x && out && batch > 0, "bad modUp arguments");
    KB_REQUIRE(c->gamma <= KS_MAXG, "gamma too large");
    const long total = (long)batch * c->ri_blocks * c->n;
    hipLaunchKernelGGL(k_modup_pq_to_t, dim3(blocks_for(total, KS_BS)), dim3(KS_BS), 0, st, c->t, x, out,
                       c->pn + c->qn, c->tn, c->gamma, c->ri_blocks, c->oT, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_inner_product(const kbfv_context *c, const uint64_t *cin, const uint64_t *key, uint64_t *out,
                     int batch, hipStream_t st) {
    KB_REQUIRE(c && cin && key && out && batch > 0, "bad inner-product arguments");
    if (batch >= 16) {  // enough ciphertexts to fill the matrix-core column tiles
        const int r = kb_inner_product_mfma(c, cin, key, out, batch, st);
        if (r != 1) return r;
    }
    const int D = c->dnum, B = c->ri_blocks, tn = c->tn, n = c->n;
    const unsigned grid = (unsigned)((n / 64) * tn * ((B + IP_WAVES - 1) / IP_WAVES));
#define KB_IP(MD)                                                                                        \
    hipLaunchKernelGGL(k_inner_product<MD>, dim3(grid), dim3(64 * IP_WAVES), 0, st, c->t, cin, key, out, D, \
                       B, c->key_dnum, c->key_blocks, tn, c->oT, n, batch)
    if (D <= 4)
        KB_IP(4);
    else if (D <= 8)
        KB_IP(8);
    else if (D <= 16)
        KB_IP(16);
    else if (D <= 32)
        KB_IP(32);
    else {
        const long total = (long)batch * 2 * B * tn * n;
        hipLaunchKernelGGL(k_inner_product_generic, dim3(blocks_for(total, KS_BS)), dim3(KS_BS), 0, st, c->t,
                           cin, key, out, D, B, c->key_dnum, c->key_blocks, tn, c->oT, n, total);
    }
#undef KB_IP
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_modup_t_to_pq(const kbfv_context *c, const uint64_t *in, uint64_t *out, int batch, hipStream_t st) {
    KB_REQUIRE(c && in && out && batch > 0, "bad modUp arguments");
    KB_REQUIRE(c->tn <= KS_MAXG, "too many T primes");
    KB_REQUIRE(2L * batch <= 65535, "batch too large for one modUp launch");
    hipLaunchKernelGGL(k_modup_t_to_pq, dim3(c->n / KS_BS, c->ri_blocks, 2 * batch), dim3(KS_BS), 0, st, c->t,
                       in, out, c->pn + c->qn, c->tn, c->gamma, c->ri_blocks, c->oT, c->n);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_moddown(const kbfv_context *c, const uint64_t *in, uint64_t *out, const uint64_t *add,
               long add_stride, int batch, hipStream_t st) {
    KB_REQUIRE(c && in && out && batch > 0, "bad modDown arguments");
    KB_REQUIRE(2L * batch <= 65535, "batch too large for one modDown launch");
    hipLaunchKernelGGL(k_moddown, dim3(c->n / KS_BS, c->qn, 2 * batch), dim3(KS_BS), 0, st, c->t, in, out, add,
                       add_stride, c->pn, c->qn, c->oQ, c->n);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_digits_to_t_ntt(const kbfv_context *c, const uint64_t *x_coeff, long x_stride, uint64_t *out,
                       int batch, uint64_t *tmp, hipStream_t st) {
    const long N = c->n;
    const int D = c->dnum, tn = c->tn;
    if (c->digits_direct_t) {
        // digit d is Q limb d itself: NTT over each T prime straight from the Q
        // limb (x < q_d < 8 t_j is a valid lazy input of the forward butterfly);
        // other chains reduce the digit into T first (k_modup_qj_to_t)
        return kb_ntt(c, +1, x_coeff, out, c->oT, tn, D, batch, x_stride, N, 0, (long)D * tn * N,
                      (long)tn * N, N, st);
    }
    (void)tmp;
    int rc = kb_modup_qj_to_t(c, x_coeff, x_stride, out, batch, st);
    if (rc) return rc;
    return kb_ntt(c, +1, out, out, c->oT, tn, D, batch, (long)D * tn * N, (long)tn * N, N,
                  (long)D * tn * N, (long)tn * N, N, st);
}
