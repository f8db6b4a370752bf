// RGSW x RLWE external product (gadget base 2^w over each Q limb).
//
// No counterpart in the reference: its "external_product_T" is the key-switch
// inner product (include/ExternalProduct.cuh:11), which this engine implements
// in ks.hip.  Gadget: row (h, i, j) of an RGSW encryption of m carries
// m * 2^(w*j) on limb i of component h (0 = a, 1 = b) and 0 on the other limbs;
// the RLWE input (a, b) is decomposed into the base-2^w digits of each Q
// residue, d_{h,i,j} = (x_{h,i} >> w*j) mod 2^w, the same small integer on every
// limb.  out = sum_{h,i,j} NTT(d_{h,i,j}) * row_{h,i,j}.
//
// The MAC is row-stationary: a thread keeps its rows' RGSW values in registers
// and walks the batch, so one RGSW (shared by the batch) is read once per call.
#include "kbfv_internal.h"
#include "kbfv_ops.h"

namespace {

constexpr int XP_BS = 256;
constexpr int XP_MAXRC = 8;  // rows per MAC call

inline unsigned blocks_for(long total, int bs) { return (unsigned)((total + bs - 1) / bs); }

// coef [batch][2][qn][N] -> dig [batch][rc][qn][N] for rows r0 .. r0+rc-1
__global__ void k_digit_decompose(kb_dev_tables t, const uint64_t *__restrict__ coef,
                                  uint64_t *__restrict__ dig, int qn, int oQ, int nd, int bb, int r0,
                                  int rc, int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = (int)(i % n);
    long rest = i / n;
    const int r = (int)(rest % rc);
    const long bt = rest / rc;
    const int row = r0 + r;
    const int h = row / (qn * nd), li = (row / nd) % qn, j = row % nd;
    const uint64_t x = coef[(bt * 2 + h) * (long)qn * n + (long)li * n + k];
    const uint64_t mask = (bb >= 64) ? ~0ULL : ((1ULL << bb) - 1);
    const uint64_t d = (x >> (bb * j)) & mask;
    uint64_t *o = dig + (bt * rc + r) * (long)qn * n + k;
    for (int l = 0; l < qn; ++l) {
        const uint64_t q = t.mod[oQ + l];
        o[(long)l * n] = d < q ? d : d % q;
    }
}

// out [batch][2][qn][N] += sum_r dig[batch][r][qn][N] * rgsw[r0+r][2][qn][N]
template <int RC>
__global__ void __launch_bounds__(XP_BS)
    k_gadget_mac(kb_dev_tables t, const uint64_t *__restrict__ dig, const uint64_t *__restrict__ rgsw,
                 uint64_t *__restrict__ out, int qn, int oQ, int n, int batch, int rc, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;  // total = qn * n
    const int k = (int)(i % n);
    const int l = (int)(i / n);
    const int m = oQ + l;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    const long L = (long)qn * n;
    uint64_t ga[RC], gb[RC];
#pragma unroll
    for (int r = 0; r < RC; ++r)
        if (r < rc) {
            ga[r] = rgsw[(long)r * 2 * L + i];
            gb[r] = rgsw[(long)r * 2 * L + L + i];
        }
    // two ciphertexts per iteration: twice the digit loads in flight (the
    // MAC waits on memory; 422 against 445-452 ms per 128 products at
    // N = 2^16, 28 primes, base 2^20, with unrolling by 4 no better)
#pragma unroll 2
    for (int bt = 0; bt < batch; ++bt) {
        kbfv_u128 a0{0, 0}, a1{0, 0};
#pragma unroll
        for (int r = 0; r < RC; ++r)
            if (r < rc) {
                const uint64_t d = dig[((long)bt * rc + r) * L + i];
                kb_mac128(a0, d, ga[r]);
                kb_mac128(a1, d, gb[r]);
            }
        uint64_t *o = out + (long)bt * 2 * L + i;
        o[0] = kb_addmod(o[0], kb_barrett128(a0, q, mlo, mhi), q);
        o[L] = kb_addmod(o[L], kb_barrett128(a1, q, mlo, mhi), q);
    }
    (void)k;
}

// rgsw [rows][2][qn][N] += gadget(m): row (h,i,j) gets m_ntt[i] * 2^(bb*j) on limb i of component h
__global__ void k_rgsw_add_gadget(kb_dev_tables t, uint64_t *__restrict__ rgsw,
                                  const uint64_t *__restrict__ m_ntt, int qn, int oQ, int nd, int bb,
                                  int n, long total) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;  // total = rows * n
    const int k = (int)(i % n);
    const int row = (int)(i / n);
    const int h = row / (qn * nd), li = (row / nd) % qn, j = row % nd;
    const int m = oQ + li;
    const uint64_t q = t.mod[m], mlo = t.mu[2 * m], mhi = t.mu[2 * m + 1];
    // 2^(bb*j) mod q by square-and-multiply (bb*j < 64 * 64)
    uint64_t g = 1, base = 2, e = (uint64_t)bb * j;
    while (e) {
        if (e & 1) g = kb_mulmod(g, base, q, mlo, mhi);
        base = kb_mulmod(base, base, q, mlo, mhi);
        e >>= 1;
    }
    const uint64_t add = kb_mulmod(m_ntt[(long)li * n + k], g, q, mlo, mhi);
    uint64_t *dst = rgsw + ((long)row * 2 + h) * qn * (long)n + (long)li * n + k;
    *dst = kb_addmod(*dst, add, q);
}

}  // namespace

int kb_digit_decompose(const kbfv_context *c, const uint64_t *coef, uint64_t *dig, int batch, int base_bits,
                       int nd, int r0, int rc, hipStream_t st) {
    const long total = (long)batch * rc * c->n;
    hipLaunchKernelGGL(k_digit_decompose, dim3(blocks_for(total, XP_BS)), dim3(XP_BS), 0, st, c->t, coef, dig,
                       c->qn, c->oQ, nd, base_bits, r0, rc, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_gadget_mac(const kbfv_context *c, const uint64_t *dig, const uint64_t *rgsw_rows, uint64_t *out,
                  int batch, int rc, hipStream_t st) {
    KB_REQUIRE(rc >= 1 && rc <= XP_MAXRC, "row chunk out of range");
    const long total = (long)c->qn * c->n;
    hipLaunchKernelGGL(k_gadget_mac<XP_MAXRC>, dim3(blocks_for(total, XP_BS)), dim3(XP_BS), 0, st, c->t, dig,
                       rgsw_rows, out, c->qn, c->oQ, c->n, batch, rc, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}

int kb_rgsw_add_gadget(const kbfv_context *c, uint64_t *rgsw, const uint64_t *m_ntt, int base_bits, int nd,
                       hipStream_t st) {
    const int rows = 2 * c->qn * nd;
    const long total = (long)rows * c->n;
    hipLaunchKernelGGL(k_rgsw_add_gadget, dim3(blocks_for(total, XP_BS)), dim3(XP_BS), 0, st, c->t, rgsw, m_ntt,
                       c->qn, c->oQ, nd, base_bits, c->n, total);
    KB_LAUNCH_CHECK();
    return KBFV_OK;
}
