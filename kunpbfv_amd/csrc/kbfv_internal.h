// Internal host-side definitions shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "kbfv_device.h"
#include "kunpbfv.h"

typedef unsigned __int128 kb_u128_host;

// ---------------------------------------------------------------- errors
void kb_set_error(const char *fmt, ...);
#define KB_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) {                                                            \
            kb_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
            return KBFV_EHIP;                                                              \
        }                                                                                  \
    } while (0)
#define KB_REQUIRE(cond, ...)           \
    do {                                \
        if (!(cond)) {                  \
            kb_set_error(__VA_ARGS__);  \
            return KBFV_EINVAL;         \
        }                               \
    } while (0)
#define KB_TRY(x)            \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)
#define KB_LAUNCH_CHECK()                                                                    \
    do {                                                                                     \
        hipError_t _e = hipGetLastError();                                                   \
        if (_e != hipSuccess) {                                                              \
            kb_set_error("%s:%d kernel launch -> %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
            return KBFV_EHIP;                                                                \
        }                                                                                    \
    } while (0)

// ---------------------------------------------------------------- context
// Device pointers of every precomputed table.  Small tables live in one
// device allocation (blob); twiddles in their own.
struct kb_dev_tables {
    const uint64_t *mod;   // [nmod]
    const uint64_t *mu;    // [nmod][2] Barrett floor(2^128/q): lo, hi
    const kbfv_tw *tw;     // [nmod][2][N]: forward psi table, inverse psi table (bit-reversed)
    const kbfv_tw *ninv;   // [nmod] n^-1 with Shoup companion

    // HPS base conversion Q -> R and R -> Q
    const kbfv_tw *q_hatinv;       // [qn]
    const uint64_t *q_hat_mod_r;   // [rn][qn]
    const uint64_t *alphaQ_mod_r;  // [qn+1][rn]
    const double *q_inv;           // [qn]
    const kbfv_tw *r_hatinv;       // [rn]
    const uint64_t *r_hat_mod_q;   // [qn][rn]
    const uint64_t *alphaR_mod_q;  // [rn+1][qn]
    const double *r_inv;           // [rn]
    // HPS scale-and-round QR -> R
    const uint64_t *sr_modr;  // [rn][qn+1]
    const double *sr_frac;    // [qn]
    // HPS decrypt
    const kbfv_tw *dec_modt;    // [qn] (value, Shoup mod t)
    const double *dec_frac;     // [qn]
    const kbfv_tw *dec_B_modt;  // [qn]
    const double *dec_B_frac;   // [qn]
    // encrypt
    const kbfv_tw *tinv_mod_q;  // [qn]
    // key switching (full level)
    const kbfv_tw *ri_hatinv;        // [ri_blocks][gamma]
    const uint64_t *ri_hat_mod_t;    // [ri_blocks][tn][gamma]
    const uint64_t *ri_mod_t;        // [tn][ri_blocks]
    const kbfv_tw *qj_hatinv;        // [dnum][pn]
    const uint64_t *qj_hat_mod_t;    // [dnum][tn][pn]
    const uint64_t *qj_mod_t;        // [dnum][tn]
    const kbfv_tw *t_hatinv;         // [tn]
    const uint64_t *t_hat_mod_pq;    // [pn+qn][tn]
    const uint64_t *t_mod_pq;        // [pn+qn]
    const uint64_t *halfT_mod;       // [pn+qn+tn] in the order P, Q, T
    const uint64_t *halfT_t;         // [tn] = halfT_mod + pn + qn (the T part, at any level)
    const uint64_t *p_hat_mod_q;     // [qn][pn]
    const kbfv_tw *p_hatinv;         // [pn]
    const kbfv_tw *Pinv_mod_q;       // [qn]
    const uint64_t *P_mod_q;         // [qn]
    // modulus switching: [l][i] = q_l^-1 mod q_i (with Shoup companion), i < l
    const kbfv_tw *ql_inv_mod_qi;    // [qn][qn]
    // int8-MFMA operand tables: [16*ceil(J/16)][8][8*ceil(K/8)] packed digits of
    // 2^(8u) M_ij mod p_j, then [16*ceil(J/16)][2] offsets (null if too large)
    const uint64_t *mx_qr;  // base conversion Q -> R, K = qn + 1 (alpha row [-Q]_r)
    const uint64_t *mx_rq;  // base conversion R -> Q, K = rn + 1 (alpha row [-R]_q)
    const uint64_t *mx_sr;  // scale-and-round, K = qn
    // SEAL slot order: NTT index of slot i (batching only, else null)
    const uint64_t *seal_map;  // [N]
    // for inputs that are N times the INTT (kb_intt_noscale): n^-1 folded in
    const kbfv_tw *q_hatinv_n;  // [qn] [q_hatinv n^-1]_q
    const kbfv_tw *t_hatinv_n;  // [tn] [t_hatinv n^-1]_t
    const uint64_t *halfT_n;    // [tn] [N (T/2 mod t)]_t
    const kbfv_tw *inv_w1;      // [nmod] inverse table entry 1 without its n^-1
};

struct kbfv_context {
    int log_n = 0, n = 0;
    int qn = 0, pn = 0, tn = 0, rn = 0, gamma = 0, dnum = 0, ri_blocks = 0;
    int batching = 0, nmod = 0, oP = 0, oQ = 0, oR = 0, oT = 0, oPlain = 0;
    uint64_t plain = 0;
    int qMSB = 0, sizeQMSB = 0, tMSB = 0;
    uint64_t negQ_mod_t = 0, negQ_mod_t_shoup = 0;
    int device = 0;
    // K = 1 and every Q prime below 8 * min(T): a Q limb is a valid lazy input
    // of the forward NTT over every T prime, so the digit -> T conversion is
    // that NTT reading the Q limb (kb_digits_to_t_ntt)
    bool digits_direct_t = false;

    // level views: a context for ciphertexts at level `level` < qn_full - 1
    // (Q_l = the first level + 1 Q primes, R_l = the level + 2 moduli after
    // them in the [Q | R] table, oR = oQ + level + 1)
    // shares the parent's moduli, twiddles, key-switching tables and NTT plans
    // and has its own HPS, decryption and encryption tables.  The full-level
    // context has parent == nullptr and level == qn - 1.
    const kbfv_context *parent = nullptr;
    int level = 0;
    // shape of keys made on the full-level context ([2][key_dnum][key_blocks][tn][N]),
    // the strides the key-switching kernels read them with at every level
    int key_dnum = 0, key_blocks = 0;
    std::vector<kbfv_context *> views;  // full-level context only: views[l] for l < qn - 1
    void *d_level_blob = nullptr;       // a view's own tables

    // host copies (exported to parity tests)
    std::vector<uint64_t> mod, psi;
    std::vector<std::pair<std::string, std::vector<uint64_t>>> host_u64;
    std::vector<std::pair<std::string, std::vector<double>>> host_f64;

    void *d_blob = nullptr;
    kbfv_tw *d_tw = nullptr;
    kb_dev_tables t{};

    // NTT plan per direction (forward, inverse) and transform-count bucket;
    // 0 = the built-in plan.  Set by kbfv_ntt_plan_set / kbfv_ntt_autotune.
    std::atomic<uint32_t> ntt_plan[2][KBFV_NTT_BUCKETS] = {};
};

// resolves the allocation-free HomMult's scratch address on `device` (which
// must be current); once per device, at context creation (scheme.hip)
int kb_noalloc_setup(int device);

// the full-level context a view belongs to (the context itself otherwise)
static inline const kbfv_context *kb_root(const kbfv_context *c) { return c->parent ? c->parent : c; }

// per-limb modulus helpers
static inline uint64_t kb_host_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((kb_u128_host)a * b) % q);
}

// ---------------------------------------------------------------- launchers (ntt.hip)
// dir: +1 forward, -1 inverse.  Layout addressed as
//   src + bi*sB + y*sY + l*sL   (elements),   dst + bi*dB + y*dY + l*dL
// for batch bi < batch, poly y < polys, limb l < limbs; limb l uses modulus mod0 + l.
int kb_ntt(const kbfv_context *c, int dir, const uint64_t *src, uint64_t *dst, int mod0, int limbs,
           int polys, int batch, long sB, long sY, long sL, long dB, long dY, long dL,
           hipStream_t st);
// inverse NTT without the final n^-1 and reduction: dst = N * INTT(src) as lazy
// values in [0, 4q), for consumers that fold n^-1 into their first constant
int kb_intt_noscale(const kbfv_context *c, const uint64_t *src, uint64_t *dst, int mod0, int limbs, int polys,
                    int batch, long sB, long sY, long sL, long dB, long dY, long dL, hipStream_t st);
// contiguous [batch*polys][limbs][N] in place
int kb_ntt_inplace(const kbfv_context *c, int dir, uint64_t *data, int mod0, int limbs, int polys,
                   hipStream_t st);
