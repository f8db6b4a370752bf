// BFV context: parameter derivation and every precomputed table, uploaded once
// per context (no __constant__ globals, so contexts of different N coexist).
//
// Replaces BFVContext::BFVContext / getPrimeBFV / preComputeOnCPU /
// copyMemoryToGPU (include/BFVcontext.cuh:13-101, 130-542, 545-1395;
// include/Params.cuh:7-148) and BaseConverter::init
// (include/BasisConv_QR.cuh:53-189).  Host arithmetic uses unsigned __int128
// and a small multi-word integer where products of many primes are needed.
#include <errno.h>
#include <sys/random.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "kbfv_internal.h"
#include "kbfv_mfma.h"

// ---------------------------------------------------------------- errors
static thread_local char g_err[512] = "";
void kb_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
extern "C" const char *kbfv_last_error(void) { return g_err; }

// ---------------------------------------------------------------- number theory
namespace {

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((kb_u128_host)a * b) % q); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}
uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a % q, q - 2, q); }
uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((kb_u128_host)w << 64) / q); }
kbfv_tw tw_of(uint64_t w, uint64_t q) { return kbfv_tw{w, shoup(w, q)}; }

bool is_prime(uint64_t n) {
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : bases)
        if (n % b == 0) return n == b;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, s++;
    for (uint64_t b : bases) {
        uint64_t x = powmod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool ok = false;
        for (int r = 1; r < s && !ok; r++) {
            x = mulmod(x, x, n);
            ok = (x == n - 1);
        }
        if (!ok) return false;
    }
    return true;
}

int bits_of(uint64_t v) {
    int b = 0;
    while (v) b++, v >>= 1;
    return b;
}

// The reference's root choice: the smallest primitive root g, with the prime
// factors of q-1 found by trial division bounded by sqrt of the shrinking
// cofactor (include/Utils.cuh:55-88); psi = g^((q-1)/2N) (include/Utils.cuh:90).
uint64_t primitive_root(uint64_t q) {
    std::vector<uint64_t> fac;
    uint64_t number = q - 1;
    if (number % 2 == 0) {
        fac.push_back(2);
        while (number % 2 == 0) number /= 2;
    }
    for (uint64_t i = 3; (double)i < std::sqrt((double)number); i++) {
        if (number % i == 0) {
            fac.push_back(i);
            while (number % i == 0) number /= i;
        }
    }
    if (number > 2 && std::find(fac.begin(), fac.end(), number) == fac.end()) fac.push_back(number);
    uint64_t phi = q - 1;
    for (uint64_t r = 2; r <= phi; r++) {
        bool bad = false;
        for (uint64_t f : fac)
            if (powmod(r, phi / f, q) == 1) {
                bad = true;
                break;
            }
        if (!bad) return r;
    }
    return 0;
}

uint64_t bitrev(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) r = (r << 1) | (x & 1), x >>= 1;
    return r;
}

// multi-word unsigned integer (little endian words)
struct Big {
    std::vector<uint64_t> w{0};
    explicit Big(uint64_t v = 0) { w[0] = v; }
    void mul(uint64_t v) {
        kb_u128_host carry = 0;
        for (auto &x : w) {
            kb_u128_host p = (kb_u128_host)x * v + carry;
            x = (uint64_t)p;
            carry = p >> 64;
        }
        if (carry) w.push_back((uint64_t)carry);
    }
    uint64_t divmod(uint64_t d) {
        kb_u128_host rem = 0;
        for (size_t i = w.size(); i-- > 0;) {
            kb_u128_host cur = (rem << 64) | w[i];
            w[i] = (uint64_t)(cur / d);
            rem = cur % d;
        }
        while (w.size() > 1 && w.back() == 0) w.pop_back();
        return (uint64_t)rem;
    }
    uint64_t mod(uint64_t d) const {
        kb_u128_host rem = 0;
        for (size_t i = w.size(); i-- > 0;) rem = ((rem << 64) | w[i]) % d;
        return (uint64_t)rem;
    }
    Big mulBig(const Big &o) const {
        Big r;
        r.w.assign(w.size() + o.w.size() + 1, 0);
        for (size_t i = 0; i < w.size(); i++) {
            kb_u128_host carry = 0;
            for (size_t j = 0; j < o.w.size(); j++) {
                kb_u128_host cur = (kb_u128_host)w[i] * o.w[j] + r.w[i + j] + carry;
                r.w[i + j] = (uint64_t)cur;
                carry = cur >> 64;
            }
            r.w[i + o.w.size()] += (uint64_t)carry;
        }
        while (r.w.size() > 1 && r.w.back() == 0) r.w.pop_back();
        return r;
    }
    int cmp(const Big &o) const {
        size_t n = std::max(w.size(), o.w.size());
        for (size_t i = n; i-- > 0;) {
            uint64_t a = i < w.size() ? w[i] : 0, b = i < o.w.size() ? o.w[i] : 0;
            if (a != b) return a > b ? 1 : -1;
        }
        return 0;
    }
};

// product of m[lo..hi) except index skip, mod q
uint64_t prod_except(const std::vector<uint64_t> &m, int lo, int hi, int skip, uint64_t q) {
    uint64_t r = 1 % q;
    for (int i = lo; i < hi; i++)
        if (i != skip) r = mulmod(r, m[i] % q, q);
    return r;
}

}  // namespace

// ---------------------------------------------------------------- parameter API
// 32 bytes from the operating system's entropy source (getrandom(2), which
// blocks only until the kernel's pool is first initialised).  The reference
// draws its key and encryption randomness from std::random_device
// (include/Utils.cuh:133-134, include/RNG.cuh:27-28).
extern "C" int kbfv_random_seed(uint8_t out[32]) {
    KB_REQUIRE(out, "null argument");
    size_t got = 0;
    while (got < 32) {
        const ssize_t r = getrandom(out + got, 32 - got, 0);
        if (r < 0) {
            if (errno == EINTR) continue;
            kb_set_error("getrandom failed: %s", strerror(errno));
            return KBFV_EINVAL;
        }
        got += (size_t)r;
    }
    return KBFV_OK;
}

extern "C" int kbfv_primes_below(uint64_t ntt_size, uint64_t upper, int count, uint64_t *out) {
    // include/uintarith_bfv.h:983 get_primes_below
    uint64_t factor = 2 * ntt_size, value = upper - factor;
    uint64_t lower = 1ULL << (bits_of(upper) - 1);
    int got = 0;
    while (got < count && value > lower) {
        if (is_prime(value)) out[got++] = value;
        value -= factor;
    }
    if (got != count) {
        kb_set_error("not enough primes below %llu", (unsigned long long)upper);
        return KBFV_EINVAL;
    }
    return KBFV_OK;
}

static int primes_desc(uint64_t below, uint64_t step, int count, uint64_t *out) {
    uint64_t v = below - step + 1;
    int got = 0;
    while (got < count && v > step) {
        if (is_prime(v)) out[got++] = v;
        v -= step;
    }
    return got == count ? KBFV_OK : KBFV_EINVAL;
}

extern "C" int kbfv_chain_60bit(int qn, int pn, int tn, uint64_t *q, uint64_t *p, uint64_t *t) {
    KB_REQUIRE(qn > 0 && pn > 0 && tn > 0 && qn + pn + tn <= 256, "bad chain sizes");
    std::vector<uint64_t> all(qn + pn + tn);
    KB_REQUIRE(primes_desc(1ULL << 60, 1ULL << 17, (int)all.size(), all.data()) == KBFV_OK,
               "prime search failed");
    for (int i = 0; i < tn; i++) t[i] = all[i];
    for (int i = 0; i < pn; i++) p[i] = all[tn + i];
    for (int i = 0; i < qn; i++) q[i] = all[tn + pn + i];
    return KBFV_OK;
}

extern "C" int kbfv_default_params(int log_n, uint64_t *q, int *qn, uint64_t *p, int *pn,
                                   uint64_t *t, int *tn, int *gamma, uint64_t *plain) {
    switch (log_n) {
        case 12: {
            // SEAL-style 4096 set: two 36-bit q and one 37-bit special prime, 1 mod 8192
            KB_REQUIRE(primes_desc(1ULL << 36, 8192, 2, q) == KBFV_OK, "prime search");
            KB_REQUIRE(primes_desc(1ULL << 37, 8192, 1, p) == KBFV_OK, "prime search");
            *qn = 2, *pn = 1, *tn = 2, *gamma = 1, *plain = 65537;
            return primes_desc(1ULL << 60, 1ULL << 17, 2, t);
        }
        case 13: {
            *qn = 2, *pn = 1, *tn = 3, *gamma = 1, *plain = 65537;
            return kbfv_chain_60bit(2, 1, 3, q, p, t);
        }
        case 14: {
            // the reference's logN=14 set (include/Params.cuh:31-46)
            static const uint64_t Q[] = {0x7FFE28001, 0x7FFF18001, 0x7FFF80001, 0x7FFFB0001};
            static const uint64_t Pp[] = {0xFFFFCA8001, 0xFFFFE80001};
            static const uint64_t T[] = {0xffffffffffc0001, 0xfffffffff840001, 0xfffffffff6a0001,
                                         0xfffffffff5a0001};
            memcpy(q, Q, sizeof(Q)), memcpy(p, Pp, sizeof(Pp)), memcpy(t, T, sizeof(T));
            *qn = 4, *pn = 2, *tn = 4, *gamma = 2, *plain = 23068673;
            return KBFV_OK;
        }
        case 15: {
            // the reference's logN=15 set (include/Params.cuh:120-145) with gamma = 2:
            // its gamma = 4 violates T > 4*dnum*N*q*R_b and breaks relinearisation.
            static const uint64_t Q[] = {
                36028797001138177, 36028797003563009, 36028797003694081, 36028797005135873,
                36028797005529089, 36028797005856769, 36028797009985537, 36028797010444289,
                36028797012606977, 36028797013000193, 36028797013327873, 36028797014376449,
                36028797014573057, 36028797014704129, 36028797017456641};
            static const uint64_t Pp[] = {72057594037338113};
            static const uint64_t T[] = {0xffffffffffc0001, 0xfffffffff840001, 0xfffffffff6a0001,
                                         0xfffffffff5a0001};
            memcpy(q, Q, sizeof(Q)), memcpy(p, Pp, sizeof(Pp)), memcpy(t, T, sizeof(T));
            *qn = 15, *pn = 1, *tn = 4, *gamma = 2, *plain = 0xC0001;
            return KBFV_OK;
        }
        case 16: {
            *qn = 28, *pn = 1, *tn = 4, *gamma = 2, *plain = 65537;
            return kbfv_chain_60bit(28, 1, 4, q, p, t);
        }
        default:
            kb_set_error("no default parameter set for logN=%d", log_n);
            return KBFV_EINVAL;
    }
}

// ---------------------------------------------------------------- context creation
namespace {

struct Blob {
    std::vector<uint8_t> bytes;
    template <class T>
    size_t add(const std::vector<T> &v) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + v.size() * sizeof(T) + 8);
        if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

std::vector<kbfv_tw> tws(const std::vector<uint64_t> &v, const std::vector<uint64_t> &mods) {
    std::vector<kbfv_tw> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = tw_of(v[i], mods[i]);
    return r;
}

// int8-MFMA operand tables.  For a K x J matrix M mod p_j: the expanded matrix
// [JP][8][KP] of packed digit bytes (entry E[j][v][i] packs, in byte u, digit
// v of the balanced digits of 2^(8u) M_ij mod p_j), padded to 16 ceil(J/16)
// rows and 8 ceil(K/8) columns with zeros; then per output row j KB_MX_ROW
// words: the offset (-KB_MX_ACC_BIAS * sum_v 2^(8v)) mod p_j that cancels the
// accumulator bias, 2^64 - p_j, the offset and 1/p_j as doubles, and the
// diagonal coefficient of scale-and-round with its Shoup companion.
template <class Entry, class Diag>
std::vector<uint64_t> mfma_table(int K, int J, const std::vector<uint64_t> &pj, Entry entry, Diag diag) {
    const int KP = 8 * ((K + 7) / 8), JP = 16 * ((J + 15) / 16);
    std::vector<uint64_t> v((size_t)JP * 8 * KP + KB_MX_ROW * (size_t)JP, 0);
    kb_u128_host bias_sum = 0;
    for (int u = 0; u < 8; u++) bias_sum += (kb_u128_host)KB_MX_ACC_BIAS << (8 * u);
    for (int j = 0; j < J; j++) {
        for (int i = 0; i < K; i++) {
            const uint64_t m = entry(i, j) % pj[j];
            for (int u = 0; u < 8; u++) {
                const uint64_t mu = (uint64_t)(((kb_u128_host)m << (8 * u)) % pj[j]);
                const uint64_t d = (mu + KB_DIGIT_BIAS) ^ KB_DIGIT_BIAS;
                for (int w = 0; w < 8; w++) v[((size_t)j * 8 + w) * KP + i] |= ((d >> (8 * w)) & 0xff) << (8 * u);
            }
        }
        uint64_t *row = v.data() + (size_t)JP * 8 * KP + KB_MX_ROW * (size_t)j;
        const uint64_t off = (pj[j] - (uint64_t)(bias_sum % pj[j])) % pj[j];
        const double offd = (double)off, invp = 1.0 / (double)pj[j];
        const uint64_t dg = diag(j) % pj[j];
        row[0] = off;
        row[1] = 0 - pj[j];
        std::memcpy(&row[2], &offd, 8);
        std::memcpy(&row[3], &invp, 8);
        row[4] = dg;
        row[5] = shoup(dg, pj[j]);
    }
    return v;
}

// The tables that depend on the ciphertext modulus Q (and the HPS basis R
// that goes with it): HPS base conversion Q <-> R and scale-and-round
// (include/BasisConv_QR.cuh:53-189, include/BFVcontext.cuh:1268-1388), HPS
// decryption (:1024-1137), the encryption constant -Q mod t (:881-924) and the
// int8-MFMA operands of the first two.  Built once for the full Q and once
// per level view (Q_l, R_l).
struct LevelTables {
    std::vector<uint64_t> q_hatinv, q_hat_mod_r, alphaQ, r_hatinv, r_hat_mod_q, alphaR, sr_modr;
    std::vector<double> q_inv, r_inv, sr_frac, dec_frac, dec_B_frac;
    std::vector<uint64_t> dec_modt, dec_B_modt, q_hatinv_n, mx_qr, mx_rq, mx_sr;
    int qMSB = 0, sizeQMSB = 0, tMSB = 0;
    uint64_t negQ_mod_t = 0;
    bool mx_ok = false;
};

LevelTables level_tables(const std::vector<uint64_t> &Q, const std::vector<uint64_t> &R, uint64_t plain, int N) {
    const int qn = (int)Q.size(), rn = (int)R.size();
    LevelTables L;
    // HPS base converters (include/BasisConv_QR.cuh:53-189)
    L.q_hatinv.resize(qn), L.q_hat_mod_r.resize((size_t)rn * qn), L.alphaQ.resize((size_t)(qn + 1) * rn);
    L.q_inv.resize(qn), L.r_inv.resize(rn);
    L.r_hatinv.resize(rn), L.r_hat_mod_q.resize((size_t)qn * rn), L.alphaR.resize((size_t)(rn + 1) * qn);
    for (int i = 0; i < qn; i++) {
        L.q_hatinv[i] = invmod(prod_except(Q, 0, qn, i, Q[i]), Q[i]);
        L.q_inv[i] = 1.0 / (double)Q[i];
    }
    for (int j = 0; j < rn; j++) {
        for (int i = 0; i < qn; i++) L.q_hat_mod_r[j * qn + i] = prod_except(Q, 0, qn, i, R[j]);
        uint64_t Qm = prod_except(Q, 0, qn, -1, R[j]);
        for (int a = 0; a <= qn; a++) L.alphaQ[a * rn + j] = mulmod(a, Qm, R[j]);
    }
    for (int i = 0; i < rn; i++) {
        L.r_hatinv[i] = invmod(prod_except(R, 0, rn, i, R[i]), R[i]);
        L.r_inv[i] = 1.0 / (double)R[i];
    }
    for (int j = 0; j < qn; j++) {
        for (int i = 0; i < rn; i++) L.r_hat_mod_q[j * rn + i] = prod_except(R, 0, rn, i, Q[j]);
        uint64_t Rm = prod_except(R, 0, rn, -1, Q[j]);
        for (int a = 0; a <= rn; a++) L.alphaR[a * qn + j] = mulmod(a, Rm, Q[j]);
    }
    // HPS scale-and-round tables (include/BFVcontext.cuh:1268-1388)
    L.sr_modr.assign((size_t)rn * (qn + 1), 0);
    L.sr_frac.assign(qn, 0);
    {
        std::vector<uint64_t> S(Q);
        S.insert(S.end(), R.begin(), R.end());
        int sn = qn + rn;
        Big tR(plain);
        for (uint64_t r : R) tR.mul(r);
        for (int i = 0; i < sn; i++) {
            uint64_t shatinv = invmod(prod_except(S, 0, sn, i, S[i]), S[i]);
            Big x = tR;
            x.mul(shatinv);
            if (i < qn) L.sr_frac[i] = (double)x.mod(S[i]) / (double)S[i];
            x.divmod(S[i]);
            if (i < qn)
                for (int j = 0; j < rn; j++) L.sr_modr[j * (qn + 1) + i] = x.mod(R[j]);
            else
                L.sr_modr[(i - qn) * (qn + 1) + qn] = x.mod(R[i - qn]);
        }
    }
    // HPS decrypt (include/BFVcontext.cuh:1024-1137)
    L.dec_modt.resize(qn), L.dec_B_modt.resize(qn), L.dec_frac.resize(qn), L.dec_B_frac.resize(qn);
    {
        uint64_t qmax = *std::max_element(Q.begin(), Q.end());
        L.qMSB = bits_of(qmax), L.sizeQMSB = bits_of((uint64_t)qn), L.tMSB = bits_of(plain);
        int hf = L.qMSB >> 1;
        for (int i = 0; i < qn; i++) {
            kb_u128_host x = (kb_u128_host)plain * L.q_hatinv[i];
            L.dec_modt[i] = (uint64_t)((x / Q[i]) % plain);
            L.dec_frac[i] = (double)(uint64_t)(x % Q[i]) / (double)Q[i];
            uint64_t hb = (uint64_t)(((kb_u128_host)L.q_hatinv[i] << hf) % Q[i]);
            kb_u128_host y = (kb_u128_host)plain * hb;
            L.dec_B_modt[i] = (uint64_t)((y / Q[i]) % plain);
            L.dec_B_frac[i] = (double)(uint64_t)(y % Q[i]) / (double)Q[i];
        }
    }
    // encrypt (include/BFVcontext.cuh:881-924)
    L.negQ_mod_t = plain - prod_except(Q, 0, qn, -1, plain);
    // [q_hatinv n^-1]_q for inputs that are N times the INTT
    L.q_hatinv_n.resize(qn);
    for (int i = 0; i < qn; i++) L.q_hatinv_n[i] = mulmod(L.q_hatinv[i], invmod((uint64_t)N % Q[i], Q[i]), Q[i]);
    // every matrix of the MFMA path has K, J <= 32, and its epilogue's quotient
    // estimate needs output primes of at least 2^58
    L.mx_ok = rn + 1 <= 32;
    for (uint64_t m : Q) L.mx_ok = L.mx_ok && m >= (1ULL << 58);
    for (uint64_t m : R) L.mx_ok = L.mx_ok && m >= (1ULL << 58);
    if (L.mx_ok) {
        auto none = [](int) { return (uint64_t)0; };
        L.mx_qr = mfma_table(
            qn + 1, rn, R,
            [&](int i, int j) { return i < qn ? L.q_hat_mod_r[j * qn + i] : (R[j] - L.alphaQ[rn + j]) % R[j]; },
            none);
        L.mx_rq = mfma_table(
            rn + 1, qn, Q,
            [&](int i, int j) { return i < rn ? L.r_hat_mod_q[j * rn + i] : (Q[j] - L.alphaR[qn + j]) % Q[j]; },
            none);
        L.mx_sr = mfma_table(
            qn, rn, R, [&](int i, int j) { return L.sr_modr[j * (qn + 1) + i]; },
            [&](int j) { return L.sr_modr[j * (qn + 1) + qn]; });
    }
    return L;
}

struct LevelOffsets {
    size_t qhi, qhr, aq, qinv, rhi, rhq, ar, rinv, srm, srf, dm, df, dbm, dbf, qhin, mxqr, mxrq, mxsr;
};

LevelOffsets add_level(Blob &b, const LevelTables &L, const std::vector<uint64_t> &Q,
                       const std::vector<uint64_t> &R, uint64_t plain) {
    const std::vector<uint64_t> tvec_plain(Q.size(), plain);
    LevelOffsets o;
    o.qhi = b.add(tws(L.q_hatinv, Q)), o.qhr = b.add(L.q_hat_mod_r), o.aq = b.add(L.alphaQ);
    o.qinv = b.add(L.q_inv);
    o.rhi = b.add(tws(L.r_hatinv, R)), o.rhq = b.add(L.r_hat_mod_q), o.ar = b.add(L.alphaR);
    o.rinv = b.add(L.r_inv);
    o.srm = b.add(L.sr_modr), o.srf = b.add(L.sr_frac);
    o.dm = b.add(tws(L.dec_modt, tvec_plain)), o.df = b.add(L.dec_frac);
    o.dbm = b.add(tws(L.dec_B_modt, tvec_plain)), o.dbf = b.add(L.dec_B_frac);
    o.qhin = b.add(tws(L.q_hatinv_n, Q));
    o.mxqr = b.add(L.mx_qr), o.mxrq = b.add(L.mx_rq), o.mxsr = b.add(L.mx_sr);
    return o;
}

void bind_level(kb_dev_tables &d, const uint8_t *base, const LevelOffsets &o, bool mx_ok) {
    auto at = [&](size_t off) { return (const void *)(base + off); };
    d.q_hatinv = (const kbfv_tw *)at(o.qhi), d.q_hat_mod_r = (const uint64_t *)at(o.qhr);
    d.alphaQ_mod_r = (const uint64_t *)at(o.aq), d.q_inv = (const double *)at(o.qinv);
    d.r_hatinv = (const kbfv_tw *)at(o.rhi), d.r_hat_mod_q = (const uint64_t *)at(o.rhq);
    d.alphaR_mod_q = (const uint64_t *)at(o.ar), d.r_inv = (const double *)at(o.rinv);
    d.sr_modr = (const uint64_t *)at(o.srm), d.sr_frac = (const double *)at(o.srf);
    d.dec_modt = (const kbfv_tw *)at(o.dm), d.dec_frac = (const double *)at(o.df);
    d.dec_B_modt = (const kbfv_tw *)at(o.dbm), d.dec_B_frac = (const double *)at(o.dbf);
    d.q_hatinv_n = (const kbfv_tw *)at(o.qhin);
    d.mx_qr = mx_ok ? (const uint64_t *)at(o.mxqr) : nullptr;
    d.mx_rq = mx_ok ? (const uint64_t *)at(o.mxrq) : nullptr;
    d.mx_sr = mx_ok ? (const uint64_t *)at(o.mxsr) : nullptr;
}

void set_level_scalars(kbfv_context *c, const LevelTables &L, uint64_t plain) {
    c->qMSB = L.qMSB, c->sizeQMSB = L.sizeQMSB, c->tMSB = L.tMSB;
    c->negQ_mod_t = L.negQ_mod_t;
    c->negQ_mod_t_shoup = shoup(L.negQ_mod_t, plain);
}

void free_context(kbfv_context *c) {
    for (kbfv_context *v : c->views) {
        if (v->d_level_blob) (void)hipFree(v->d_level_blob);
        delete v;
    }
    c->views.clear();
    if (c->d_blob) (void)hipFree(c->d_blob);
    if (c->d_tw) (void)hipFree(c->d_tw);
    delete c;
}

}  // namespace

extern "C" int kbfv_context_create(int log_n, const uint64_t *q, int qn, const uint64_t *p, int pn,
                                   const uint64_t *t, int tn, int gamma, uint64_t plain,
                                   kbfv_context **out) {
    KB_REQUIRE(out && q && p && t, "null argument");
    KB_REQUIRE(log_n >= 12 && log_n <= 16, "logN must be in [12, 16]");
    KB_REQUIRE(qn >= 1 && pn >= 1 && tn >= 1 && gamma >= 1, "empty basis");
    KB_REQUIRE(qn <= 64 && pn <= 8 && tn <= 12 && gamma <= 16, "basis too large");
    KB_REQUIRE(plain >= 2 && plain < (1ULL << 60), "bad plain modulus");
    *out = nullptr;
    auto c = new kbfv_context();
    c->log_n = log_n;
    c->n = 1 << log_n;
    c->qn = qn, c->pn = pn, c->tn = tn, c->rn = qn + 1, c->gamma = gamma;
    c->dnum = (qn + pn - 1) / pn;
    c->ri_blocks = (pn + qn + gamma - 1) / gamma;
    c->plain = plain;
    c->nmod = pn + qn + c->rn + tn + 1;
    c->oP = 0, c->oQ = pn, c->oR = pn + qn, c->oT = pn + qn + c->rn, c->oPlain = c->nmod - 1;
    const int N = c->n, rn = c->rn, pq = pn + qn, B = c->ri_blocks, D = c->dnum;
    std::vector<uint64_t> &mod = c->mod;
    mod.assign(c->nmod, 0);
    uint64_t qmin = ~0ULL;
    for (int i = 0; i < pn; i++) mod[c->oP + i] = p[i];
    for (int i = 0; i < qn; i++) mod[c->oQ + i] = q[i], qmin = std::min(qmin, q[i]);
    for (int i = 0; i < tn; i++) mod[c->oT + i] = t[i];
    mod[c->oPlain] = plain;
    auto fail = [&](const char *msg) {
        kb_set_error("%s", msg);
        delete c;
        return KBFV_EINVAL;
    };
    // include/BFVcontext.cuh:57-58: R basis below min(Q), stepping by 2N
    if (kbfv_primes_below(N, qmin, rn, mod.data() + c->oR) != KBFV_OK) return fail("no R basis");
    for (int i = 0; i < c->nmod; i++) {
        if (i == c->oPlain) continue;
        if (!is_prime(mod[i]) || mod[i] >= (1ULL << 60) || (mod[i] - 1) % (2 * (uint64_t)N) != 0)
            return fail("every RNS modulus must be a prime below 2^60 that is 1 mod 2N");
        for (int j = 0; j < i; j++)
            if (mod[j] == mod[i]) return fail("RNS moduli must be distinct");
    }
    c->batching = ((plain - 1) % (2 * (uint64_t)N)) == 0 && is_prime(plain);
    {
        uint64_t qmax = 0, tmin = ~0ULL;
        for (int i = 0; i < qn; i++) qmax = std::max(qmax, q[i]);
        for (int i = 0; i < tn; i++) tmin = std::min(tmin, t[i]);
        c->digits_direct_t = pn == 1 && qmax / 8 < tmin;
    }

    // twiddles: bit-reversed psi powers, n^-1 folded into psiinv[1]
    // (include/Utils.cuh:171-191)
    c->psi.assign(c->nmod, 0);
    std::vector<kbfv_tw> tw((size_t)c->nmod * 2 * N), ninv(c->nmod), inv_w1(c->nmod);
    for (int m = 0; m < c->nmod; m++) {
        if (m == c->oPlain && !c->batching) continue;
        uint64_t qm = mod[m];
        uint64_t g = primitive_root(qm);
        uint64_t psi = powmod(g, (qm - 1) / (2 * (uint64_t)N), qm);
        c->psi[m] = psi;
        uint64_t psiinv = invmod(psi, qm), ni = invmod((uint64_t)N % qm, qm);
        std::vector<uint64_t> f(N), iv(N);
        f[0] = iv[0] = 1;
        for (int i = 1; i < N; i++) {
            uint64_t pr = bitrev(i - 1, log_n), nx = bitrev(i, log_n);
            f[nx] = mulmod(psi, f[pr], qm);
            iv[nx] = mulmod(psiinv, iv[pr], qm);
        }
        inv_w1[m] = tw_of(iv[1], qm);  // the last inverse stage's twiddle without n^-1
        iv[1] = mulmod(iv[1], ni, qm);
        kbfv_tw *base = tw.data() + (size_t)m * 2 * N;
        for (int i = 0; i < N; i++) base[i] = tw_of(f[i], qm), base[N + i] = tw_of(iv[i], qm);
        ninv[m] = tw_of(ni, qm);
    }
    std::vector<uint64_t> mu(2 * c->nmod);
    for (int m = 0; m < c->nmod; m++) {
        // floor(2^128 / q) = (2^128 - 1) / q for q not a power of two
        kb_u128_host all1 = ~(kb_u128_host)0;
        kb_u128_host v = all1 / mod[m];
        mu[2 * m] = (uint64_t)v, mu[2 * m + 1] = (uint64_t)(v >> 64);
    }

    std::vector<uint64_t> Q(mod.begin() + c->oQ, mod.begin() + c->oQ + qn);
    std::vector<uint64_t> R(mod.begin() + c->oR, mod.begin() + c->oR + rn);
    std::vector<uint64_t> T(mod.begin() + c->oT, mod.begin() + c->oT + tn);
    std::vector<uint64_t> P(mod.begin() + c->oP, mod.begin() + c->oP + pn);
    std::vector<uint64_t> PQ(mod.begin(), mod.begin() + pq);

    // HPS base conversion, scale-and-round, decryption and encryption tables
    // of the full level
    const LevelTables L0 = level_tables(Q, R, plain, N);
    const std::vector<uint64_t> &q_hatinv = L0.q_hatinv, &q_hat_mod_r = L0.q_hat_mod_r, &alphaQ = L0.alphaQ;
    const std::vector<uint64_t> &r_hatinv = L0.r_hatinv, &r_hat_mod_q = L0.r_hat_mod_q, &alphaR = L0.alphaR;
    const std::vector<uint64_t> &sr_modr = L0.sr_modr, &dec_modt = L0.dec_modt, &dec_B_modt = L0.dec_B_modt;
    const std::vector<double> &q_inv = L0.q_inv, &r_inv = L0.r_inv, &sr_frac = L0.sr_frac;
    const std::vector<double> &dec_frac = L0.dec_frac, &dec_B_frac = L0.dec_B_frac;
    set_level_scalars(c, L0, plain);
    c->level = qn - 1;
    std::vector<uint64_t> tinv(qn);
    for (int i = 0; i < qn; i++) tinv[i] = invmod(plain % Q[i], Q[i]);
    // key switching (include/BFVcontext.cuh:286-490)
    std::vector<uint64_t> ri_hatinv((size_t)B * gamma), ri_hatinv_mod((size_t)B * gamma, 1);
    std::vector<uint64_t> ri_hat_mod_t((size_t)B * tn * gamma), ri_mod_t((size_t)tn * B);
    for (int b = 0; b < B; b++) {
        int lo = b * gamma, hi = std::min(lo + gamma, pq);
        for (int j = lo; j < hi; j++) {
            ri_hatinv[b * gamma + j - lo] = invmod(prod_except(PQ, lo, hi, j, PQ[j]), PQ[j]);
            ri_hatinv_mod[b * gamma + j - lo] = PQ[j];
            for (int k = 0; k < tn; k++)
                ri_hat_mod_t[(b * tn + k) * gamma + j - lo] = prod_except(PQ, lo, hi, j, T[k]);
        }
        for (int k = 0; k < tn; k++) ri_mod_t[k * B + b] = prod_except(PQ, lo, hi, -1, T[k]);
    }
    std::vector<uint64_t> qj_hatinv((size_t)D * pn), qj_hatinv_mod((size_t)D * pn, 1);
    std::vector<uint64_t> qj_hat_mod_t((size_t)D * tn * pn), qj_mod_t((size_t)D * tn);
    for (int b = 0; b < D; b++) {
        int lo = b * pn, hi = std::min(lo + pn, qn);
        for (int j = lo; j < hi; j++) {
            qj_hatinv[b * pn + j - lo] = invmod(prod_except(Q, lo, hi, j, Q[j]), Q[j]);
            qj_hatinv_mod[b * pn + j - lo] = Q[j];
            for (int k = 0; k < tn; k++)
                qj_hat_mod_t[(b * tn + k) * pn + j - lo] = prod_except(Q, lo, hi, j, T[k]);
        }
        for (int k = 0; k < tn; k++) qj_mod_t[b * tn + k] = prod_except(Q, lo, hi, -1, T[k]);
    }
    std::vector<uint64_t> t_hatinv(tn), t_hat_mod_pq((size_t)pq * tn), t_mod_pq(pq), halfT(pq + tn);
    for (int k = 0; k < tn; k++) t_hatinv[k] = invmod(prod_except(T, 0, tn, k, T[k]), T[k]);
    for (int i = 0; i < pq; i++) {
        for (int k = 0; k < tn; k++) t_hat_mod_pq[i * tn + k] = prod_except(T, 0, tn, k, PQ[i]);
        t_mod_pq[i] = prod_except(T, 0, tn, -1, PQ[i]);
    }
    Big Tbig(1);
    for (uint64_t x : T) Tbig.mul(x);
    {
        Big half = Tbig;
        half.divmod(2);
        for (int i = 0; i < pq; i++) halfT[i] = half.mod(PQ[i]);
        for (int k = 0; k < tn; k++) halfT[pq + k] = half.mod(T[k]);
    }
    // the same constants for inputs that are N times the INTT (its n^-1 left
    // out): [q_hatinv n^-1]_q, [t_hatinv n^-1]_t and [N halfT]_t
    std::vector<uint64_t> t_hatinv_n(tn), halfT_n(tn);
    for (int k = 0; k < tn; k++) {
        t_hatinv_n[k] = mulmod(t_hatinv[k], invmod((uint64_t)N % T[k], T[k]), T[k]);
        halfT_n[k] = mulmod(halfT[pq + k], (uint64_t)N % T[k], T[k]);
    }
    std::vector<uint64_t> p_hat_mod_q((size_t)qn * pn), p_hatinv(pn), P_mod_q(qn), Pinv(qn);
    for (int l = 0; l < qn; l++)
        for (int i = 0; i < pn; i++) p_hat_mod_q[l * pn + i] = prod_except(P, 0, pn, i, Q[l]);
    for (int i = 0; i < pn; i++) p_hatinv[i] = invmod(prod_except(P, 0, pn, i, P[i]), P[i]);
    for (int i = 0; i < qn; i++) {
        P_mod_q[i] = prod_except(P, 0, pn, -1, Q[i]);
        Pinv[i] = invmod(P_mod_q[i], Q[i]);
    }
    // T must hold the inner product: T > 4 * dnum * N * max(Q_j) * max(R_b)
    {
        Big maxq(0), maxr(0);
        for (int b = 0; b < D; b++) {
            Big x(1);
            for (int j = b * pn; j < std::min(b * pn + pn, qn); j++) x.mul(Q[j]);
            if (x.cmp(maxq) > 0) maxq = x;
        }
        for (int b = 0; b < B; b++) {
            Big x(1);
            for (int j = b * gamma; j < std::min(b * gamma + gamma, pq); j++) x.mul(PQ[j]);
            if (x.cmp(maxr) > 0) maxr = x;
        }
        Big lhs = maxq.mulBig(maxr);
        lhs.mul(4ULL * (uint64_t)D * (uint64_t)N);
        if (Tbig.cmp(lhs) <= 0)
            return fail("T basis too small for the key-switch inner product (raise tn or lower gamma)");
    }

    // host copies for parity tests
    auto keep = [&](const char *name, const std::vector<uint64_t> &v) { c->host_u64.push_back({name, v}); };
    auto keepf = [&](const char *name, const std::vector<double> &v) { c->host_f64.push_back({name, v}); };
    keep("q_hatinv", q_hatinv), keep("q_hat_mod_r", q_hat_mod_r), keep("alphaQ_mod_r", alphaQ);
    keep("r_hatinv", r_hatinv), keep("r_hat_mod_q", r_hat_mod_q), keep("alphaR_mod_q", alphaR);
    keep("sr_modr", sr_modr), keep("dec_modt", dec_modt), keep("dec_B_modt", dec_B_modt);
    keep("tinv_mod_q", tinv), keep("negQ_mod_t", {c->negQ_mod_t});
    keep("ri_hatinv", ri_hatinv), keep("ri_hat_mod_t", ri_hat_mod_t), keep("ri_mod_t", ri_mod_t);
    keep("qj_hatinv", qj_hatinv), keep("qj_hat_mod_t", qj_hat_mod_t), keep("qj_mod_t", qj_mod_t);
    keep("t_hatinv", t_hatinv), keep("t_hat_mod_pq", t_hat_mod_pq), keep("t_mod_pq", t_mod_pq);
    keep("halfT_mod", halfT), keep("p_hat_mod_q", p_hat_mod_q), keep("p_hatinv", p_hatinv);
    keep("P_mod_q", P_mod_q), keep("Pinv_mod_q", Pinv);
    keep("dec_msb", {(uint64_t)c->qMSB, (uint64_t)c->sizeQMSB, (uint64_t)c->tMSB});
    keep("psi", c->psi);
    keepf("q_inv", q_inv), keepf("r_inv", r_inv), keepf("sr_frac", sr_frac);
    keepf("dec_frac", dec_frac), keepf("dec_B_frac", dec_B_frac);

    // SEAL slot order (include/BFVcontext.cuh:102-129): row 0 slot i at the NTT
    // index of psi^(5^i), row 1 at that of psi^(-5^i)
    std::vector<uint64_t> seal_map;
    if (c->batching) {
        seal_map.resize(N);
        uint64_t pos = 1, m2 = 2 * (uint64_t)N;
        for (int i = 0; i < N / 2; i++) {
            seal_map[i] = bitrev((pos - 1) >> 1, log_n);
            seal_map[N / 2 + i] = bitrev((m2 - pos - 1) >> 1, log_n);
            pos = pos * 5 % m2;
        }
    }

    // device upload: one blob for the small tables
    Blob blob;
    size_t o_mod = blob.add(mod), o_mu = blob.add(mu), o_ninv = blob.add(ninv);
    const LevelOffsets o_lvl = add_level(blob, L0, Q, R, plain);
    size_t o_ti = blob.add(tws(tinv, Q));
    size_t o_rih = blob.add(tws(ri_hatinv, ri_hatinv_mod)), o_rht = blob.add(ri_hat_mod_t);
    size_t o_rmt = blob.add(ri_mod_t);
    size_t o_qjh = blob.add(tws(qj_hatinv, qj_hatinv_mod)), o_qjt = blob.add(qj_hat_mod_t);
    size_t o_qjm = blob.add(qj_mod_t);
    size_t o_th = blob.add(tws(t_hatinv, T)), o_thp = blob.add(t_hat_mod_pq), o_tmp = blob.add(t_mod_pq);
    size_t o_half = blob.add(halfT), o_phq = blob.add(p_hat_mod_q), o_phi = blob.add(tws(p_hatinv, P));
    size_t o_pinv = blob.add(tws(Pinv, Q)), o_pmq = blob.add(P_mod_q);
    // modulus switching (the reference's qiInvVecModql, include/BFVScheme.cuh:403-418)
    std::vector<kbfv_tw> ql_inv((size_t)qn * qn, kbfv_tw{0, 0});
    for (int l = 0; l < qn; l++)
        for (int i = 0; i < l; i++) ql_inv[(size_t)l * qn + i] = tw_of(invmod(Q[l] % Q[i], Q[i]), Q[i]);
    size_t o_qli = blob.add(ql_inv);
    size_t o_seal = blob.add(seal_map);
    size_t o_thn = blob.add(tws(t_hatinv_n, T)), o_hfn = blob.add(halfT_n);
    size_t o_iw1 = blob.add(inv_w1);

    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) {
        const int rc = kb_noalloc_setup(c->device);
        if (rc != KBFV_OK) {
            delete c;
            return rc;
        }
    }
    if (e == hipSuccess) e = hipMalloc(&c->d_blob, blob.bytes.size());
    if (e == hipSuccess) e = hipMemcpy(c->d_blob, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_tw, tw.size() * sizeof(kbfv_tw));
    if (e == hipSuccess) e = hipMemcpy(c->d_tw, tw.data(), tw.size() * sizeof(kbfv_tw), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        kb_set_error("context upload: %s", hipGetErrorString(e));
        if (c->d_blob) (void)hipFree(c->d_blob);
        if (c->d_tw) (void)hipFree(c->d_tw);
        delete c;
        return KBFV_EHIP;
    }
    auto at = [&](size_t off) { return (const void *)((const uint8_t *)c->d_blob + off); };
    kb_dev_tables &d = c->t;
    d.mod = (const uint64_t *)at(o_mod), d.mu = (const uint64_t *)at(o_mu);
    d.tw = c->d_tw, d.ninv = (const kbfv_tw *)at(o_ninv);
    bind_level(d, (const uint8_t *)c->d_blob, o_lvl, L0.mx_ok);
    d.tinv_mod_q = (const kbfv_tw *)at(o_ti);
    d.ri_hatinv = (const kbfv_tw *)at(o_rih), d.ri_hat_mod_t = (const uint64_t *)at(o_rht);
    d.ri_mod_t = (const uint64_t *)at(o_rmt);
    d.qj_hatinv = (const kbfv_tw *)at(o_qjh), d.qj_hat_mod_t = (const uint64_t *)at(o_qjt);
    d.qj_mod_t = (const uint64_t *)at(o_qjm);
    d.t_hatinv = (const kbfv_tw *)at(o_th), d.t_hat_mod_pq = (const uint64_t *)at(o_thp);
    d.t_mod_pq = (const uint64_t *)at(o_tmp), d.halfT_mod = (const uint64_t *)at(o_half);
    d.halfT_t = d.halfT_mod + pq;
    d.p_hat_mod_q = (const uint64_t *)at(o_phq), d.p_hatinv = (const kbfv_tw *)at(o_phi);
    d.Pinv_mod_q = (const kbfv_tw *)at(o_pinv), d.P_mod_q = (const uint64_t *)at(o_pmq);
    d.ql_inv_mod_qi = (const kbfv_tw *)at(o_qli);
    if (c->batching) d.seal_map = (const uint64_t *)at(o_seal);
    d.t_hatinv_n = (const kbfv_tw *)at(o_thn);
    d.halfT_n = (const uint64_t *)at(o_hfn), d.inv_w1 = (const kbfv_tw *)at(o_iw1);
    c->key_dnum = c->dnum, c->key_blocks = c->ri_blocks;

    // level views l = 0 .. qn - 2: Q_l = q_0..q_l and R_l = the l + 2 moduli
    // that follow q_l in the [Q | R] table, so that Q_l R_l is again one
    // contiguous range of the table (the HomMult pipeline transforms and
    // tensors over it as one range); keys of the full level are read on their
    // first dnum_l digits and ri_blocks_l blocks (the last block's limbs past
    // Q_l are produced and dropped)
    for (int l = 0; l + 1 < qn; ++l) {
        const std::vector<uint64_t> Ql(mod.begin() + c->oQ, mod.begin() + c->oQ + l + 1),
            Rl(mod.begin() + c->oQ + l + 1, mod.begin() + c->oQ + 2 * l + 3);
        const LevelTables Ll = level_tables(Ql, Rl, plain, N);
        Blob lb;
        const LevelOffsets ol = add_level(lb, Ll, Ql, Rl, plain);
        // digit groups of pn primes over Q_l (with pn > 1 the last one may be
        // shorter than at the full level)
        const int Dl = (l + 1 + pn - 1) / pn;
        std::vector<uint64_t> lqh((size_t)Dl * pn), lqh_mod((size_t)Dl * pn, 1), lqht((size_t)Dl * tn * pn),
            lqm((size_t)Dl * tn);
        for (int b = 0; b < Dl; b++) {
            int lo = b * pn, hi = std::min(lo + pn, l + 1);
            for (int j = lo; j < hi; j++) {
                lqh[b * pn + j - lo] = invmod(prod_except(Ql, lo, hi, j, Ql[j]), Ql[j]);
                lqh_mod[b * pn + j - lo] = Ql[j];
                for (int k = 0; k < tn; k++) lqht[(b * tn + k) * pn + j - lo] = prod_except(Ql, lo, hi, j, T[k]);
            }
            for (int k = 0; k < tn; k++) lqm[b * tn + k] = prod_except(Ql, lo, hi, -1, T[k]);
        }
        const size_t o_lqh = lb.add(tws(lqh, lqh_mod)), o_lqht = lb.add(lqht), o_lqm = lb.add(lqm);
        auto v = new kbfv_context();
        c->views.push_back(v);
        v->parent = c, v->level = l;
        v->log_n = c->log_n, v->n = c->n;
        v->qn = l + 1, v->pn = pn, v->tn = tn, v->rn = l + 2, v->gamma = gamma;
        v->dnum = (l + 1 + pn - 1) / pn;
        v->ri_blocks = (pn + l + 1 + gamma - 1) / gamma;
        v->key_dnum = c->dnum, v->key_blocks = c->ri_blocks;
        v->batching = c->batching, v->nmod = c->nmod;
        v->oP = c->oP, v->oQ = c->oQ, v->oR = c->oQ + l + 1, v->oT = c->oT, v->oPlain = c->oPlain;
        v->plain = plain, v->device = c->device, v->digits_direct_t = c->digits_direct_t;
        v->mod = c->mod;
        set_level_scalars(v, Ll, plain);
        v->t = c->t;
        hipError_t ev = hipMalloc(&v->d_level_blob, lb.bytes.size());
        if (ev == hipSuccess)
            ev = hipMemcpy(v->d_level_blob, lb.bytes.data(), lb.bytes.size(), hipMemcpyHostToDevice);
        if (ev != hipSuccess) {
            kb_set_error("level tables upload: %s", hipGetErrorString(ev));
            free_context(c);
            return KBFV_EHIP;
        }
        bind_level(v->t, (const uint8_t *)v->d_level_blob, ol, Ll.mx_ok);
        const uint8_t *vb = (const uint8_t *)v->d_level_blob;
        v->t.qj_hatinv = (const kbfv_tw *)(vb + o_lqh);
        v->t.qj_hat_mod_t = (const uint64_t *)(vb + o_lqht);
        v->t.qj_mod_t = (const uint64_t *)(vb + o_lqm);
    }
    *out = c;
    return KBFV_OK;
}

extern "C" int kbfv_context_level(const kbfv_context *c, int level, const kbfv_context **view) {
    KB_REQUIRE(c && view, "null argument");
    const kbfv_context *root = kb_root(c);
    KB_REQUIRE(level >= 0 && level < root->qn, "level %d out of range [0, %d]", level, root->qn - 1);
    *view = level == root->qn - 1 ? root : root->views[level];
    return KBFV_OK;
}

extern "C" int kbfv_context_destroy(kbfv_context *c) {
    if (!c) return KBFV_OK;
    KB_REQUIRE(!c->parent, "a level view belongs to its context and is destroyed with it");
    free_context(c);
    return KBFV_OK;
}

extern "C" int kbfv_context_info(const kbfv_context *c, int *out) {
    KB_REQUIRE(c && out, "null argument");
    int v[16] = {c->log_n, c->n, c->qn, c->pn, c->tn, c->rn, c->gamma, c->dnum,
                 c->ri_blocks, c->batching, c->nmod, c->oP, c->oQ, c->oR, c->oT, c->oPlain};
    memcpy(out, v, sizeof(v));
    return KBFV_OK;
}

extern "C" int kbfv_context_moduli(const kbfv_context *c, uint64_t *out) {
    KB_REQUIRE(c && out, "null argument");
    memcpy(out, c->mod.data(), c->mod.size() * sizeof(uint64_t));
    return c->nmod;
}

extern "C" int kbfv_context_export(const kbfv_context *c, const char *name, uint64_t *out, int cap) {
    KB_REQUIRE(c && name && out, "null argument");
    for (auto &kv : c->host_u64)
        if (kv.first == name) {
            KB_REQUIRE((int)kv.second.size() <= cap, "export buffer too small");
            memcpy(out, kv.second.data(), kv.second.size() * sizeof(uint64_t));
            return (int)kv.second.size();
        }
    kb_set_error("unknown table %s", name);
    return KBFV_EINVAL;
}

extern "C" int kbfv_context_export_f64(const kbfv_context *c, const char *name, double *out, int cap) {
    KB_REQUIRE(c && name && out, "null argument");
    for (auto &kv : c->host_f64)
        if (kv.first == name) {
            KB_REQUIRE((int)kv.second.size() <= cap, "export buffer too small");
            memcpy(out, kv.second.data(), kv.second.size() * sizeof(double));
            return (int)kv.second.size();
        }
    kb_set_error("unknown table %s", name);
    return KBFV_EINVAL;
}
