// C++ drop-in classes (include/kunpbfv/bfv.hpp) on the C ABI.
#include "kunpbfv/bfv.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

void ok(int rc, const char *what) {
    if (rc != KBFV_OK) throw std::runtime_error(std::string(what) + ": " + kbfv_last_error());
}
void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
uint64_tt *dalloc(size_t elems) {
    void *p = nullptr;
    hip_ok(hipMalloc(&p, elems * sizeof(uint64_tt)), "hipMalloc");
    return (uint64_tt *)p;
}
void dfree(void *p) {
    if (p) (void)hipFree(p);
}

}  // namespace

// ---------------------------------------------------------------- Plaintext / Ciphertext
Plaintext::Plaintext(int N_, int L_, int l_, int slots_) : N(N_), L(L_), l(l_), slots(slots_) {
    mx_device = dalloc((size_t)N * (L + 1));
    hip_ok(hipMemset(mx_device, 0, sizeof(uint64_tt) * N * (L + 1)), "hipMemset");
}
Plaintext::Plaintext(const Plaintext &o) : N(o.N), L(o.L), l(o.l), slots(o.slots) {
    mx_device = dalloc((size_t)N * (L + 1));
    hip_ok(hipMemcpy(mx_device, o.mx_device, sizeof(uint64_tt) * N * (L + 1), hipMemcpyDeviceToDevice), "copy");
}
Plaintext &Plaintext::operator=(const Plaintext &o) {
    if (this == &o) return *this;
    if (N != o.N || L != o.L) {
        dfree(mx_device);
        N = o.N, L = o.L;
        mx_device = dalloc((size_t)N * (L + 1));
    }
    l = o.l, slots = o.slots;
    hip_ok(hipMemcpy(mx_device, o.mx_device, sizeof(uint64_tt) * N * (L + 1), hipMemcpyDeviceToDevice), "copy");
    return *this;
}
Plaintext::~Plaintext() { dfree(mx_device); }

Ciphertext::Ciphertext(int N_, int L_, int l_, int slots_) : N(N_), l(l_), L(L_), slots(slots_) {
    cipher_device = dalloc((size_t)2 * N * (L + 1));
    ax_device = cipher_device;
    bx_device = cipher_device + (size_t)N * (L + 1);
}
Ciphertext::Ciphertext(const Ciphertext &o) : N(o.N), l(o.l), L(o.L), slots(o.slots) {
    cipher_device = dalloc((size_t)2 * N * (L + 1));
    ax_device = cipher_device;
    bx_device = cipher_device + (size_t)N * (L + 1);
    hip_ok(hipMemcpy(cipher_device, o.cipher_device, sizeof(uint64_tt) * 2 * N * (L + 1), hipMemcpyDeviceToDevice),
           "copy");
}
Ciphertext &Ciphertext::operator=(const Ciphertext &o) {
    if (this == &o) return *this;
    if (N != o.N || L != o.L) {
        dfree(cipher_device);
        N = o.N, L = o.L;
        cipher_device = dalloc((size_t)2 * N * (L + 1));
        ax_device = cipher_device;
        bx_device = cipher_device + (size_t)N * (L + 1);
    }
    l = o.l, slots = o.slots;
    hip_ok(hipMemcpy(cipher_device, o.cipher_device, sizeof(uint64_tt) * 2 * N * (L + 1), hipMemcpyDeviceToDevice),
           "copy");
    return *this;
}
Ciphertext::~Ciphertext() { dfree(cipher_device); }

// ---------------------------------------------------------------- BFVContext
BFVContext::BFVContext(size_t poly_modulus_degree) {
    int log_n = 0;
    while ((size_t(1) << log_n) < poly_modulus_degree) log_n++;
    if ((size_t(1) << log_n) != poly_modulus_degree) throw std::invalid_argument("N must be a power of two");
    uint64_t q[64], p[64], t[64], plain;
    int qn, pn, tn, g;
    ok(kbfv_default_params(log_n, q, &qn, p, &pn, t, &tn, &g, &plain), "default parameters");
    init(log_n, std::vector<uint64_tt>(q, q + qn), std::vector<uint64_tt>(p, p + pn),
         std::vector<uint64_tt>(t, t + tn), g, plain);
}

BFVContext::BFVContext(int log_n, const std::vector<uint64_tt> &q, const std::vector<uint64_tt> &p,
                       const std::vector<uint64_tt> &t, int g, uint64_tt plain) {
    init(log_n, q, p, t, g, plain);
}

void BFVContext::init(int log_n, const std::vector<uint64_tt> &q, const std::vector<uint64_tt> &p,
                      const std::vector<uint64_tt> &t, int g, uint64_tt plain) {
    ok(kbfv_context_create(log_n, (const uint64_t *)q.data(), (int)q.size(), (const uint64_t *)p.data(),
                           (int)p.size(), (const uint64_t *)t.data(), (int)t.size(), g, plain, &handle),
       "context");
    int info[16];
    ok(kbfv_context_info(handle, info), "context info");
    logN = info[0];
    N = info[1];
    q_num = info[2], p_num = info[3], t_num = info[4], r_num = info[5], gamma = info[6], dnum = info[7];
    Ri_blockNum = info[8];
    K = p_num, alpha = p_num, L = q_num - 1;
    Qj_blockNum = (q_num + p_num - 1) / p_num;
    mod_num = p_num + q_num + t_num;
    logNh = logN - 1, logslots = logN, slots = (int)N, M = 2 * N, Nh = N / 2;
    plain_modulus = plain;
    qVec = q, pVec = p, tVec = t;
    std::vector<uint64_t> mods(info[10]);
    kbfv_context_moduli(handle, mods.data());
    rVec.assign(mods.begin() + info[13], mods.begin() + info[13] + r_num);
}

BFVContext::~BFVContext() { kbfv_context_destroy(handle); }

void BFVContext::encode(uint64_tt *vals, Plaintext &msg) {
    ok(kbfv_encode(handle, (uint64_t *)vals, 1, stream), "encode");
    for (int i = 0; i <= L; ++i)
        hip_ok(hipMemcpyAsync(msg.mx_device + (size_t)i * N, vals, sizeof(uint64_tt) * N, hipMemcpyDeviceToDevice,
                              stream),
               "encode copy");
    hip_ok(hipStreamSynchronize(stream), "encode");
}

void BFVContext::encode_ntt(uint64_tt *vals, Plaintext &msg) {
    for (int i = 0; i <= L; ++i)
        hip_ok(hipMemcpyAsync(msg.mx_device + (size_t)i * N, vals, sizeof(uint64_tt) * N, hipMemcpyDeviceToDevice,
                              stream),
               "encode_ntt copy");
    const int64_t n = N;
    // modulus index of q_0 is p_num (the table is [P | Q | R | T | t])
    ok(kbfv_ntt(handle, +1, (uint64_t *)msg.mx_device, (uint64_t *)msg.mx_device, p_num, msg.l + 1, 1, 1, 0,
                (L + 1) * n, n, 0, (L + 1) * n, n, stream),
       "encode_ntt");
    hip_ok(hipStreamSynchronize(stream), "encode_ntt");
}

void BFVContext::decode(Plaintext &msg, uint64_tt *vals) {
    hip_ok(hipMemcpyAsync(vals, msg.mx_device, sizeof(uint64_tt) * N, hipMemcpyDeviceToDevice, stream), "decode");
    ok(kbfv_decode(handle, (uint64_t *)vals, 1, stream), "decode");
    hip_ok(hipStreamSynchronize(stream), "decode");
}

// the reference indexes moduli as [P | Q | T]; this engine's table is [P | Q | R | T | t]
static int pqt_to_table(const BFVContext &c, int idx) {
    return idx >= c.p_num + c.q_num ? idx + c.r_num : idx;
}

void BFVContext::ToNTTInplace(uint64_tt *a, int start_poly_idx, int start_mod_idx, int poly_num, int mod_num_,
                              int poly_mod_len) {
    const int64_t n = N;
    start_mod_idx = pqt_to_table(*this, start_mod_idx);
    uint64_t *base = (uint64_t *)a + (int64_t)start_poly_idx * n;
    ok(kbfv_ntt(handle, +1, base, base, start_mod_idx, mod_num_, poly_num, 1, 0, poly_mod_len * n, n, 0,
                poly_mod_len * n, n, stream),
       "ToNTTInplace");
}

void BFVContext::FromNTTInplace(uint64_tt *a, int start_poly_idx, int start_mod_idx, int poly_num, int mod_num_,
                                int poly_mod_len) {
    const int64_t n = N;
    start_mod_idx = pqt_to_table(*this, start_mod_idx);
    uint64_t *base = (uint64_t *)a + (int64_t)start_poly_idx * n;
    ok(kbfv_ntt(handle, -1, base, base, start_mod_idx, mod_num_, poly_num, 1, 0, poly_mod_len * n, n, 0,
                poly_mod_len * n, n, stream),
       "FromNTTInplace");
}

void BFVContext::autotuneNTT(int batch) {
    if (batch < 1) throw std::invalid_argument("autotuneNTT: batch must be positive");
    // over the QR basis: 1 polynomial, 3 (one tensor), a mid-size batch and the
    // whole batch's 3 tensor polynomials, one size per transform-count bucket
    const int limbs = q_num + r_num;
    int polys_of[KBFV_NTT_BUCKETS] = {0};
    for (int polys : {1, 3, std::max(1, 2048 / limbs), 3 * batch}) {
        const int b = kbfv_ntt_plan_bucket((int64_t)limbs * polys);
        if (b >= 0) polys_of[b] = std::max(polys_of[b], polys);
    }
    int most = 0;
    for (int p : polys_of) most = std::max(most, p);
    uint64_tt *scratch = dalloc((size_t)most * limbs * N);
    try {
        for (int dir : {+1, -1})
            for (int b = 0; b < KBFV_NTT_BUCKETS; ++b)
                if (polys_of[b])
                    ok(kbfv_ntt_autotune(handle, dir, p_num, limbs, polys_of[b], (uint64_t *)scratch, 3, stream,
                                         nullptr, nullptr),
                       "autotuneNTT");
    } catch (...) {
        dfree(scratch);
        throw;
    }
    dfree(scratch);
}

void BFVContext::external_product_T_swk_reuse(uint64_tt *output, uint64_tt *cipher_modUp_QjtoT,
                                              uint64_tt *swk_modUp_RitoT, int l, int batch_size) {
    if (l != L) throw std::invalid_argument("external_product_T_swk_reuse: ciphertexts are at full level");
    if (batch_size <= 0) return;
    ok(kbfv_inner_product(handle, (const uint64_t *)cipher_modUp_QjtoT, (const uint64_t *)swk_modUp_RitoT,
                          (uint64_t *)output, batch_size, stream),
       "external_product_T_swk_reuse");
    hip_ok(hipStreamSynchronize(stream), "external_product_T_swk_reuse");
}

// ---------------------------------------------------------------- keys
static uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

SecretKey::SecretKey(BFVContext &context, hipStream_t stream) : N((int)context.N), L(context.L), K(context.K) {
    uint8_t seed[32];  // fresh from the OS, as the reference's std::random_device (include/Utils.cuh:133)
    ok(kbfv_random_seed(seed), "secret key seed");
    sx_device = dalloc((size_t)N * (K + L + 1));
    ok(kbfv_keygen_sk(context.handle, seed, (uint64_t *)sx_device, stream), "secret key");
    hip_ok(hipStreamSynchronize(stream), "secret key");
}

SecretKey::SecretKey(BFVContext &context, const uint8_t seed[32], hipStream_t stream)
    : N((int)context.N), L(context.L), K(context.K) {
    sx_device = dalloc((size_t)N * (K + L + 1));
    ok(kbfv_keygen_sk(context.handle, seed, (uint64_t *)sx_device, stream), "secret key");
    hip_ok(hipStreamSynchronize(stream), "secret key");
}
SecretKey::~SecretKey() { dfree(sx_device); }

Key::Key(int N_, int L_) : N(N_), L(L_) {
    cipher_device = dalloc((size_t)2 * N * (L + 1));
    ax_device = cipher_device;
    bx_device = cipher_device + (size_t)N * (L + 1);
}
Key::~Key() { dfree(cipher_device); }

Key_decomp::Key_decomp(int N_, int dnum_, int t_num_, int blocks) : N(N_), dnum(dnum_), t_num(t_num_), Ri_blockNum(blocks) {
    const size_t half = (size_t)dnum * blocks * t_num * N;
    cipher_device = dalloc(2 * half);
    ax_device = cipher_device;
    bx_device = cipher_device + half;
}
Key_decomp::~Key_decomp() { dfree(cipher_device); }

// ---------------------------------------------------------------- BFVScheme
BFVScheme::BFVScheme(BFVContext &ctx)
    : context(ctx), rotKey_vec_23((size_t)ctx.N, nullptr), autoKey_vec_23((size_t)ctx.logN, nullptr) {
    // fresh randomness per scheme (the reference re-seeds from std::random_device,
    // include/RNG.cuh:27-28); setSeed makes keys and encryptions reproducible
    uint8_t s[32];
    ok(kbfv_random_seed(s), "scheme seed");
    memcpy(&seed_state, s, sizeof(seed_state));
}
BFVScheme::~BFVScheme() {
    delete publicKey;
    delete rlk_23;
    for (auto *k : rotKey_vec_23) delete k;
    for (auto *k : autoKey_vec_23) delete k;
    delete ConjKey_23;
    dfree(ws);
    dfree(d01);
    dfree(batch_a);
    dfree(batch_b);
}

void BFVScheme::setSeed(uint64_t seed) { seed_state = seed; }

void BFVScheme::seed_next(uint8_t out[32]) {
    for (int i = 0; i < 4; ++i) {
        uint64_t v = splitmix(seed_state);
        memcpy(out + 8 * i, &v, 8);
    }
}

void *BFVScheme::ensure_ws(size_t bytes) {
    if (bytes > ws_bytes) {
        dfree(ws);
        ws = nullptr;
        hip_ok(hipMalloc(&ws, bytes), "workspace");
        ws_bytes = bytes;
    }
    return ws;
}

void BFVScheme::mallocMemory() {
    kbfv_context *h = context.handle;
    size_t need = kbfv_mult_workspace_bytes(h, 1);
    need = std::max(need, kbfv_encrypt_workspace_bytes(h, 1));
    need = std::max(need, kbfv_decrypt_workspace_bytes(h, 1));
    need = std::max(need, kbfv_keygen_rlk_workspace_bytes(h));
    ensure_ws(need);
}

void BFVScheme::addEncKey(SecretKey &sk, hipStream_t stream) {
    if (!publicKey) publicKey = new Key((int)context.N, context.L);
    uint8_t seed[32];
    seed_next(seed);
    ok(kbfv_keygen_pk(context.handle, seed, (uint64_t *)sk.sx_device, (uint64_t *)publicKey->cipher_device, stream),
       "public key");
    hip_ok(hipStreamSynchronize(stream), "public key");
}

void BFVScheme::addMultKey_23(SecretKey &sk, hipStream_t stream) {
    if (rlk_23) return;  // as the reference: generated once
    rlk_23 = new Key_decomp((int)context.N, context.dnum, context.t_num, context.Ri_blockNum);
    uint8_t seed[32];
    seed_next(seed);
    void *w = ensure_ws(kbfv_keygen_rlk_workspace_bytes(context.handle));
    ok(kbfv_keygen_rlk(context.handle, seed, (uint64_t *)sk.sx_device, (uint64_t *)rlk_23->cipher_device, w,
                       ws_bytes, stream),
       "relinearisation key");
    hip_ok(hipStreamSynchronize(stream), "relinearisation key");
}

Key_decomp *BFVScheme::galois_key(SecretKey &sk, uint64_t g, hipStream_t stream) {
    auto *key = new Key_decomp((int)context.N, context.dnum, context.t_num, context.Ri_blockNum);
    uint8_t seed[32];
    seed_next(seed);
    void *w = ensure_ws(kbfv_keygen_galois_workspace_bytes(context.handle));
    ok(kbfv_keygen_galois(context.handle, seed, (uint64_t *)sk.sx_device, g, (uint64_t *)key->cipher_device, w,
                          ws_bytes, stream),
       "Galois key");
    hip_ok(hipStreamSynchronize(stream), "Galois key");
    return key;
}

void BFVScheme::addLeftRotKey_23(SecretKey &sk, long rot, hipStream_t stream) {
    if (rot < 0 || rot >= (long)rotKey_vec_23.size()) throw std::invalid_argument("rotation out of range");
    if (rotKey_vec_23[rot]) return;  // as the reference: generated once
    rotKey_vec_23[rot] = galois_key(sk, kbfv_galois_elt(context.handle, rot), stream);
}

static uint64_t auto_elt(const BFVContext &c, int d) { return ((1ULL << d) + 1) % (2 * (uint64_t)c.N); }

void BFVScheme::addAutoKey_23(SecretKey &sk, int d, hipStream_t stream) {
    if (d < 0 || d >= (int)autoKey_vec_23.size()) throw std::invalid_argument("autoKey only for d in [0, logN)");
    if (autoKey_vec_23[d]) return;
    autoKey_vec_23[d] = galois_key(sk, auto_elt(context, d), stream);
}

void BFVScheme::automorphismAndAdd(Ciphertext &cipher, int d) {
    if (d < 0 || d >= (int)autoKey_vec_23.size() || !autoKey_vec_23[d])
        throw std::invalid_argument("autoKey_23 not exists");
    Ciphertext tmp(cipher.N, cipher.L, cipher.l, cipher.slots);
    apply_galois(tmp, cipher, autoKey_vec_23[d], auto_elt(context, d));
    add(cipher, cipher, tmp);
    hip_ok(hipStreamSynchronize(context.stream), "automorphismAndAdd");
}

Key_decomp *BFVScheme::addSWKey_23(SecretKey &sk, uint64_tt *s2, hipStream_t stream) {
    auto *key = new Key_decomp((int)context.N, context.dnum, context.t_num, context.Ri_blockNum);
    uint8_t seed[32];
    seed_next(seed);
    try {
        void *w = ensure_ws(kbfv_keygen_rlk_workspace_bytes(context.handle));
        ok(kbfv_keygen_swk(context.handle, seed, (uint64_t *)sk.sx_device, (const uint64_t *)s2,
                           (uint64_t *)key->cipher_device, w, ws_bytes, stream),
           "switching key");
        hip_ok(hipStreamSynchronize(stream), "switching key");
    } catch (...) {
        delete key;
        throw;
    }
    return key;
}

void BFVScheme::addConjKey_23(SecretKey &sk, hipStream_t stream) {
    if (ConjKey_23) return;
    ConjKey_23 = galois_key(sk, kbfv_galois_elt(context.handle, -1), stream);
}

void BFVScheme::apply_galois(Ciphertext &res, Ciphertext &cipher, Key_decomp *key, uint64_t g) {
    if (cipher.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    res.l = cipher.l;
    void *w = ensure_ws(kbfv_rotate_workspace_bytes(context.handle, 1));
    ok(kbfv_rotate(context.handle, (uint64_t *)key->cipher_device, g, (uint64_t *)cipher.cipher_device,
                   (uint64_t *)res.cipher_device, 1, w, ws_bytes, context.stream),
       "rotate");
    hip_ok(hipStreamSynchronize(context.stream), "rotate");
}

void BFVScheme::leftRotateAndEqual_23(Ciphertext &cipher, long rotSlots) {
    if (rotSlots < 0 || rotSlots >= (long)rotKey_vec_23.size() || !rotKey_vec_23[rotSlots])
        throw std::invalid_argument("rotKey_23 not exists");
    apply_galois(cipher, cipher, rotKey_vec_23[rotSlots], kbfv_galois_elt(context.handle, rotSlots));
}

void BFVScheme::rightRotateAndEqual_23(Ciphertext &cipher, long rotSlots) {
    // include/KeySwitch_23.cuh:580-584
    leftRotateAndEqual_23(cipher, (long)(context.N / 2) - (1L << rotSlots));
}

void BFVScheme::conjugate_23(Ciphertext &cipher_res, Ciphertext &cipher) {
    if (!ConjKey_23) throw std::invalid_argument("conjKey_23 not exists");
    apply_galois(cipher_res, cipher, ConjKey_23, kbfv_galois_elt(context.handle, -1));
}

void BFVScheme::conjugateAndEqual_23(Ciphertext &cipher) { conjugate_23(cipher, cipher); }

void BFVScheme::encryptMsg(Ciphertext &cipher, Plaintext &message, hipStream_t stream) {
    if (!publicKey) throw std::invalid_argument("public key not generated (call addEncKey)");
    cipher.l = message.l;
    uint8_t seed[32];
    seed_next(seed);
    void *w = ensure_ws(kbfv_encrypt_workspace_bytes(context.handle, 1));
    ok(kbfv_encrypt(context.handle, seed, (uint64_t *)publicKey->cipher_device, (uint64_t *)message.mx_device,
                    (uint64_t *)cipher.cipher_device, 1, w, ws_bytes, stream),
       "encrypt");
    hip_ok(hipStreamSynchronize(stream), "encrypt");
}

void BFVScheme::encryptZero(Ciphertext &cipher, int l, int slots, hipStream_t stream) {
    if (!publicKey) throw std::invalid_argument("public key not generated (call addEncKey)");
    if (l != context.L) throw std::invalid_argument("encryptZero: this engine encrypts at the full level L");
    cipher.l = l;
    cipher.slots = slots;
    const size_t enc = kbfv_encrypt_workspace_bytes(context.handle, 1);
    // the zero message sits in the workspace, behind the encryption's own part
    uint8_t *w = (uint8_t *)ensure_ws(enc + 256 + (size_t)context.N * 8);
    uint64_t *zero = (uint64_t *)(w + ((enc + 255) & ~size_t(255)));
    hip_ok(hipMemsetAsync(zero, 0, (size_t)context.N * 8, stream), "encryptZero");
    uint8_t seed[32];
    seed_next(seed);
    ok(kbfv_encrypt(context.handle, seed, (uint64_t *)publicKey->cipher_device, zero,
                    (uint64_t *)cipher.cipher_device, 1, w, enc, stream),
       "encryptZero");
    hip_ok(hipStreamSynchronize(stream), "encryptZero");
}

void BFVScheme::hps_decrypt_scale_and_round(uint64_tt *dst, uint64_tt *src, const hipStream_t &stream) const {
    ok(kbfv_hps_decrypt(context.handle, (const uint64_t *)src, (uint64_t *)dst, 1, stream),
       "hps_decrypt_scale_and_round");
}

void BFVScheme::compute_t_QInv_c0c1c2(uint64_tt *axbx1, uint64_tt *axa, uint64_tt *bxbx, Ciphertext &c1,
                                      Ciphertext &c2) {
    if (c1.l == 0 || c2.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    c1.l = std::min(c1.l, c2.l);
    const size_t poly = (size_t)context.N * context.q_num;
    const size_t mws = kbfv_mult_workspace_bytes(context.handle, 1);
    uint8_t *w = (uint8_t *)ensure_ws(mws + 256 + 3 * poly * 8);
    uint64_t *out3 = (uint64_t *)(w + ((mws + 255) & ~size_t(255)));
    ok(kbfv_mult_tensor(context.handle, (uint64_t *)c1.cipher_device, (uint64_t *)c2.cipher_device, out3, 1, w,
                        mws, context.stream),
       "compute_t_QInv_c0c1c2");
    hip_ok(hipMemcpyAsync(axa, out3, poly * 8, hipMemcpyDeviceToDevice, context.stream), "tensor copy");
    hip_ok(hipMemcpyAsync(axbx1, out3 + poly, poly * 8, hipMemcpyDeviceToDevice, context.stream), "tensor copy");
    hip_ok(hipMemcpyAsync(bxbx, out3 + 2 * poly, poly * 8, hipMemcpyDeviceToDevice, context.stream), "tensor copy");
    hip_ok(hipStreamSynchronize(context.stream), "compute_t_QInv_c0c1c2");
}

// ---------------------------------------------------------------- levels
const kbfv_context *BFVScheme::level_view(int l) const {
    const kbfv_context *v = nullptr;
    ok(kbfv_context_level(context.handle, l, &v), "level");
    return v;
}

// c (level c.l >= l, limbs 0..c.l of each component at stride (L+1) N) ->
// dst [2][l + 1][N]; ws: scratch for the switch when c.l > l
void BFVScheme::to_compact(Ciphertext &c, int l, uint64_tt *dst, void *w, size_t wlen) {
    const size_t N = context.N, full = (size_t)(context.L + 1) * N;
    const size_t have = (size_t)(c.l + 1) * N;
    if (c.l == l) {
        hip_ok(hipMemcpy2DAsync(dst, have * 8, c.cipher_device, full * 8, have * 8, 2, hipMemcpyDeviceToDevice,
                                context.stream),
               "compact");
        return;
    }
    // the switch reads its own compact copy: [2][c.l + 1][N] then its workspace
    uint64_tt *src = (uint64_tt *)w;
    const size_t src_bytes = (2 * have * 8 + 255) & ~size_t(255);
    hip_ok(hipMemcpy2DAsync(src, have * 8, c.cipher_device, full * 8, have * 8, 2, hipMemcpyDeviceToDevice,
                            context.stream),
           "compact");
    ok(kbfv_mod_switch(context.handle, (uint64_t *)src, c.l, (uint64_t *)dst, l, 1, (uint8_t *)w + src_bytes,
                       wlen - src_bytes, context.stream),
       "mod switch");
}

void BFVScheme::from_compact(const uint64_tt *src, int l, Ciphertext &c) {
    const size_t N = context.N, full = (size_t)(context.L + 1) * N, have = (size_t)(l + 1) * N;
    hip_ok(hipMemcpy2DAsync(c.cipher_device, full * 8, src, have * 8, have * 8, 2, hipMemcpyDeviceToDevice,
                            context.stream),
           "expand");
    c.l = l;
}

void BFVScheme::rescaleToAndEqual(Ciphertext &cipher, int level) {
    if (level < 0 || level > cipher.l) throw std::invalid_argument("rescaleToAndEqual: level out of range");
    if (level == cipher.l) return;
    const size_t N = context.N, ct = 2 * (size_t)(cipher.l + 1) * N * 8;
    const size_t msw = kbfv_mod_switch_workspace_bytes(context.handle, cipher.l, 1);
    uint8_t *w = (uint8_t *)ensure_ws(ct + 256 + ct + 256 + msw);
    uint64_tt *out = (uint64_tt *)w;
    to_compact(cipher, level, out, w + ((ct + 255) & ~size_t(255)), ws_bytes - ((ct + 255) & ~size_t(255)));
    from_compact(out, level, cipher);
    hip_ok(hipStreamSynchronize(context.stream), "rescaleTo");
}

void BFVScheme::rescaleAndEqual(Ciphertext &cipher) {
    if (cipher.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    rescaleToAndEqual(cipher, cipher.l - 1);
}

void BFVScheme::decryptMsg(Plaintext &plain, SecretKey &sk, Ciphertext &cipher, hipStream_t stream) {
    plain.l = cipher.l;
    plain.slots = cipher.slots;
    if (cipher.l < context.L) {
        const int l = cipher.l;
        const size_t ct = 2 * (size_t)(l + 1) * context.N * 8, dws = kbfv_decrypt_workspace_bytes(level_view(l), 1);
        uint8_t *w = (uint8_t *)ensure_ws(((ct + 255) & ~size_t(255)) + dws);
        uint64_tt *cc = (uint64_tt *)w;
        to_compact(cipher, l, cc, nullptr, 0);
        ok(kbfv_decrypt(level_view(l), (uint64_t *)sk.sx_device, (uint64_t *)cc, (uint64_t *)plain.mx_device, 1,
                        w + ((ct + 255) & ~size_t(255)), dws, context.stream),
           "decrypt");
    } else {
        void *w = ensure_ws(kbfv_decrypt_workspace_bytes(context.handle, 1));
        ok(kbfv_decrypt(context.handle, (uint64_t *)sk.sx_device, (uint64_t *)cipher.cipher_device,
                        (uint64_t *)plain.mx_device, 1, w, ws_bytes, stream),
           "decrypt");
    }
    // the result is replicated into every limb, as the reference's HPS decrypt does
    for (int i = 1; i <= plain.L; ++i)
        hip_ok(hipMemcpyAsync(plain.mx_device + (size_t)i * plain.N, plain.mx_device, sizeof(uint64_tt) * plain.N,
                              hipMemcpyDeviceToDevice, stream),
               "decrypt copy");
    hip_ok(hipStreamSynchronize(stream), "decrypt");
}

void BFVScheme::add(Ciphertext &res, Ciphertext &c1, Ciphertext &c2) {
    const int l = std::min(c1.l, c2.l);
    if (l == context.L) {
        res.l = l;
        ok(kbfv_add(context.handle, (uint64_t *)c1.cipher_device, (uint64_t *)c2.cipher_device,
                    (uint64_t *)res.cipher_device, 1, context.stream),
           "add");
        return;
    }
    const size_t ct = ((2 * (size_t)(context.L + 1) * context.N * 8) + 255) & ~size_t(255);
    const size_t msw = kbfv_mod_switch_workspace_bytes(context.handle, context.L, 1) + ct;
    uint8_t *w = (uint8_t *)ensure_ws(2 * ct + msw);
    uint64_tt *a = (uint64_tt *)w, *b = (uint64_tt *)(w + ct);
    to_compact(c1, l, a, w + 2 * ct, ws_bytes - 2 * ct);
    to_compact(c2, l, b, w + 2 * ct, ws_bytes - 2 * ct);
    ok(kbfv_add(level_view(l), (uint64_t *)a, (uint64_t *)b, (uint64_t *)a, 1, context.stream), "add");
    from_compact(a, l, res);
    hip_ok(hipStreamSynchronize(context.stream), "add");
}
void BFVScheme::addAndEqual(Ciphertext &c1, Ciphertext &c2) { add(c1, c1, c2); }
void BFVScheme::sub(Ciphertext &res, Ciphertext &c1, Ciphertext &c2) {
    const int l = std::min(c1.l, c2.l);
    if (l == context.L) {
        res.l = l;
        ok(kbfv_poly_sub(context.handle, (uint64_t *)c1.cipher_device, (uint64_t *)c2.cipher_device,
                         (uint64_t *)res.cipher_device, context.p_num, context.q_num, 2, context.stream),
           "sub");
        return;
    }
    const size_t ct = ((2 * (size_t)(context.L + 1) * context.N * 8) + 255) & ~size_t(255);
    const size_t msw = kbfv_mod_switch_workspace_bytes(context.handle, context.L, 1) + ct;
    uint8_t *w = (uint8_t *)ensure_ws(2 * ct + msw);
    uint64_tt *a = (uint64_tt *)w, *b = (uint64_tt *)(w + ct);
    to_compact(c1, l, a, w + 2 * ct, ws_bytes - 2 * ct);
    to_compact(c2, l, b, w + 2 * ct, ws_bytes - 2 * ct);
    ok(kbfv_poly_sub(level_view(l), (uint64_t *)a, (uint64_t *)b, (uint64_t *)a, context.p_num, l + 1, 2,
                     context.stream),
       "sub");
    from_compact(a, l, res);
    hip_ok(hipStreamSynchronize(context.stream), "sub");
}
void BFVScheme::subAndEqual(Ciphertext &c1, Ciphertext &c2) { sub(c1, c1, c2); }

// include/BFVScheme.cuh:192-326
void BFVScheme::negate(Ciphertext &res, Ciphertext &cipher) {
    res.l = cipher.l;
    ok(kbfv_negate(context.handle, (uint64_t *)cipher.cipher_device, (uint64_t *)res.cipher_device, 1,
                   context.stream),
       "negate");
}
void BFVScheme::negateAndEqual(Ciphertext &cipher) { negate(cipher, cipher); }

void BFVScheme::addConstAndEqual(Ciphertext &cipher, Plaintext &cnst) {
    cipher.l = std::min(cipher.l, cnst.l);
    void *w = ensure_ws(kbfv_add_plain_workspace_bytes(context.handle, 1));
    ok(kbfv_add_plain(context.handle, (uint64_t *)cipher.cipher_device, (uint64_t *)cnst.mx_device,
                      (uint64_t *)cipher.cipher_device, 1, w, ws_bytes, context.stream),
       "addConst");
    hip_ok(hipStreamSynchronize(context.stream), "addConst");
}

void BFVScheme::addConstAndEqual(Ciphertext &cipher, uint64_tt cnst) {
    // the constant polynomial cnst: every slot (batching) or the constant term
    std::vector<uint64_t> m((size_t)context.N, 0);
    m[0] = cnst % context.plain_modulus;
    uint64_t *dm = nullptr;
    hip_ok(hipMalloc(&dm, m.size() * 8), "hipMalloc");
    hip_ok(hipMemcpy(dm, m.data(), m.size() * 8, hipMemcpyHostToDevice), "copy");
    void *w = ensure_ws(kbfv_add_plain_workspace_bytes(context.handle, 1));
    const int rc = kbfv_add_plain(context.handle, (uint64_t *)cipher.cipher_device, dm,
                                  (uint64_t *)cipher.cipher_device, 1, w, ws_bytes, context.stream);
    (void)hipStreamSynchronize(context.stream);
    (void)hipFree(dm);
    ok(rc, "addConst");
}

void BFVScheme::mult_23(Ciphertext &res, Ciphertext &c1, Ciphertext &c2) {
    if (c1.l == 0 || c2.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    if (!rlk_23) throw std::invalid_argument("rotKey_23 not exists");
    const int l = std::min(c1.l, c2.l);
    if (l < context.L) {
        // both operands at level l (the higher one switched down), HomMult on the level view
        const kbfv_context *v = level_view(l);
        const size_t ct = ((2 * (size_t)(context.L + 1) * context.N * 8) + 255) & ~size_t(255);
        const size_t mws = kbfv_mult_workspace_bytes(v, 1);
        const size_t msw = kbfv_mod_switch_workspace_bytes(context.handle, context.L, 1) + ct;
        uint8_t *w = (uint8_t *)ensure_ws(2 * ct + std::max(mws, msw));
        uint64_tt *a = (uint64_tt *)w, *b = (uint64_tt *)(w + ct);
        to_compact(c1, l, a, w + 2 * ct, ws_bytes - 2 * ct);
        to_compact(c2, l, b, w + 2 * ct, ws_bytes - 2 * ct);
        ok(kbfv_mult_relin(v, (uint64_t *)rlk_23->cipher_device, (uint64_t *)a, (uint64_t *)b, (uint64_t *)a, 1,
                           w + 2 * ct, ws_bytes - 2 * ct, context.stream),
           "mult");
        from_compact(a, l, res);
        hip_ok(hipStreamSynchronize(context.stream), "mult");
        return;
    }
    res.l = l;
    void *w = ensure_ws(kbfv_mult_workspace_bytes(context.handle, 1));
    ok(kbfv_mult_relin(context.handle, (uint64_t *)rlk_23->cipher_device, (uint64_t *)c1.cipher_device,
                       (uint64_t *)c2.cipher_device, (uint64_t *)res.cipher_device, 1, w, ws_bytes, context.stream),
       "mult");
    hip_ok(hipStreamSynchronize(context.stream), "mult");
}

void BFVScheme::multAndEqual_23(Ciphertext &c1, Ciphertext &c2) { mult_23(c1, c1, c2); }

// The linear terms travel between the halves in the caller's two buffers (the
// reference's layout); the C ABI keeps them as one [2][q_num][N] block, so
// they are copied through the scheme-owned d01 on either side.
void BFVScheme::multAndEqual_beforeIP_23(Ciphertext &c1, Ciphertext &c2, uint64_tt *IP_input,
                                         uint64_tt *axbx1_mul_batch, uint64_tt *bxbx_mul_batch) {
    if (c1.l == 0 || c2.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    if (!rlk_23) throw std::invalid_argument("rotKey_23 not exists");
    c1.l = std::min(c1.l, c2.l);
    const size_t poly = (size_t)context.N * context.q_num;
    if (!d01) d01 = dalloc(2 * poly);
    void *w = ensure_ws(kbfv_mult_workspace_bytes(context.handle, 1));
    ok(kbfv_mult_before_ip(context.handle, (uint64_t *)c1.cipher_device, (uint64_t *)c2.cipher_device,
                           (uint64_t *)IP_input, (uint64_t *)d01, 1, w, ws_bytes, context.stream),
       "multAndEqual_beforeIP_23");
    hip_ok(hipMemcpyAsync(axbx1_mul_batch, d01, poly * 8, hipMemcpyDeviceToDevice, context.stream), "beforeIP");
    hip_ok(hipMemcpyAsync(bxbx_mul_batch, d01 + poly, poly * 8, hipMemcpyDeviceToDevice, context.stream), "beforeIP");
    hip_ok(hipStreamSynchronize(context.stream), "multAndEqual_beforeIP_23");
}

void BFVScheme::multAndEqual_afterIP_23(Ciphertext &c1, Ciphertext &c2, uint64_tt *IP_output,
                                        uint64_tt *axbx1_mul_batch, uint64_tt *bxbx_mul_batch) {
    if (c1.l == 0 || c2.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    c1.l = std::min(c1.l, c2.l);
    const size_t poly = (size_t)context.N * context.q_num;
    if (!d01) d01 = dalloc(2 * poly);
    hip_ok(hipMemcpyAsync(d01, axbx1_mul_batch, poly * 8, hipMemcpyDeviceToDevice, context.stream), "afterIP");
    hip_ok(hipMemcpyAsync(d01 + poly, bxbx_mul_batch, poly * 8, hipMemcpyDeviceToDevice, context.stream), "afterIP");
    void *w = ensure_ws(kbfv_mult_workspace_bytes(context.handle, 1));
    ok(kbfv_mult_after_ip(context.handle, (const uint64_t *)IP_output, (const uint64_t *)d01,
                          (uint64_t *)c1.cipher_device, 1, w, ws_bytes, context.stream),
       "multAndEqual_afterIP_23");
    hip_ok(hipStreamSynchronize(context.stream), "multAndEqual_afterIP_23");
}

// All pairs in one batched pipeline (the relinearisation key is read once per
// chunk).  Ciphertexts that already form one contiguous array (c[i] at
// c[0] + i * 2 * N * (L+1) words) are used in place; otherwise they are
// gathered into scheme-owned buffers that grow once to the largest batch and
// are reused, so no call after the first allocates device memory.
void BFVScheme::multAndEqual_23_batch(std::vector<Ciphertext *> &c1, std::vector<Ciphertext *> &c2) {
    if (c1.size() != c2.size()) throw std::invalid_argument("batch sizes differ");
    if (!rlk_23) throw std::invalid_argument("rotKey_23 not exists");
    const int B = (int)c1.size();
    if (B == 0) return;
    for (int i = 0; i < B; ++i) {  // every check before any device work
        if (!c1[i] || !c2[i]) throw std::invalid_argument("null ciphertext");
        if (c1[i]->l == 0 || c2[i]->l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
        if (c1[i]->N != (int)context.N || c1[i]->L != context.L || c2[i]->N != (int)context.N ||
            c2[i]->L != context.L)
            throw std::invalid_argument("ciphertext shape differs from the context");
    }
    bool below = false;
    for (int i = 0; i < B; ++i) below = below || c1[i]->l < context.L || c2[i]->l < context.L;
    if (below) {  // pairs below the full level: one HomMult each, at its level
        for (int i = 0; i < B; ++i) mult_23(*c1[i], *c1[i], *c2[i]);
        return;
    }
    const size_t ct = (size_t)2 * context.N * (context.L + 1);
    auto contiguous = [&](std::vector<Ciphertext *> &v) {
        for (int i = 1; i < B; ++i)
            if (v[i]->cipher_device != v[0]->cipher_device + i * ct) return false;
        return true;
    };
    const bool in1 = contiguous(c1), in2 = contiguous(c2);
    if ((!in1 || !in2) && batch_cap < (size_t)B) {
        dfree(batch_a);
        dfree(batch_b);
        batch_a = batch_b = nullptr;
        batch_cap = 0;
        batch_a = dalloc(ct * B);
        batch_b = dalloc(ct * B);
        batch_cap = (size_t)B;
    }
    hipStream_t st = context.stream;
    uint64_tt *a = in1 ? c1[0]->cipher_device : batch_a, *b = in2 ? c2[0]->cipher_device : batch_b;
    for (int i = 0; i < B; ++i) {
        if (!in1) hip_ok(hipMemcpyAsync(a + i * ct, c1[i]->cipher_device, ct * 8, hipMemcpyDeviceToDevice, st), "gather");
        if (!in2) hip_ok(hipMemcpyAsync(b + i * ct, c2[i]->cipher_device, ct * 8, hipMemcpyDeviceToDevice, st), "gather");
    }
    // the workspace sized for the whole batch when it fits in 1/4 of free device
    // memory, else for chunks (kbfv_mult_relin chunks the batch to fit)
    size_t want = kbfv_mult_workspace_bytes(context.handle, B);
    if (want > ws_bytes) {
        size_t free_b = 0, total_b = 0;
        hip_ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const size_t one = kbfv_mult_workspace_bytes(context.handle, 1);
        want = std::max(one, std::min(want, free_b / 4));
    }
    void *w = ensure_ws(want);
    ok(kbfv_mult_relin(context.handle, (uint64_t *)rlk_23->cipher_device, (uint64_t *)a, (uint64_t *)b,
                       (uint64_t *)a, B, w, ws_bytes, st),
       "batched mult");
    for (int i = 0; i < B; ++i) {
        c1[i]->l = std::min(c1[i]->l, c2[i]->l);
        if (!in1) hip_ok(hipMemcpyAsync(c1[i]->cipher_device, a + i * ct, ct * 8, hipMemcpyDeviceToDevice, st), "scatter");
    }
    hip_ok(hipStreamSynchronize(st), "batched mult");
}

void BFVScheme::multConstAndEqual(Ciphertext &cipher, uint64_tt cnst) {
    if (cnst == 1) return;
    if (cipher.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    ok(kbfv_mul_scalar(context.handle, (uint64_t *)cipher.cipher_device, cnst, nullptr,
                       (uint64_t *)cipher.cipher_device, 1, context.stream),
       "multConst");
}

void BFVScheme::multConstAndAddCipherEqual(Ciphertext &c1, Ciphertext &c2, uint64_tt cnst) {
    if (c2.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    c1.l = std::min(c1.l, c2.l);
    ok(kbfv_mul_scalar(context.handle, (uint64_t *)c2.cipher_device, cnst, (uint64_t *)c1.cipher_device,
                       (uint64_t *)c1.cipher_device, 1, context.stream),
       "multConstAndAddCipher");
}

// include/KeySwitch_23.cuh:375-466 (no secret-key shortcut there either)
void BFVScheme::square(Ciphertext &res, Ciphertext &cipher) { mult_23(res, cipher, cipher); }
void BFVScheme::squareAndEqual(Ciphertext &cipher) { mult_23(cipher, cipher, cipher); }

void BFVScheme::multConstAndEqual(Ciphertext &cipher, Plaintext &cnst) {
    if (cipher.l == 0) throw std::invalid_argument("Ciphertexts are on level 0");
    const size_t n = (size_t)context.N * (context.L + 1);
    ok(kbfv_poly_mul(context.handle, (uint64_t *)cipher.ax_device, (uint64_t *)cnst.mx_device,
                     (uint64_t *)cipher.ax_device, context.p_num, context.q_num, 1, context.stream),
       "multConst");
    ok(kbfv_poly_mul(context.handle, (uint64_t *)cipher.bx_device, (uint64_t *)cnst.mx_device,
                     (uint64_t *)cipher.bx_device, context.p_num, context.q_num, 1, context.stream),
       "multConst");
    (void)n;
    hip_ok(hipStreamSynchronize(context.stream), "multConst");
}

// ---------------------------------------------------------------- wire format
namespace {

std::vector<uint8_tt> wire_save(BFVContext &c, int kind, const uint64_tt *dev, int count, uint64_t aux) {
    const size_t nb = kbfv_wire_bytes(c.handle, kind, count, aux);
    if (nb == 0) throw std::runtime_error(std::string("wire size: ") + kbfv_last_error());
    std::vector<uint8_tt> blob(nb);
    ok(kbfv_wire_save(c.handle, kind, (const uint64_t *)dev, count, aux, blob.data(), nb, c.stream), "wire save");
    return blob;
}

void wire_load(BFVContext &c, const std::vector<uint8_tt> &blob, int kind, uint64_tt *dev, int count) {
    ok(kbfv_wire_load(c.handle, blob.data(), blob.size(), kind, (uint64_t *)dev, count, c.stream), "wire load");
}

}  // namespace

std::vector<uint8_tt> BFVContext::save(const Ciphertext &c) {
    return wire_save(*this, KBFV_WIRE_CIPHERTEXT, c.cipher_device, 1, 0);
}
std::vector<uint8_tt> BFVContext::save(const Plaintext &m) {
    return wire_save(*this, KBFV_WIRE_PLAINTEXT, m.mx_device, 1, 0);
}
std::vector<uint8_tt> BFVContext::save(const SecretKey &sk) {
    return wire_save(*this, KBFV_WIRE_SECRET_KEY, sk.sx_device, 1, 0);
}
std::vector<uint8_tt> BFVContext::save(const Key &pk) {
    return wire_save(*this, KBFV_WIRE_PUBLIC_KEY, pk.cipher_device, 1, 0);
}
std::vector<uint8_tt> BFVContext::save(const Key_decomp &key, uint64_tt galois_elt) {
    return wire_save(*this, galois_elt ? KBFV_WIRE_GALOIS_KEY : KBFV_WIRE_RELIN_KEY, key.cipher_device, 1,
                     galois_elt);
}
void BFVContext::load(const std::vector<uint8_tt> &blob, Ciphertext &c) {
    wire_load(*this, blob, KBFV_WIRE_CIPHERTEXT, c.cipher_device, 1);
}
void BFVContext::load(const std::vector<uint8_tt> &blob, Plaintext &m) {
    wire_load(*this, blob, KBFV_WIRE_PLAINTEXT, m.mx_device, 1);
    for (int i = 1; i <= m.L; ++i)  // replicate into every limb (encode's layout)
        hip_ok(hipMemcpyAsync(m.mx_device + (size_t)i * m.N, m.mx_device, (size_t)m.N * 8, hipMemcpyDeviceToDevice,
                              stream),
               "plaintext limbs");
    hip_ok(hipStreamSynchronize(stream), "plaintext limbs");
}
void BFVContext::load(const std::vector<uint8_tt> &blob, SecretKey &sk) {
    wire_load(*this, blob, KBFV_WIRE_SECRET_KEY, sk.sx_device, 1);
}
void BFVContext::load(const std::vector<uint8_tt> &blob, Key &pk) {
    wire_load(*this, blob, KBFV_WIRE_PUBLIC_KEY, pk.cipher_device, 1);
}
uint64_tt BFVContext::load(const std::vector<uint8_tt> &blob, Key_decomp &key) {
    int kind = 0, count = 0;
    uint64_t aux = 0;
    ok(kbfv_wire_peek(blob.data(), blob.size(), &kind, &count, &aux), "wire peek");
    if (kind != KBFV_WIRE_RELIN_KEY && kind != KBFV_WIRE_GALOIS_KEY)
        throw std::invalid_argument("wire blob holds no key-switching key");
    wire_load(*this, blob, kind, key.cipher_device, 1);
    return kind == KBFV_WIRE_GALOIS_KEY ? aux : 0;
}
