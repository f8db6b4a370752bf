// C++ drop-in surface of the kunpbfv engine: the classes a caller of the
// reference library (QiiPatri/kunpBFV) uses, with the same names, members and
// argument meaning, implemented on the C ABI of libkunpbfv.so (kunpbfv.h).
//
// Interfaces kept (reference file:line):
//   BFVContext    include/BFVcontext.h:17-391   (ctor, encode/encode_ntt/decode,
//                                                ToNTTInplace/FromNTTInplace, public counts)
//   BFVScheme     include/BFVScheme.h:12-179    (mallocMemory, addEncKey, addMultKey_23,
//                                                encryptMsg, decryptMsg, add*, mult*_23,
//                                                multConstAndEqual)
//   Ciphertext    include/Ciphertext.cuh:5-89   (cipher_device = [ax | bx], N, L, l, slots)
//   Plaintext     include/Plaintext.cuh:6-69    (mx_device, N, L, l, slots)
//   SecretKey     include/SecretKey.cuh:10-63   (sx_device over K+L+1 limbs, NTT form)
//   Key           include/Key.cuh:6-81          (public key [ax | bx])
//   Key_decomp    include/Key_decomp.cuh:6-70   (relinearisation key, [2][dnum][blocks][T][N])
// Errors: std::invalid_argument for level-0 ciphertexts and a missing
// relinearisation key (include/KeySwitch_23.cuh:246-253), std::runtime_error for
// a failed device call (the reference never checks them).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <map>
#include <vector>

#include "../kunpbfv.h"

typedef unsigned long long uint64_tt;
typedef unsigned int uint32_tt;
typedef unsigned char uint8_tt;

class Plaintext {
   public:
    uint64_tt *mx_device = nullptr;
    int N = 0, L = 0, l = 0, slots = 0;
    Plaintext() = default;
    Plaintext(int N, int L, int l, int slots);
    Plaintext(const Plaintext &o);
    Plaintext &operator=(const Plaintext &o);
    ~Plaintext();
};

class Ciphertext {
   public:
    uint64_tt *cipher_device = nullptr;
    uint64_tt *ax_device = nullptr;
    uint64_tt *bx_device = nullptr;
    int N = 0, l = 0, L = 0, slots = 0;
    Ciphertext() = default;
    Ciphertext(int N, int L, int l, int slots);
    Ciphertext(const Ciphertext &o);
    Ciphertext &operator=(const Ciphertext &o);
    ~Ciphertext();
};

class SecretKey;
class Key;
class Key_decomp;

class BFVContext {
   public:
    // default parameter set of the ring degree (logN 12..16)
    explicit BFVContext(size_t poly_modulus_degree);
    // explicit chain: q (ciphertext), p (special), t (auxiliary key-switch basis)
    BFVContext(int logN, const std::vector<uint64_tt> &q, const std::vector<uint64_tt> &p,
               const std::vector<uint64_tt> &t, int gamma, uint64_tt plain_modulus);
    ~BFVContext();
    BFVContext(const BFVContext &) = delete;
    BFVContext &operator=(const BFVContext &) = delete;

    // slots -> plaintext: INTT over t in place on `vals` (device, N values), then
    // the coefficients replicated into every limb of msg (include/Encoder.cuh:23)
    void encode(uint64_tt *vals, Plaintext &msg);
    // coefficients replicated into every limb and NTT'd over Q (include/Encoder.cuh:64)
    void encode_ntt(uint64_tt *vals, Plaintext &msg);
    // limb 0 of msg -> vals, NTT over t (include/Encoder.cuh:81)
    void decode(Plaintext &msg, uint64_tt *vals);
    // forward / inverse NTT of poly_num polynomials of poly_mod_len limbs, limbs
    // [start_mod_idx, start_mod_idx + mod_num) of the [P | Q] basis
    // (include/BFVcontext.cuh:1513-1552)
    void ToNTTInplace(uint64_tt *device_a, int start_poly_idx, int start_mod_idx, int poly_num, int mod_num,
                      int poly_mod_len);
    void FromNTTInplace(uint64_tt *device_a, int start_poly_idx, int start_mod_idx, int poly_num, int mod_num,
                        int poly_mod_len);

    // NTT plans (no counterpart in the reference, whose NTT has one shape):
    // time every plan of both directions at the transform counts HomMult+relin
    // of `batch` ciphertexts launches and keep the fastest per bucket
    // (kbfv_ntt_autotune).  Call it before other work on the context.
    void autotuneNTT(int batch);

    // key-stationary inner product over a batch (include/ExternalProduct.cuh:169):
    // cipher_modUp_QjtoT [batch_size][Qj_blockNum][t_num][N] (NTT over T, as
    // BFVScheme::multAndEqual_beforeIP_23 writes it) times swk_modUp_RitoT (a
    // Key_decomp's cipher_device) -> output [batch_size][2][Ri_blockNum][t_num][N];
    // BFV ciphertexts stay at full level, so l must be L
    void external_product_T_swk_reuse(uint64_tt *output, uint64_tt *cipher_modUp_QjtoT, uint64_tt *swk_modUp_RitoT,
                                      int l, int batch_size);

    // Wire format (no counterpart in the reference, which has no serialization;
    // kunpbfv.h KBFV_WIRE_*): blobs that load only into a context with the same
    // parameters.  A Plaintext travels as its limb-0 coefficients mod t and is
    // replicated into every limb on load (the encode() layout); an encode_ntt
    // plaintext (residues mod q) is refused by save with std::runtime_error.  A Key_decomp is
    // a relinearisation key, or a Galois key when galois_elt != 0.
    std::vector<uint8_tt> save(const Ciphertext &c);
    std::vector<uint8_tt> save(const Plaintext &m);
    std::vector<uint8_tt> save(const SecretKey &sk);
    std::vector<uint8_tt> save(const Key &pk);
    std::vector<uint8_tt> save(const Key_decomp &key, uint64_tt galois_elt = 0);
    void load(const std::vector<uint8_tt> &blob, Ciphertext &c);
    void load(const std::vector<uint8_tt> &blob, Plaintext &m);
    void load(const std::vector<uint8_tt> &blob, SecretKey &sk);
    void load(const std::vector<uint8_tt> &blob, Key &pk);
    // returns the Galois element the blob carries (0 for a relinearisation key)
    uint64_tt load(const std::vector<uint8_tt> &blob, Key_decomp &key);

    int logN, logNh, logslots, slots, L, q_num, K, dnum, alpha, p_num, t_num, r_num, mod_num, gamma;
    int Ri_blockNum, Qj_blockNum;
    long N, M, Nh;
    uint64_tt plain_modulus;
    std::vector<uint64_tt> qVec, pVec, tVec, rVec;

    kbfv_context *handle = nullptr;
    hipStream_t stream = nullptr;

   private:
    void init(int logN, const std::vector<uint64_tt> &q, const std::vector<uint64_tt> &p,
              const std::vector<uint64_tt> &t, int gamma, uint64_tt plain);
};

class SecretKey {
   public:
    uint64_tt *sx_device = nullptr;
    int N = 0, L = 0, K = 0;
    // a fresh key from OS entropy (the reference's SecretKey(context))
    explicit SecretKey(BFVContext &context, hipStream_t stream = 0);
    // a reproducible key from an explicit seed
    SecretKey(BFVContext &context, const uint8_t seed[32], hipStream_t stream = 0);
    ~SecretKey();
    SecretKey(const SecretKey &) = delete;
    SecretKey &operator=(const SecretKey &) = delete;
};

class Key {
   public:
    uint64_tt *cipher_device = nullptr, *ax_device = nullptr, *bx_device = nullptr;
    int N = 0, L = 0;
    Key(int N, int L);
    ~Key();
    Key(const Key &) = delete;
    Key &operator=(const Key &) = delete;
};

class Key_decomp {
   public:
    uint64_tt *cipher_device = nullptr, *ax_device = nullptr, *bx_device = nullptr;
    int N = 0, dnum = 0, t_num = 0, Ri_blockNum = 0;
    Key_decomp(int N, int dnum, int t_num, int Ri_blockNum);
    ~Key_decomp();
    Key_decomp(const Key_decomp &) = delete;
    Key_decomp &operator=(const Key_decomp &) = delete;
};

class BFVScheme {
   public:
    explicit BFVScheme(BFVContext &context);
    ~BFVScheme();
    BFVScheme(const BFVScheme &) = delete;
    BFVScheme &operator=(const BFVScheme &) = delete;

    // scratch for single-ciphertext operations (grows on demand for batches)
    void mallocMemory();
    // Every key / encryption draws its seed from a generator that starts from
    // OS entropy; setSeed restarts it from `seed`, which makes keys and
    // encryptions reproducible (tests, benchmarks).
    void setSeed(uint64_t seed);

    void addEncKey(SecretKey &secretKey, hipStream_t stream = 0);
    void addMultKey_23(SecretKey &secretKey, hipStream_t stream = 0);
    // Galois keys: left rotation by rot (g = 5^rot mod 2N) and conjugation (g = 2N - 1)
    void addLeftRotKey_23(SecretKey &secretKey, long rot, hipStream_t stream = 0);
    void addConjKey_23(SecretKey &secretKey, hipStream_t stream = 0);
    // automorphism X -> X^(2^d + 1) (include/KeySwitch_23.cuh:171-190, a stub there)
    void addAutoKey_23(SecretKey &secretKey, int d, hipStream_t stream = 0);
    // key-switching key from s2 (q_num limbs over Q, NTT form) to the secret key
    // (include/KeySwitch_23.cuh:6-16, a stub there); the caller owns the result
    Key_decomp *addSWKey_23(SecretKey &secretKey, uint64_tt *s2, hipStream_t stream = 0);

    // encryption of zero (include/BFVScheme.cuh:87-116); l must be the full level L
    // (bring a ciphertext down with rescaleToAndEqual)
    void encryptZero(Ciphertext &cipher, int l, int slots, hipStream_t stream = 0);
    // modulus switching (include/BFVScheme.cuh:450-487): drop the top limbs
    // down to `level` (rescaleAndEqual: one limb), x_i <- (x_i - x_l) q_l^-1 mod
    // q_i; the ciphertext keeps its buffer of L + 1 limbs per component and
    // sets cipher.l.  Ciphertexts below the full level decrypt, add, subtract,
    // negate and multiply (with the full-level relinearisation key) at their
    // level; operands of two different levels meet at the lower one.
    void rescaleToAndEqual(Ciphertext &cipher, int level);
    void rescaleAndEqual(Ciphertext &cipher);
    void encryptMsg(Ciphertext &cipher, Plaintext &message, hipStream_t stream = 0);
    void decryptMsg(Plaintext &plain, SecretKey &secretKey, Ciphertext &cipher, hipStream_t stream = 0);

    void add(Ciphertext &cipher_res, Ciphertext &cipher1, Ciphertext &cipher2);
    void addAndEqual(Ciphertext &cipher1, Ciphertext &cipher2);
    void sub(Ciphertext &cipher_res, Ciphertext &cipher1, Ciphertext &cipher2);
    void subAndEqual(Ciphertext &cipher1, Ciphertext &cipher2);
    void negate(Ciphertext &cipher_res, Ciphertext &cipher);
    void negateAndEqual(Ciphertext &cipher);
    // plaintext (from encode) or constant added to every slot
    void addConstAndEqual(Ciphertext &cipher, Plaintext &cnst);
    void addConstAndEqual(Ciphertext &cipher, uint64_tt cnst);

    void mult_23(Ciphertext &cipher_res, Ciphertext &cipher1, Ciphertext &cipher2);
    void multAndEqual_23(Ciphertext &cipher1, Ciphertext &cipher2);
    // the batched form: every kernel covers all pairs, the key is read once
    void multAndEqual_23_batch(std::vector<Ciphertext *> &c1, std::vector<Ciphertext *> &c2);
    // the split-IP seam (include/KeySwitch_23.cuh:312-372): beforeIP writes the
    // pair's inner-product input (Qj_blockNum*t_num*N words) and its linear terms
    // (q_num*N words each, opaque to the caller); once
    // context.external_product_T_swk_reuse has run over a batch of such inputs,
    // afterIP finishes cipher1 = cipher1 * cipher2 relinearised, adding the
    // linear terms back (the reference leaves that addition commented out)
    void multAndEqual_beforeIP_23(Ciphertext &cipher1, Ciphertext &cipher2, uint64_tt *IP_input,
                                  uint64_tt *axbx1_mul_batch, uint64_tt *bxbx_mul_batch);
    void multAndEqual_afterIP_23(Ciphertext &cipher1, Ciphertext &cipher2, uint64_tt *IP_output,
                                 uint64_tt *axbx1_mul_batch, uint64_tt *bxbx_mul_batch);
    // plaintext (from encode_ntt) times ciphertext (include/BFVScheme.cuh:226)
    void multConstAndEqual(Ciphertext &cipher, Plaintext &cnst);
    void multConstAndEqual(Ciphertext &cipher, uint64_tt cnst);
    void multConstAndAddCipherEqual(Ciphertext &c1, Ciphertext &c2, uint64_tt cnst);  // c1 += c2 * cnst
    void square(Ciphertext &cipher_res, Ciphertext &cipher);
    void squareAndEqual(Ciphertext &cipher);

    // automorphisms with key switching (include/KeySwitch_23.cuh:466-584)
    void leftRotateAndEqual_23(Ciphertext &cipher, long rotSlots);
    void rightRotateAndEqual_23(Ciphertext &cipher, long rotSlots);  // left by N/2 - 2^rotSlots
    void conjugate_23(Ciphertext &cipher_res, Ciphertext &cipher);
    void conjugateAndEqual_23(Ciphertext &cipher);
    // f(X) -> f(X) + f(X^(2^d + 1)) (include/KeySwitch_23.cuh:587-589, empty there)
    void automorphismAndAdd(Ciphertext &cipher, int d);

    // HPS decryption rounding of b + a*s in coefficient form (q_num limbs) into
    // dst [N] (include/BFVScheme.cuh:358-398)
    void hps_decrypt_scale_and_round(uint64_tt *dst, uint64_tt *src, const hipStream_t &stream) const;
    // the t/Q-scaled tensor of HomMult, each [q_num][N] in NTT form:
    // axbx1 = a1b2 + a2b1, axa = a1a2, bxbx = b1b2 (include/KeySwitch_23.cuh:703-926)
    void compute_t_QInv_c0c1c2(uint64_tt *axbx1, uint64_tt *axa, uint64_tt *bxbx, Ciphertext &cipher1,
                               Ciphertext &cipher2);

    BFVContext &context;
    Key *publicKey = nullptr;
    Key_decomp *rlk_23 = nullptr;
    std::vector<Key_decomp *> rotKey_vec_23;   // [N], by rotation amount (null: not generated)
    Key_decomp *ConjKey_23 = nullptr;
    std::vector<Key_decomp *> autoKey_vec_23;  // [logN], by d of X -> X^(2^d + 1)

   private:
    // level-l view of the context and the compact [2][l + 1][N] form of a
    // ciphertext (switched down from its own level when that is higher)
    const kbfv_context *level_view(int l) const;
    void to_compact(Ciphertext &c, int l, uint64_tt *dst, void *ws, size_t ws_len);
    void from_compact(const uint64_tt *src, int l, Ciphertext &c);
    void seed_next(uint8_t out[32]);
    Key_decomp *galois_key(SecretKey &sk, uint64_t g, hipStream_t stream);
    void apply_galois(Ciphertext &res, Ciphertext &cipher, Key_decomp *key, uint64_t g);
    void *ensure_ws(size_t bytes);
    uint64_t seed_state = 0;  // set from OS entropy by the constructor
    void *ws = nullptr;
    size_t ws_bytes = 0;
    uint64_tt *d01 = nullptr;  // [2][q_num][N] linear terms of the split-IP seam
    // gather buffers of multAndEqual_23_batch for callers whose ciphertexts are
    // not one contiguous array; grown once to the largest batch seen
    uint64_tt *batch_a = nullptr, *batch_b = nullptr;
    size_t batch_cap = 0;
};
