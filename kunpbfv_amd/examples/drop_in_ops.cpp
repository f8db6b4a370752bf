// The rest of the reference's BFVScheme surface on the C++ drop-in API:
// left rotation, conjugation, negation, constant multiply / add, plaintext add,
// square.  Each result is decrypted and checked slot by slot on the host.
// Slots are in the reference's order (NTT index j is the evaluation at
// psi^(2 brv(j) + 1)), so the automorphism X -> X^g moves slot j to the slot
// whose point is psi^((2 brv(j) + 1) g).
// Usage: drop_in_ops [N]   (default 16384)
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "kunpbfv/bfv.hpp"

static uint64_tt brv(uint64_tt x, int bits) {
    uint64_tt r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

int main(int argc, char *argv[]) {
    size_t n_arg = argc > 1 ? strtoull(argv[1], nullptr, 10) : 16384;
    BFVContext context(n_arg);
    context.autotuneNTT(2);  // everything below runs on the tuned NTT plans
    BFVScheme scheme(context);
    SecretKey sk(context);
    scheme.mallocMemory();
    scheme.addEncKey(sk);
    scheme.addMultKey_23(sk);
    scheme.addLeftRotKey_23(sk, 1);
    scheme.addConjKey_23(sk);

    const int N = (int)context.N, L = context.L, slots = context.slots;
    const uint64_tt t = context.plain_modulus;
    int logn = 0;
    while ((1 << logn) < N) ++logn;
    std::vector<uint64_tt> m1(slots), m2(slots);
    srand(3);
    for (int i = 0; i < slots; ++i) m1[i] = rand() % t, m2[i] = rand() % t;

    uint64_tt *d1, *d2, *dd;
    (void)hipMalloc(&d1, 8 * slots);
    (void)hipMalloc(&d2, 8 * slots);
    (void)hipMalloc(&dd, 8 * slots);
    auto up = [&](uint64_tt *d, const std::vector<uint64_tt> &h) {
        (void)hipMemcpy(d, h.data(), 8 * slots, hipMemcpyHostToDevice);
    };
    Plaintext p1(N, L, L, slots), p2(N, L, L, slots), pd(N, L, L, slots);
    Ciphertext c(N, L, L, slots), c2(N, L, L, slots), r(N, L, L, slots);
    up(d1, m1);
    up(d2, m2);
    context.encode(d1, p1);
    context.encode(d2, p2);
    auto fresh = [&]() { scheme.encryptMsg(c, p1); };
    auto check = [&](const char *what, Ciphertext &ct, const std::vector<uint64_tt> &want) {
        scheme.decryptMsg(pd, sk, ct);
        context.decode(pd, dd);
        std::vector<uint64_tt> got(slots);
        (void)hipMemcpy(got.data(), dd, 8 * slots, hipMemcpyDeviceToHost);
        int bad = 0;
        for (int i = 0; i < slots; ++i) bad += got[i] != want[i];
        printf("%-22s slots_wrong=%d\n", what, bad);
        return bad;
    };
    // slot permutation of X -> X^g in the reference's slot order
    auto perm = [&](uint64_tt g) {
        std::map<uint64_tt, int> at;
        const uint64_tt m2n = 2 * (uint64_tt)N;
        for (int j = 0; j < N; ++j) at[(2 * brv(j, logn) + 1) % m2n] = j;
        std::vector<uint64_tt> w(slots);
        for (int j = 0; j < N; ++j) w[j] = m1[at[(2 * brv(j, logn) + 1) * g % m2n]];
        return w;
    };
    int bad = 0;
    std::vector<uint64_tt> w(slots);

    fresh();
    scheme.leftRotateAndEqual_23(c, 1);
    uint64_tt g5 = 5;
    bad += check("leftRotateAndEqual_23", c, perm(g5));

    fresh();
    scheme.conjugate_23(r, c);
    bad += check("conjugate_23", r, perm(2 * (uint64_tt)N - 1));

    fresh();
    scheme.negateAndEqual(c);
    for (int i = 0; i < slots; ++i) w[i] = (t - m1[i]) % t;
    bad += check("negateAndEqual", c, w);

    fresh();
    scheme.multConstAndEqual(c, (uint64_tt)12345);
    for (int i = 0; i < slots; ++i) w[i] = (uint64_tt)((unsigned __int128)m1[i] * 12345 % t);
    bad += check("multConstAndEqual", c, w);

    fresh();
    scheme.addConstAndEqual(c, (uint64_tt)777);
    for (int i = 0; i < slots; ++i) w[i] = (m1[i] + 777) % t;
    bad += check("addConstAndEqual(u64)", c, w);

    fresh();
    scheme.addConstAndEqual(c, p2);
    for (int i = 0; i < slots; ++i) w[i] = (m1[i] + m2[i]) % t;
    bad += check("addConstAndEqual(pt)", c, w);

    fresh();
    scheme.encryptMsg(c2, p2);
    scheme.multConstAndAddCipherEqual(c, c2, 3);
    for (int i = 0; i < slots; ++i) w[i] = (uint64_tt)((m1[i] + (unsigned __int128)3 * m2[i]) % t);
    bad += check("multConstAndAddCipher", c, w);

    fresh();
    scheme.squareAndEqual(c);
    for (int i = 0; i < slots; ++i) w[i] = (uint64_tt)((unsigned __int128)m1[i] * m1[i] % t);
    bad += check("squareAndEqual", c, w);

    // wire format: ciphertext, plaintext and secret key through blobs
    fresh();
    {
        std::vector<uint8_tt> blob = context.save(c);
        Ciphertext back(N, L, L, slots);
        context.load(blob, back);
        bad += check("wire ciphertext", back, m1);
        SecretKey sk2(context);
        context.load(context.save(sk), sk2);
        scheme.decryptMsg(pd, sk2, back);
        context.decode(pd, dd);
        std::vector<uint64_tt> got(slots);
        (void)hipMemcpy(got.data(), dd, 8 * slots, hipMemcpyDeviceToHost);
        int wrong = 0;
        for (int i = 0; i < slots; ++i) wrong += got[i] != m1[i];
        printf("%-22s slots_wrong=%d\n", "wire secret key", wrong);
        bad += wrong;
        Plaintext pl(N, L, L, slots);
        context.load(context.save(p2), pl);
        scheme.encryptMsg(c2, pl);
        bad += check("wire plaintext", c2, m2);
        int rejected = 0;
        try {
            Key pk_wrong(N, L);
            context.load(blob, pk_wrong);  // a ciphertext blob is no public key
        } catch (const std::runtime_error &) {
            rejected = 1;
        }
        printf("%-22s %s\n", "wire kind check", rejected ? "rejected" : "ACCEPTED");
        bad += !rejected;
    }

    // split-IP seam over two pairs with one key-stationary inner product
    // between the halves (the reference's include/advanced/SchemeAlgo.cuh:460-490)
    {
        const size_t ipi = (size_t)context.Qj_blockNum * context.t_num * N;
        const size_t ipo = (size_t)2 * context.Ri_blockNum * context.t_num * N;
        const size_t lin = (size_t)context.q_num * N;
        uint64_tt *ip_in, *ip_out, *axbx1, *bxbx;
        (void)hipMalloc(&ip_in, 8 * 2 * ipi);
        (void)hipMalloc(&ip_out, 8 * 2 * ipo);
        (void)hipMalloc(&axbx1, 8 * 2 * lin);
        (void)hipMalloc(&bxbx, 8 * 2 * lin);
        Ciphertext a0(N, L, L, slots), b0(N, L, L, slots), a1(N, L, L, slots), b1(N, L, L, slots);
        scheme.encryptMsg(a0, p1);
        scheme.encryptMsg(b0, p2);
        scheme.encryptMsg(a1, p2);
        scheme.encryptMsg(b1, p2);
        Ciphertext *x[2] = {&a0, &a1}, *y[2] = {&b0, &b1};
        for (int i = 0; i < 2; ++i)
            scheme.multAndEqual_beforeIP_23(*x[i], *y[i], ip_in + i * ipi, axbx1 + i * lin, bxbx + i * lin);
        context.external_product_T_swk_reuse(ip_out, ip_in, scheme.rlk_23->cipher_device, L, 2);
        for (int i = 0; i < 2; ++i)
            scheme.multAndEqual_afterIP_23(*x[i], *y[i], ip_out + i * ipo, axbx1 + i * lin, bxbx + i * lin);
        std::vector<uint64_tt> w0(slots), w1(slots);
        for (int i = 0; i < slots; ++i) {
            w0[i] = (uint64_tt)((unsigned __int128)m1[i] * m2[i] % t);
            w1[i] = (uint64_tt)((unsigned __int128)m2[i] * m2[i] % t);
        }
        bad += check("split-IP pair 0", a0, w0);
        bad += check("split-IP pair 1", a1, w1);
        (void)hipFree(ip_in);
        (void)hipFree(ip_out);
        (void)hipFree(axbx1);
        (void)hipFree(bxbx);
    }

    printf("N=%d total_wrong=%d\n", N, bad);
    (void)hipFree(d1);
    (void)hipFree(d2);
    (void)hipFree(dd);
    return bad == 0 ? 0 : 1;
}
