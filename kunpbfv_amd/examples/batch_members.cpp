// The batched HomMult entry and the BFVScheme members a reference caller can
// reach besides multAndEqual_23 (include/BFVScheme.h:55-174), on the C++
// drop-in API, each checked on the host:
//   multAndEqual_23_batch   equals mult_23 bit for bit on every pair (two calls,
//                           the second smaller: the gather buffers are reused)
//   encryptZero             decrypts to zero
//   compute_t_QInv_c0c1c2 + hps_decrypt_scale_and_round
//                           axa*s^2 + axbx1*s + bxbx rounds to m1*m2
//   addSWKey_23(s^2)        relinearises like the relinearisation key
//   addAutoKey_23 / automorphismAndAdd
//                           decrypts to m(X) + m(X^(2^d+1)) slot by slot
//   rotKey_vec_23           a vector indexed by the rotation amount
//   rescaleToAndEqual / rescaleAndEqual
//                           switched-down ciphertexts decrypt, add across levels
//                           and multiply below the full level
// Usage: batch_members [N]   (default 16384)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
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
    BFVScheme scheme(context);
    uint8_t kseed[32];
    for (int i = 0; i < 32; ++i) kseed[i] = (uint8_t)(7 * i + 1);
    SecretKey sk(context, kseed);
    scheme.setSeed(42);
    scheme.mallocMemory();
    scheme.addEncKey(sk);
    scheme.addMultKey_23(sk);

    const int N = (int)context.N, L = context.L, slots = context.slots, qn = context.q_num;
    const uint64_tt t = context.plain_modulus;
    int logn = 0;
    while ((1 << logn) < N) ++logn;
    const size_t ctw = (size_t)2 * N * (L + 1), poly = (size_t)N * qn;
    int total_wrong = 0;

    uint64_tt *dv, *dd;
    (void)hipMalloc(&dv, 8 * slots);
    (void)hipMalloc(&dd, 8 * slots);
    Plaintext pt(N, L, L, slots), pd(N, L, L, slots);
    auto encrypt_vec = [&](Ciphertext &c, const std::vector<uint64_tt> &m) {
        (void)hipMemcpy(dv, m.data(), 8 * slots, hipMemcpyHostToDevice);
        context.encode(dv, pt);
        scheme.encryptMsg(c, pt);
    };
    auto decrypt_vec = [&](Ciphertext &c) {
        scheme.decryptMsg(pd, sk, c);
        context.decode(pd, dd);
        std::vector<uint64_tt> got(slots);
        (void)hipMemcpy(got.data(), dd, 8 * slots, hipMemcpyDeviceToHost);
        return got;
    };
    auto report = [&](const char *what, int bad) {
        printf("%-34s wrong=%d\n", what, bad);
        total_wrong += bad;
    };
    auto host = [&](const uint64_tt *d, size_t words) {
        std::vector<uint64_tt> h(words);
        (void)hipMemcpy(h.data(), d, 8 * words, hipMemcpyDeviceToHost);
        return h;
    };

    // ---- multAndEqual_23_batch against mult_23, twice (5 pairs, then 3)
    srand(11);
    const int B = 5;
    std::vector<std::vector<uint64_tt>> ms(2 * B, std::vector<uint64_tt>(slots));
    std::vector<std::unique_ptr<Ciphertext>> a, b, ref;
    for (int i = 0; i < B; ++i) {
        for (auto *m : {&ms[2 * i], &ms[2 * i + 1]})
            for (auto &v : *m) v = rand() % t;
        a.emplace_back(new Ciphertext(N, L, L, slots));
        b.emplace_back(new Ciphertext(N, L, L, slots));
        ref.emplace_back(new Ciphertext(N, L, L, slots));
        encrypt_vec(*a[i], ms[2 * i]);
        encrypt_vec(*b[i], ms[2 * i + 1]);
        scheme.mult_23(*ref[i], *a[i], *b[i]);
    }
    for (int round = 0; round < 2; ++round) {
        const int nb = round == 0 ? B : 3;
        std::vector<std::unique_ptr<Ciphertext>> x;
        std::vector<Ciphertext *> v1, v2;
        for (int i = 0; i < nb; ++i) {
            x.emplace_back(new Ciphertext(*a[i]));  // separate allocations: the gather path
            v1.push_back(x.back().get());
            v2.push_back(b[i].get());
        }
        scheme.multAndEqual_23_batch(v1, v2);
        int bad = 0;
        for (int i = 0; i < nb; ++i)
            bad += host(v1[i]->cipher_device, ctw) != host(ref[i]->cipher_device, ctw);
        report(round == 0 ? "multAndEqual_23_batch (5 pairs)" : "multAndEqual_23_batch (3 pairs)", bad);
    }
    {
        std::vector<uint64_tt> got = decrypt_vec(*ref[0]);
        int bad = 0;
        for (int i = 0; i < slots; ++i) bad += got[i] != (ms[0][i] * ms[1][i]) % t;
        report("mult_23 decrypts to m1*m2", bad);
    }

    // ---- encryptZero
    {
        Ciphertext z(N, L, L, slots);
        scheme.encryptZero(z, L, slots);
        std::vector<uint64_tt> got = decrypt_vec(z);
        int bad = 0;
        for (auto v : got) bad += v != 0;
        report("encryptZero", bad);
    }

    // ---- compute_t_QInv_c0c1c2 + hps_decrypt_scale_and_round
    {
        uint64_tt *axa, *axbx, *bxbx, *x, *tmp;
        (void)hipMalloc(&axa, 8 * poly);
        (void)hipMalloc(&axbx, 8 * poly);
        (void)hipMalloc(&bxbx, 8 * poly);
        (void)hipMalloc(&x, 8 * poly);
        (void)hipMalloc(&tmp, 8 * poly);
        Ciphertext c1(*a[0]);
        scheme.compute_t_QInv_c0c1c2(axbx, axa, bxbx, c1, *b[0]);
        // x = (axa * s + axbx1) * s + bxbx over Q, NTT form; then coefficients
        const uint64_tt *s_q = sk.sx_device + (size_t)context.p_num * N;
        kbfv_context *h = context.handle;
        const int oq = context.p_num;
        int rc = kbfv_poly_mul(h, (const uint64_t *)axa, (const uint64_t *)s_q, (uint64_t *)tmp, oq, qn, 1, 0);
        rc |= kbfv_poly_add(h, (const uint64_t *)tmp, (const uint64_t *)axbx, (uint64_t *)tmp, oq, qn, 1, 0);
        rc |= kbfv_poly_mul(h, (const uint64_t *)tmp, (const uint64_t *)s_q, (uint64_t *)tmp, oq, qn, 1, 0);
        rc |= kbfv_poly_add(h, (const uint64_t *)tmp, (const uint64_t *)bxbx, (uint64_t *)x, oq, qn, 1, 0);
        rc |= kbfv_ntt(h, -1, (const uint64_t *)x, (uint64_t *)x, oq, qn, 1, 1, 0, poly, N, 0, poly, N, 0);
        scheme.hps_decrypt_scale_and_round(dd, x, 0);
        rc |= kbfv_decode(h, (uint64_t *)dd, 1, 0);
        (void)hipDeviceSynchronize();
        std::vector<uint64_tt> got(slots);
        (void)hipMemcpy(got.data(), dd, 8 * slots, hipMemcpyDeviceToHost);
        int bad = rc != 0;
        for (int i = 0; i < slots; ++i) bad += got[i] != (ms[0][i] * ms[1][i]) % t;
        report("compute_t_QInv_c0c1c2 + HPS round", bad);
        (void)hipFree(axa), (void)hipFree(axbx), (void)hipFree(bxbx), (void)hipFree(x), (void)hipFree(tmp);
    }

    // ---- addSWKey_23 with s^2 as the source: a relinearisation key
    {
        uint64_tt *s2;
        (void)hipMalloc(&s2, 8 * poly);
        const uint64_tt *s_q = sk.sx_device + (size_t)context.p_num * N;
        (void)kbfv_poly_mul(context.handle, (const uint64_t *)s_q, (const uint64_t *)s_q, (uint64_t *)s2,
                            context.p_num, qn, 1, 0);
        Key_decomp *swk = scheme.addSWKey_23(sk, s2);
        Key_decomp *keep = scheme.rlk_23;
        scheme.rlk_23 = swk;
        Ciphertext r(N, L, L, slots);
        scheme.mult_23(r, *a[1], *b[1]);
        scheme.rlk_23 = keep;
        std::vector<uint64_tt> got = decrypt_vec(r);
        int bad = host(swk->cipher_device, 8) == host(keep->cipher_device, 8);  // fresh randomness
        for (int i = 0; i < slots; ++i) bad += got[i] != (ms[2][i] * ms[3][i]) % t;
        report("addSWKey_23(s^2) relinearises", bad);
        delete swk;
        (void)hipFree(s2);
    }

    // ---- addAutoKey_23 / automorphismAndAdd: m(X) + m(X^g), g = 2^d + 1
    {
        const int d = 2;
        const uint64_tt g = (1ULL << d) + 1, m2n = 2 * (uint64_tt)N;
        scheme.addAutoKey_23(sk, d);
        int bad = scheme.autoKey_vec_23.size() != (size_t)logn || !scheme.autoKey_vec_23[d];
        Ciphertext c(*a[2]);
        scheme.automorphismAndAdd(c, d);
        std::map<uint64_tt, int> at;
        for (int j = 0; j < N; ++j) at[(2 * brv(j, logn) + 1) % m2n] = j;
        std::vector<uint64_tt> got = decrypt_vec(c);
        const std::vector<uint64_tt> &m = ms[4];
        for (int j = 0; j < N; ++j) bad += got[j] != (m[j] + m[at[(2 * brv(j, logn) + 1) * g % m2n]]) % t;
        report("automorphismAndAdd(d = 2)", bad);
    }

    // ---- rotKey_vec_23: the reference's vector, indexed by the rotation amount
    {
        scheme.addLeftRotKey_23(sk, 3);
        int bad = scheme.rotKey_vec_23.size() != (size_t)N || !scheme.rotKey_vec_23[3] || scheme.rotKey_vec_23[1];
        report("rotKey_vec_23 layout", bad);
    }

    // ---- levels: rescaleToAndEqual, then decrypt, add and HomMult below the
    // full level (operands of two levels meet at the lower one), rescaleAndEqual
    if (L >= 2) {
        std::vector<uint64_tt> m1(slots), m2(slots);
        for (auto &v : m1) v = (uint64_tt)rand() % t;
        for (auto &v : m2) v = (uint64_tt)rand() % t;
        Ciphertext a(N, L, L, slots), b(N, L, L, slots), s(N, L, L, slots);
        encrypt_vec(a, m1);
        encrypt_vec(b, m2);
        scheme.rescaleToAndEqual(a, L - 1);  // a at L - 1, b at L
        int bad = a.l != L - 1;
        auto g = decrypt_vec(a);
        for (int i = 0; i < slots; ++i) bad += g[i] != m1[i];
        report("rescaleToAndEqual + decrypt", bad);
        scheme.add(s, a, b);
        bad = s.l != L - 1;
        g = decrypt_vec(s);
        for (int i = 0; i < slots; ++i) bad += g[i] != (m1[i] + m2[i]) % t;
        report("add across levels", bad);
        scheme.multAndEqual_23(a, b);
        bad = a.l != L - 1;
        g = decrypt_vec(a);
        for (int i = 0; i < slots; ++i) bad += g[i] != (m1[i] * m2[i]) % t;
        report("multAndEqual_23 below the full level", bad);
        scheme.rescaleAndEqual(a);
        bad = a.l != L - 2;
        g = decrypt_vec(a);
        for (int i = 0; i < slots; ++i) bad += g[i] != (m1[i] * m2[i]) % t;
        report("rescaleAndEqual after the product", bad);
    }

    printf("total_wrong=%d\n", total_wrong);
    return total_wrong ? 1 : 0;
}
