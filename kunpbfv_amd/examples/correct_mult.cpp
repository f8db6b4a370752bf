// HomMult correctness driver on the C++ drop-in API: the same calls, in the same
// order, as the reference's correct_mult.cu (encode -> encrypt -> multAndEqual_23
// -> multConstAndEqual -> decrypt -> decode; expected m1*m2*m1 mod t).
// Usage: correct_mult [N] [--no-const]   (default 32768; --no-const skips the
// plaintext multiply and expects m1*m2, for chains too short for two products)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "kunpbfv/bfv.hpp"

int main(int argc, char *argv[]) {
    size_t poly_modulus_degree = argc > 1 ? strtoull(argv[1], nullptr, 10) : 32768;
    const bool with_const = !(argc > 2 && std::string(argv[2]) == "--no-const");
    BFVContext context(poly_modulus_degree);
    BFVScheme scheme(context);
    SecretKey sk(context);
    scheme.mallocMemory();
    scheme.addEncKey(sk);
    scheme.addMultKey_23(sk);
    scheme.addLeftRotKey_23(sk, 1);

    int N = (int)context.N, L = context.L, slots = context.slots;
    std::vector<uint64_tt> mes1(slots), mes2(slots);
    srand(1);
    for (int i = 0; i < slots; ++i) {
        mes1[i] = rand() % context.plain_modulus;
        mes2[i] = rand() % context.plain_modulus;
    }
    uint64_tt *d_msg1, *d_msg2, *d_dec;
    (void)hipMalloc(&d_msg1, sizeof(uint64_tt) * slots);
    (void)hipMalloc(&d_msg2, sizeof(uint64_tt) * slots);
    (void)hipMalloc(&d_dec, sizeof(uint64_tt) * slots);
    (void)hipMemcpy(d_msg1, mes1.data(), sizeof(uint64_tt) * slots, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_msg2, mes2.data(), sizeof(uint64_tt) * slots, hipMemcpyHostToDevice);

    Plaintext plain1(N, L, L, slots), plain2(N, L, L, slots), plain1_ntt(N, L, L, slots);
    Ciphertext c1(N, L, L, slots), c2(N, L, L, slots);
    context.encode(d_msg1, plain1);
    context.encode(d_msg2, plain2);
    context.encode_ntt(d_msg1, plain1_ntt);
    scheme.encryptMsg(c1, plain1);
    scheme.encryptMsg(c2, plain2);
    scheme.multAndEqual_23(c1, c2);
    if (with_const) scheme.multConstAndEqual(c1, plain1_ntt);
    Plaintext plain_dec(N, L, L, slots);
    scheme.decryptMsg(plain_dec, sk, c1);
    context.decode(plain_dec, d_dec);
    std::vector<uint64_tt> host_dec(slots);
    (void)hipMemcpy(host_dec.data(), d_dec, sizeof(uint64_tt) * slots, hipMemcpyDeviceToHost);

    int bad = 0;
    for (int i = 0; i < slots; ++i) {
        unsigned __int128 e = (unsigned __int128)mes1[i] * mes2[i] % context.plain_modulus;
        if (with_const) e = e * mes1[i] % context.plain_modulus;
        bad += (uint64_tt)e != host_dec[i];
    }
    printf("expected: ");
    for (int i = 0; i < 8; ++i) {
        unsigned __int128 e = (unsigned __int128)mes1[i] * mes2[i] % context.plain_modulus;
        if (with_const) e = e * mes1[i] % context.plain_modulus;
        printf("%llu ", (unsigned long long)e);
    }
    printf("\nactual:   ");
    for (int i = 0; i < 8; ++i) printf("%llu ", (unsigned long long)host_dec[i]);
    printf("\nN=%d slots_wrong=%d\n", N, bad);
    (void)hipFree(d_msg1);
    (void)hipFree(d_msg2);
    (void)hipFree(d_dec);
    return bad == 0 ? 0 : 1;
}
