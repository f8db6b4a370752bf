/*
 * Oracle: scalar number theory and NTT twiddle tables.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows:
 *   modpow128 / mulMod128 / x_Shoup / modinv128      include/Utils.cuh:14-53
 *   findPrimeFactors / findPrimitiveRoot             include/Utils.cuh:55-88
 *   findMthRootOfUnity                               include/Utils.cuh:90-101
 *   fillTablePsi128_special / fillTablePsi_shoup128  include/Utils.cuh:171-191
 *   is_prime / get_primes_below                      include/uintarith_bfv.h:971-1016
 *   singleBarrett_new                                include/uint128.cuh:126-151
 *
 * Pinned by (tests/test_oracle_nt.py): Python big-integer arithmetic for
 * mulmod/powmod/inverse/Shoup/Barrett on random and edge operands; a
 * Miller-Rabin/trial-division cross-check of primality; the prime chains and
 * primitive roots the reference's own sets produce (R basis and psi of the
 * logN=15 chain, generator 10 of T0) recorded in tests/golden; the defining
 * property psi^N = -1 and psi^(2N) = 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

/* include/Utils.cuh:14 modpow128 */
uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = or_mulmod(r, a, q);
        a = or_mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

/* include/Utils.cuh:48 modinv128: Fermat inverse, q prime */
uint64_t or_invmod(uint64_t a, uint64_t q) { return or_powmod(a % q, q - 2, q); }

/* include/Utils.cuh:40 x_Shoup: floor(w * 2^64 / q) */
uint64_t or_shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

/* Deterministic Miller-Rabin; exact for all 64-bit n.  Gives the same answer as
 * the reference's trial division is_prime (include/uintarith_bfv.h:971). */
int or_is_prime(uint64_t n) {
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n % bases[i] == 0) return n == bases[i];
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        s++;
    }
    for (int i = 0; i < 12; i++) {
        uint64_t x = or_powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int ok = 0;
        for (int r = 1; r < s; r++) {
            x = or_mulmod(x, x, n);
            if (x == n - 1) {
                ok = 1;
                break;
            }
        }
        if (!ok) return 0;
    }
    return 1;
}

/* include/Utils.cuh:55 findPrimeFactors, including its loop bound
 * `i < sqrt(number)` evaluated in double on the shrinking cofactor. */
static int prime_factors(uint64_t number, uint64_t *fac) {
    int nf = 0;
    int has2 = 0;
    while (number % 2 == 0) {
        has2 = 1;
        number /= 2;
    }
    if (has2) fac[nf++] = 2;
    for (uint64_t i = 3; (double)i < sqrt((double)number); i++) {
        int hit = 0;
        while (number % i == 0) {
            hit = 1;
            number /= i;
        }
        if (hit) fac[nf++] = i;
    }
    if (number > 2) {
        int dup = 0;
        for (int k = 0; k < nf; k++) dup |= (fac[k] == number);
        if (!dup) fac[nf++] = number;
    }
    return nf;
}

/* include/Utils.cuh:71 findPrimitiveRoot: smallest r >= 2 with r^(phi/f) != 1
 * for every recorded factor f of phi = q - 1. */
uint64_t or_primitive_root(uint64_t q) {
    uint64_t fac[64];
    uint64_t phi = q - 1;
    int nf = prime_factors(phi, fac);
    for (uint64_t r = 2; r <= phi; r++) {
        int bad = 0;
        for (int k = 0; k < nf; k++) {
            if (or_powmod(r, phi / fac[k], q) == 1) {
                bad = 1;
                break;
            }
        }
        if (!bad) return r;
    }
    return (uint64_t)-1;
}

/* include/Utils.cuh:90 findMthRootOfUnity */
uint64_t or_mth_root(uint64_t m, uint64_t q) {
    if ((q - 1) % m != 0) return (uint64_t)-1;
    uint64_t g = or_primitive_root(q);
    return or_powmod(g, (q - 1) / m, q);
}

static int bit_count(uint64_t v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

/* include/uintarith_bfv.h:983 get_primes_below: start at upper_bound - 2N and
 * step down by 2N while above 2^(bits(upper_bound)-1). */
int or_primes_below(uint64_t ntt_size, uint64_t upper_bound, int count, uint64_t *out) {
    uint64_t factor = 2 * ntt_size;
    uint64_t value = upper_bound - factor;
    uint64_t lower = (uint64_t)1 << (bit_count(upper_bound) - 1);
    int got = 0;
    while (got < count && value > lower) {
        if (or_is_prime(value)) out[got++] = value;
        value -= factor;
    }
    return got == count ? 0 : -1;
}

/* Largest primes strictly below `below` that are 1 mod `step` (step a power of
 * two), in descending order.  With below = 2^60 and step = 2^17 this yields the
 * reference's tVec for the logN=15 set (include/Params.cuh:129-140). */
int or_primes_descending(uint64_t below, uint64_t step, int count, uint64_t *out) {
    uint64_t v = below - step + 1;
    int got = 0;
    while (got < count && v > step) {
        if (or_is_prime(v)) out[got++] = v;
        v -= step;
    }
    return got == count ? 0 : -1;
}

uint64_t or_bitrev(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

/* include/Utils.cuh:171 fillTablePsi128_special + :184 fillTablePsi_shoup128.
 * psi_tab[brv(i)] = psi^i, psiinv_tab[brv(i)] = psi^-i, and n^-1 folded into
 * psiinv_tab[1] (the last inverse stage's twiddle). */
void or_fill_tables(uint64_t q, uint64_t psi, int log_n, uint64_t *psi_tab,
                    uint64_t *psiinv_tab, uint64_t *psi_shoup, uint64_t *psiinv_shoup) {
    uint64_t n = (uint64_t)1 << log_n;
    uint64_t psiinv = or_invmod(psi, q);
    uint64_t ninv = or_invmod(n % q, q);
    psi_tab[0] = 1;
    psiinv_tab[0] = 1;
    for (uint64_t i = 1; i < n; i++) {
        uint64_t prev = or_bitrev(i - 1, log_n), next = or_bitrev(i, log_n);
        psi_tab[next] = or_mulmod(psi, psi_tab[prev], q);
        psiinv_tab[next] = or_mulmod(psiinv, psiinv_tab[prev], q);
    }
    if (n > 1) psiinv_tab[1] = or_mulmod(psiinv_tab[1], ninv, q);
    if (psi_shoup && psiinv_shoup) {
        for (uint64_t i = 0; i < n; i++) {
            psi_shoup[i] = or_shoup(psi_tab[i], q);
            psiinv_shoup[i] = or_shoup(psiinv_tab[i], q);
        }
    }
}

/* Canonical reduction of the 128-bit value hi:lo.  The reference's
 * singleBarrett_new / barrett_reduce_uint128_uint64 (include/uint128.cuh:126,194)
 * compute exactly this residue for every 128-bit input (quotient estimate is
 * floor(a*mu/2^128) >= floor(a/q)-1, then one conditional subtract). */
uint64_t or_barrett_single(uint64_t lo, uint64_t hi, uint64_t q) {
    u128 a = ((u128)hi << 64) | lo;
    return (uint64_t)(a % q);
}
