"""Parameter sets used by the tests (chains built with the 60-bit rule or the
reference's own sets)."""
import oracle_lib as orc

# the reference's logN=14 set (include/Params.cuh:31-46): K = 2 special primes
REF14 = dict(log_n=14, q=[0x7FFE28001, 0x7FFF18001, 0x7FFF80001, 0x7FFFB0001],
             p=[0xFFFFCA8001, 0xFFFFE80001],
             t=[0xffffffffffc0001, 0xfffffffff840001, 0xfffffffff6a0001, 0xfffffffff5a0001],
             gamma=2, plain=23068673)
# the reference's logN=15 chain (include/Params.cuh:120-145); its gamma = 4 is
# too large for its 4 T primes, gamma = 2 is the engine's default
REF15_Q = [36028797001138177, 36028797003563009, 36028797003694081, 36028797005135873,
           36028797005529089, 36028797005856769, 36028797009985537, 36028797010444289,
           36028797012606977, 36028797013000193, 36028797013327873, 36028797014376449,
           36028797014573057, 36028797014704129, 36028797017456641]
REF15 = dict(log_n=15, q=REF15_Q, p=[72057594037338113], t=REF14["t"], gamma=2, plain=0xC0001)


def chain60(log_n, qn, pn=1, tn=4, gamma=2, plain=65537):
    q, p, t = orc.chain_60bit(qn, pn, tn)
    return dict(log_n=log_n, q=q, p=p, t=t, gamma=gamma, plain=plain)


def c1_params():
    """N=4096, two 36-bit q and one 37-bit special prime, 1 mod 8192, t=65537."""
    q = orc.primes_descending(1 << 36, 8192, 2)
    p = orc.primes_descending(1 << 37, 8192, 1)
    t = orc.primes_descending(1 << 60, 1 << 17, 2)
    return dict(log_n=12, q=q, p=p, t=t, gamma=1, plain=65537)
