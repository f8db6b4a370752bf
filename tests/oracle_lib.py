"""ctypes bindings for the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.  Arrays are numpy uint64 (residues) in the layouts documented in
oracle/oracle.h.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


def _load(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    lib = C.CDLL(path)
    sig = {
        "or_mulmod": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "or_powmod": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "or_invmod": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_shoup": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_is_prime": (C.c_int, [C.c_uint64]),
        "or_primitive_root": (C.c_uint64, [C.c_uint64]),
        "or_mth_root": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_primes_below": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int, u64p]),
        "or_primes_descending": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int, u64p]),
        "or_bitrev": (C.c_uint64, [C.c_uint64, C.c_int]),
        "or_fill_tables": (None, [C.c_uint64, C.c_uint64, C.c_int, u64p, u64p, u64p, u64p]),
        "or_barrett_single": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
        "or_ntt_fwd": (None, [u64p, C.c_int, C.c_uint64, u64p]),
        "or_ntt_inv": (None, [u64p, C.c_int, C.c_uint64, u64p]),
        "or_negacyclic_mul_schoolbook": (None, [u64p, u64p, u64p, C.c_int, C.c_uint64]),
        "or_ctx_create": (C.c_void_p, [C.c_int, u64p, C.c_int, u64p, C.c_int, u64p, C.c_int,
                                       C.c_int, C.c_uint64]),
        "or_ctx_destroy": (None, [C.c_void_p]),
        "or_ctx_level": (C.c_void_p, [C.c_void_p, C.c_int]),
        "or_mod_switch": (None, [C.c_void_p, u64p, C.c_int, u64p, C.c_int]),
        "or_ctx_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
        "or_ctx_mod_count": (C.c_int, [C.c_void_p]),
        "or_ctx_moduli": (None, [C.c_void_p, u64p]),
        "or_ctx_psi": (None, [C.c_void_p, u64p]),
        "or_ctx_table": (u64p, [C.c_void_p, C.c_int, C.c_int]),
        "or_ctx_export": (C.c_int, [C.c_void_p, C.c_char_p, u64p, C.c_int]),
        "or_ctx_export_f64": (C.c_int, [C.c_void_p, C.c_char_p, f64p, C.c_int]),
        "or_ntt_fwd_range": (None, [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int]),
        "or_ntt_inv_range": (None, [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int]),
        "or_bconv_q_to_r": (None, [C.c_void_p, u64p, u64p]),
        "or_bconv_r_to_q": (None, [C.c_void_p, u64p, u64p]),
        "or_tensor": (None, [C.c_void_p, u64p, u64p, u64p, C.c_int, C.c_int]),
        "or_scale_round": (None, [C.c_void_p, u64p, u64p]),
        "or_hps_decrypt": (None, [C.c_void_p, u64p, u64p]),
        "or_add_delta_m": (None, [C.c_void_p, u64p, u64p]),
        "or_modup_qj_to_t": (None, [C.c_void_p, u64p, u64p]),
        "or_modup_pq_to_t": (None, [C.c_void_p, u64p, u64p]),
        "or_inner_product": (None, [C.c_void_p, u64p, u64p, u64p]),
        "or_modup_t_to_pq": (None, [C.c_void_p, u64p, u64p]),
        "or_moddown": (None, [C.c_void_p, u64p, u64p]),
        "or_xsalsa20_stream": (None, [u8p, u8p, C.c_uint64, u8p, C.c_size_t]),
        "or_sample_uniform": (None, [C.c_void_p, u8p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, u64p]),
        "or_sample_ternary": (None, [C.c_void_p, u8p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, u64p]),
        "or_sample_cbd": (None, [C.c_void_p, u8p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, u64p]),
        "or_sample_zo": (None, [C.c_void_p, u8p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, u64p]),
        "or_keygen_sk": (None, [C.c_void_p, u8p, u64p]),
        "or_keygen_pk": (None, [C.c_void_p, u8p, u64p, u64p]),
        "or_keygen_rlk": (None, [C.c_void_p, u8p, u64p, u64p]),
        "or_keygen_swk": (None, [C.c_void_p, u8p, u64p, u64p, u64p]),
        "or_encrypt": (None, [C.c_void_p, u8p, C.c_uint64, u64p, u64p, u64p]),
        "or_decrypt": (None, [C.c_void_p, u64p, u64p, u64p]),
        "or_mult_relin": (None, [C.c_void_p, u64p, u64p, u64p, u64p]),
        "or_mult_tensor": (None, [C.c_void_p, u64p, u64p, u64p]),
        "or_relin": (None, [C.c_void_p, u64p, u64p, u64p]),
        "or_encode": (None, [C.c_void_p, u64p, u64p]),
        "or_decode": (None, [C.c_void_p, u64p, u64p]),
        "or_rgsw_rows": (C.c_int, [C.c_void_p, C.c_int]),
        "or_rgsw_encrypt": (None, [C.c_void_p, u8p, u64p, u64p, C.c_int, u64p]),
        "or_external_product": (None, [C.c_void_p, u64p, u64p, C.c_int, u64p]),
        "or_galois_elt": (C.c_uint64, [C.c_void_p, C.c_long]),
        "or_negate": (None, [C.c_void_p, u64p, u64p]),
        "or_mul_scalar": (None, [C.c_void_p, u64p, C.c_uint64, u64p, u64p]),
        "or_add_plain": (None, [C.c_void_p, u64p, u64p, u64p]),
        "or_seal_index_map": (None, [C.c_void_p, u64p]),
        "or_encode_seal": (None, [C.c_void_p, u64p, u64p]),
        "or_decode_seal": (None, [C.c_void_p, u64p, u64p]),
        "or_automorphism": (None, [C.c_void_p, C.c_uint64, u64p, u64p, C.c_int, C.c_int, C.c_int]),
        "or_keygen_galois": (None, [C.c_void_p, u8p, u64p, C.c_uint64, u64p]),
        "or_rotate": (None, [C.c_void_p, u64p, C.c_uint64, u64p, u64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def load_variant(path):
    """Bindings for another build of the oracle sources (tests that plant a
    mistake in the oracle compile one with -DOR_MUTATION=<n>)."""
    return _load(path)


def P(a):
    """numpy uint64 array -> pointer (array must stay alive)."""
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u64p)


def B(b):
    a = np.frombuffer(bytes(b), dtype=np.uint8).copy() if not isinstance(b, np.ndarray) else b
    return a, a.ctypes.data_as(u8p)


def u64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


class Ctx:
    """Oracle context.  Moduli order [P | Q | R | T | t]."""

    def __init__(self, log_n, q, p, t, gamma, plain, _lib=None, _handle=None):
        """_lib: another build of the oracle (load_variant); the default build otherwise.
        _handle: an oracle context made elsewhere (Ctx.level), owned by this object."""
        self._lib = _lib if _lib is not None else lib
        self.q_in, self.p_in, self.t_in = u64(q), u64(p), u64(t)
        if _handle is not None:
            self.h = _handle
        else:
            self.h = self._lib.or_ctx_create(log_n, P(self.q_in), len(q), P(self.p_in), len(p),
                                             P(self.t_in), len(t), gamma, plain)
        if not self.h:
            raise ValueError("oracle rejected the parameter set")
        info = (C.c_int * 16)()
        self._lib.or_ctx_info(self.h, info)
        (self.log_n, self.n, self.qn, self.pn, self.tn, self.rn, self.gamma, self.dnum,
         self.ri_blocks, self.batching, self.nmod, self.oP, self.oQ, self.oR, self.oT,
         self.oPlain) = list(info)
        self.mod = np.zeros(self.nmod, dtype=np.uint64)
        self._lib.or_ctx_moduli(self.h, P(self.mod))
        self.psi = np.zeros(self.nmod, dtype=np.uint64)
        self._lib.or_ctx_psi(self.h, P(self.psi))
        self.plain = int(plain)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.or_ctx_destroy(self.h)
            self.h = None

    def level(self, level):
        """The oracle context of ciphertexts at `level` (or_ctx_level)."""
        h = self._lib.or_ctx_level(self.h, level)
        if not h:
            raise ValueError("bad level")
        return Ctx(self.log_n, self.q_in[:level + 1], self.p_in, self.t_in, self.gamma, self.plain,
                   _lib=self._lib, _handle=h)

    def mod_switch(self, ct, from_level, to_level):
        ct = u64(ct)
        out = np.zeros((2, to_level + 1, self.n), dtype=np.uint64)
        self._lib.or_mod_switch(self.h, P(ct), from_level, P(out), to_level)
        return out

    def table(self, m, which):
        ptr = self._lib.or_ctx_table(self.h, m, which)
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    def export(self, name):
        out = np.zeros(1 << 16, dtype=np.uint64)
        n = self._lib.or_ctx_export(self.h, name.encode(), P(out), len(out))
        assert n > 0, name
        return out[:n].copy()

    def export_f64(self, name):
        out = np.zeros(4096, dtype=np.float64)
        n = self._lib.or_ctx_export_f64(self.h, name.encode(), out.ctypes.data_as(f64p), len(out))
        assert n > 0, name
        return out[:n].copy()

    # ---- transforms ----
    def ntt_fwd(self, a, mod0, limbs, polys=1):
        a = u64(a).copy()
        self._lib.or_ntt_fwd_range(self.h, P(a), mod0, limbs, polys)
        return a

    def ntt_inv(self, a, mod0, limbs, polys=1):
        a = u64(a).copy()
        self._lib.or_ntt_inv_range(self.h, P(a), mod0, limbs, polys)
        return a

    def _call(self, fn, inputs, out_shape):
        ins = [u64(x) for x in inputs]
        out = np.zeros(out_shape, dtype=np.uint64)
        fn(self.h, *[P(x) for x in ins], P(out))
        return out

    def bconv_q_to_r(self, x):
        return self._call(self._lib.or_bconv_q_to_r, [x], (self.rn, self.n))

    def bconv_r_to_q(self, x):
        return self._call(self._lib.or_bconv_r_to_q, [x], (self.qn, self.n))

    def tensor(self, c1, c2, limbs, mod0):
        a, b = u64(c1), u64(c2)
        out = np.zeros((3, limbs, self.n), dtype=np.uint64)
        self._lib.or_tensor(self.h, P(a), P(b), P(out), limbs, mod0)
        return out

    def scale_round(self, x):
        return self._call(self._lib.or_scale_round, [x], (self.rn, self.n))

    def hps_decrypt(self, x):
        return self._call(self._lib.or_hps_decrypt, [x], (self.n,))

    def delta_m(self, m):
        return self._call(self._lib.or_add_delta_m, [m], (self.qn, self.n))

    def modup_qj_to_t(self, x):
        return self._call(self._lib.or_modup_qj_to_t, [x], (self.dnum, self.tn, self.n))

    def modup_pq_to_t(self, x):
        return self._call(self._lib.or_modup_pq_to_t, [x], (self.ri_blocks, self.tn, self.n))

    def inner_product(self, cin, key):
        return self._call(self._lib.or_inner_product, [cin, key], (2, self.ri_blocks, self.tn, self.n))

    def modup_t_to_pq(self, x):
        return self._call(self._lib.or_modup_t_to_pq, [x], (2, self.pn + self.qn, self.n))

    def moddown(self, x):
        return self._call(self._lib.or_moddown, [x], (2, self.qn, self.n))

    # ---- samplers ----
    def sample(self, kind, seed, tag, mod0, limbs, entry=0):
        s, sp = B(seed)
        out = np.zeros((limbs, self.n), dtype=np.uint64)
        getattr(lib, "or_sample_" + kind)(self.h, sp, tag, entry, mod0, limbs, P(out))
        return out

    # ---- scheme ----
    def keygen_sk(self, seed):
        s, sp = B(seed)
        sk = np.zeros((self.pn + self.qn, self.n), dtype=np.uint64)
        self._lib.or_keygen_sk(self.h, sp, P(sk))
        return sk

    def keygen_pk(self, seed, sk):
        s, sp = B(seed)
        sk = u64(sk)
        pk = np.zeros((2, self.qn, self.n), dtype=np.uint64)
        self._lib.or_keygen_pk(self.h, sp, P(sk), P(pk))
        return pk

    def keygen_rlk(self, seed, sk):
        s, sp = B(seed)
        sk = u64(sk)
        rlk = np.zeros((2, self.dnum, self.ri_blocks, self.tn, self.n), dtype=np.uint64)
        self._lib.or_keygen_rlk(self.h, sp, P(sk), P(rlk))
        return rlk

    def keygen_swk(self, seed, sk, src_q):
        s, sp = B(seed)
        sk, src = u64(sk), u64(src_q)
        key = np.zeros((2, self.dnum, self.ri_blocks, self.tn, self.n), dtype=np.uint64)
        self._lib.or_keygen_swk(self.h, sp, P(sk), P(src), P(key))
        return key

    def encrypt(self, seed, pk, m, entry=0):
        s, sp = B(seed)
        pk, m = u64(pk), u64(m)
        ct = np.zeros((2, self.qn, self.n), dtype=np.uint64)
        self._lib.or_encrypt(self.h, sp, entry, P(pk), P(m), P(ct))
        return ct

    def decrypt(self, sk, ct):
        return self._call(self._lib.or_decrypt, [sk, ct], (self.n,))

    def mult_relin(self, rlk, c1, c2):
        return self._call(self._lib.or_mult_relin, [rlk, c1, c2], (2, self.qn, self.n))

    def mult_tensor(self, c1, c2):
        return self._call(self._lib.or_mult_tensor, [c1, c2], (3, self.qn, self.n))

    def relin(self, rlk, d2):
        return self._call(self._lib.or_relin, [rlk, d2], (2, self.qn, self.n))

    def encode(self, slots):
        return self._call(self._lib.or_encode, [slots], (self.n,))

    def decode(self, m):
        return self._call(self._lib.or_decode, [m], (self.n,))

    def rgsw_rows(self, base_bits):
        return self._lib.or_rgsw_rows(self.h, base_bits)

    def rgsw_encrypt(self, seed, sk, m_coeff, base_bits):
        s, sp = B(seed)
        sk, m = u64(sk), u64(m_coeff)
        rows = self.rgsw_rows(base_bits)
        out = np.zeros((rows, 2, self.qn, self.n), dtype=np.uint64)
        self._lib.or_rgsw_encrypt(self.h, sp, P(sk), P(m), base_bits, P(out))
        return out

    def external_product(self, rgsw, ct, base_bits):
        rg, c = u64(rgsw), u64(ct)
        out = np.zeros((2, self.qn, self.n), dtype=np.uint64)
        self._lib.or_external_product(self.h, P(rg), P(c), base_bits, P(out))
        return out

    def negate(self, ct):
        return self._call(self._lib.or_negate, [ct], (2, self.qn, self.n))

    def mul_scalar(self, ct, k, add=None):
        c = u64(ct)
        out = np.zeros((2, self.qn, self.n), dtype=np.uint64)
        a = u64(add) if add is not None else None
        self._lib.or_mul_scalar(self.h, P(c), k, P(a) if a is not None else None, P(out))
        return out

    def add_plain(self, ct, m):
        return self._call(self._lib.or_add_plain, [ct, m], (2, self.qn, self.n))

    def seal_index_map(self):
        out = np.zeros(self.n, dtype=np.uint64)
        self._lib.or_seal_index_map(self.h, P(out))
        return out

    def encode_seal(self, slots):
        return self._call(self._lib.or_encode_seal, [slots], (self.n,))

    def decode_seal(self, m):
        return self._call(self._lib.or_decode_seal, [m], (self.n,))

    # ---- Galois automorphisms / rotation ----
    def galois_elt(self, steps):
        return int(self._lib.or_galois_elt(self.h, steps))

    def automorphism(self, g, a, mod0, limbs, polys=1):
        a = u64(a)
        out = np.zeros_like(a)
        self._lib.or_automorphism(self.h, g, P(a), P(out), mod0, limbs, polys)
        return out

    def keygen_galois(self, seed, sk, g):
        s, sp = B(seed)
        sk = u64(sk)
        key = np.zeros((2, self.dnum, self.ri_blocks, self.tn, self.n), dtype=np.uint64)
        self._lib.or_keygen_galois(self.h, sp, P(sk), g, P(key))
        return key

    def rotate(self, key, g, ct):
        k, c = u64(key), u64(ct)
        out = np.zeros((2, self.qn, self.n), dtype=np.uint64)
        self._lib.or_rotate(self.h, P(k), g, P(c), P(out))
        return out


def xsalsa20(key, nonce, length, block0=0):
    k, kp = B(key)
    n, np_ = B(nonce)
    out = np.zeros(length, dtype=np.uint8)
    lib.or_xsalsa20_stream(kp, np_, block0, out.ctypes.data_as(u8p), length)
    return bytes(out)


def primes_descending(below, step, count):
    out = np.zeros(count, dtype=np.uint64)
    assert lib.or_primes_descending(below, step, count, P(out)) == 0
    return [int(x) for x in out]


def chain_60bit(qn, pn, tn):
    """The 60-bit chain rule: T first, then P, then Q, all primes 1 mod 2^17 below 2^60."""
    allp = primes_descending(1 << 60, 1 << 17, tn + pn + qn)
    return allp[tn + pn:], allp[tn:tn + pn], allp[:tn]
