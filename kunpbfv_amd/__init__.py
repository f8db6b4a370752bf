"""kunpbfv_amd -- MI355X-native BFV engine, Python host binding.

Loads the in-tree ``libkunpbfv.so`` (HIP for gfx950) through its C ABI
(``include/kunpbfv.h``) and mirrors the reference's BFVContext / BFVScheme
operations on device buffers.  Device memory and streams come from PyTorch
(``torch.int64`` tensors carry the uint64 residues bit for bit); every compute
call goes to the HIP library -- there is no CPU fallback, and a missing or
unloadable library raises immediately.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkunpbfv.so")

_vp, _i64, _u64, _sz = C.c_void_p, C.c_int64, C.c_uint64, C.c_size_t
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes); the complete exported C ABI
SIGNATURES = {
    "kbfv_last_error": (C.c_char_p, []),
    "kbfv_chain_60bit": (C.c_int, [C.c_int, C.c_int, C.c_int, _u64p, _u64p, _u64p]),
    "kbfv_primes_below": (C.c_int, [_u64, _u64, C.c_int, _u64p]),
    "kbfv_random_seed": (C.c_int, [_vp]),
    "kbfv_default_params": (C.c_int, [C.c_int, _u64p, _ip, _u64p, _ip, _u64p, _ip, _ip, _u64p]),
    "kbfv_context_create": (C.c_int, [C.c_int, _u64p, C.c_int, _u64p, C.c_int, _u64p, C.c_int, C.c_int,
                                      _u64, C.POINTER(_vp)]),
    "kbfv_context_destroy": (C.c_int, [_vp]),
    "kbfv_context_info": (C.c_int, [_vp, _ip]),
    "kbfv_context_moduli": (C.c_int, [_vp, _u64p]),
    "kbfv_context_export": (C.c_int, [_vp, C.c_char_p, _u64p, C.c_int]),
    "kbfv_context_export_f64": (C.c_int, [_vp, C.c_char_p, _f64p, C.c_int]),
    "kbfv_ntt": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _i64, _i64, _i64,
                           _i64, _i64, _i64, _vp]),
    "kbfv_ntt_plan_valid": (C.c_int, [C.c_int, C.c_uint32]),
    "kbfv_ntt_plan_candidates": (C.c_int, [C.c_int, C.POINTER(C.c_uint32), C.c_int]),
    "kbfv_ntt_plan_bucket": (C.c_int, [_i64]),
    "kbfv_ntt_plan_get": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
    "kbfv_ntt_plan_set": (C.c_int, [_vp, C.c_int, C.c_int, C.c_uint32]),
    "kbfv_ntt_autotune": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_float)]),
    "kbfv_ntt_desc_upload": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _ip, _vp]),
    "kbfv_ntt_desc": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp]),
    "kbfv_poly_binop_desc_upload": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _ip, _vp]),
    "kbfv_poly_binop_desc": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp]),
    "kbfv_poly_add": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "kbfv_poly_sub": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "kbfv_poly_mul": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "kbfv_bconv_q_to_r": (C.c_int, [_vp, _vp, _i64, _vp, _i64, C.c_int, _vp]),
    "kbfv_bconv_r_to_q": (C.c_int, [_vp, _vp, _i64, _vp, _i64, C.c_int, _vp]),
    "kbfv_tensor": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "kbfv_scale_round": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_hps_decrypt": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_modup_qj_to_t": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_modup_pq_to_t": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_inner_product": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_modup_t_to_pq": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_moddown": (C.c_int, [_vp, _vp, _vp, _vp, _i64, C.c_int, _vp]),
    "kbfv_xsalsa20": (C.c_int, [_vp, _vp, _u64, _vp, _sz, _vp]),
    "kbfv_sample": (C.c_int, [_vp, C.c_int, _vp, _u64, C.c_int, C.c_int, _vp, _vp]),
    "kbfv_keygen_sk": (C.c_int, [_vp, _vp, _vp, _vp]),
    "kbfv_keygen_pk": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "kbfv_keygen_rlk_workspace_bytes": (_sz, [_vp]),
    "kbfv_keygen_rlk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "kbfv_keygen_swk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "kbfv_mult_tensor": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_encrypt_workspace_bytes": (_sz, [_vp, C.c_int]),
    "kbfv_encrypt": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_decrypt_workspace_bytes": (_sz, [_vp, C.c_int]),
    "kbfv_decrypt": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_add": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_mult_workspace_bytes": (_sz, [_vp, C.c_int]),
    "kbfv_mult_relin": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_noalloc_max_batch": (_sz, [_vp]),
    "kbfv_context_level": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "kbfv_mod_switch_workspace_bytes": (_sz, [_vp, C.c_int, C.c_int]),
    "kbfv_mod_switch": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_relin_ragged": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _ip, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_relin_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_negate": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_mul_scalar": (C.c_int, [_vp, _vp, _u64, _vp, _vp, C.c_int, _vp]),
    "kbfv_add_plain_workspace_bytes": (_sz, [_vp, C.c_int]),
    "kbfv_add_plain": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_encode_seal": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_decode_seal": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "kbfv_galois_elt": (_u64, [_vp, _i64]),
    "kbfv_automorphism": (C.c_int, [_vp, _u64, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "kbfv_keygen_galois_workspace_bytes": (_sz, [_vp]),
    "kbfv_keygen_galois": (C.c_int, [_vp, _vp, _vp, _u64, _vp, _vp, _sz, _vp]),
    "kbfv_rotate_workspace_bytes": (_sz, [_vp, C.c_int]),
    "kbfv_rotate": (C.c_int, [_vp, _vp, _u64, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_relin_graph_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _sz, C.POINTER(_vp)]),
    "kbfv_graph_launch": (C.c_int, [_vp, _vp]),
    "kbfv_graph_destroy": (None, [_vp]),
    "kbfv_encode": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "kbfv_decode": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "kbfv_rgsw_rows": (C.c_int, [_vp, C.c_int]),
    "kbfv_rgsw_encrypt_workspace_bytes": (_sz, [_vp]),
    "kbfv_rgsw_encrypt": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _sz, _vp]),
    "kbfv_external_product_workspace_bytes": (_sz, [_vp, C.c_int, C.c_int]),
    "kbfv_external_product": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_before_ip": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_mult_after_ip": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, _sz, _vp]),
    "kbfv_wire_bytes": (_sz, [_vp, C.c_int, C.c_int, _u64]),
    "kbfv_wire_save": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _u64, _vp, _sz, _vp]),
    "kbfv_wire_load": (C.c_int, [_vp, _vp, _sz, C.c_int, _vp, C.c_int, _vp]),
    "kbfv_wire_peek": (C.c_int, [_vp, _sz, _ip, _ip, _u64p]),
    "kbfv_wire_checksum": (_u64, [_vp, _sz]),
}

# wire-format kinds (include/kunpbfv.h KBFV_WIRE_*)
WIRE_CIPHERTEXT, WIRE_PLAINTEXT, WIRE_SECRET_KEY, WIRE_PUBLIC_KEY = 1, 2, 3, 4
WIRE_RELIN_KEY, WIRE_GALOIS_KEY, WIRE_RGSW = 5, 6, 7


class KbfvError(RuntimeError):
    pass


def _load(path=LIB_PATH):
    if not os.path.exists(path):
        raise KbfvError(f"HIP library not built: {path} (run __graft_entry__.build())")
    # torch ships its own libamdhip64.so with the soname libamdhip64.so.7 that
    # the library links against.  Loading torch first makes the library bind to
    # that one HIP runtime; loaded the other way round, the process holds two
    # runtimes, torch's finds no device, and torch's streams are foreign
    # handles to ours.
    import torch  # noqa: F401
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc != 0:
        raise KbfvError(f"kunpbfv error {rc}: {lib.kbfv_last_error().decode()}")
    return rc


def _arr(vals):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.uint64))
    return a, a.ctypes.data_as(_u64p)


def ntt_plan_candidates(log_n):
    """Every valid NTT plan word at N = 2^log_n (host only)."""
    n = lib.kbfv_ntt_plan_candidates(log_n, None, 0)
    check(min(n, 0))
    out = (C.c_uint32 * n)()
    lib.kbfv_ntt_plan_candidates(log_n, out, n)
    return list(out)


def ntt_plan_fields(plan, log_n):
    """Decode a plan word: (column-pass log length, column tile log, row tile log,
    column XCD order, row XCD order); 0 is the built-in plan.  The chunk field
    is ntt_plan_chunk_bytes."""
    if plan == 0:
        return ((log_n + 1) // 2, 12, 12, 1, 1)
    return (plan & 15, (plan >> 4) & 15, (plan >> 8) & 15, (plan >> 12) & 1, (plan >> 13) & 1)


def ntt_plan_chunk_bytes(plan):
    """Bytes per chunk of a chunked plan (both passes per chunk), 0 if unchunked."""
    k = (plan >> 14) & 7
    return 1 << (k + 23) if k else 0


def chain_60bit(qn, pn, tn):
    q, p, t = (np.zeros(n, np.uint64) for n in (qn, pn, tn))
    check(lib.kbfv_chain_60bit(qn, pn, tn, q.ctypes.data_as(_u64p), p.ctypes.data_as(_u64p),
                               t.ctypes.data_as(_u64p)))
    return [int(x) for x in q], [int(x) for x in p], [int(x) for x in t]


def primes_below(ntt_size, upper, count):
    out = np.zeros(count, np.uint64)
    check(lib.kbfv_primes_below(ntt_size, upper, count, out.ctypes.data_as(_u64p)))
    return [int(x) for x in out]


def default_params(log_n):
    q, p, t = (np.zeros(64, np.uint64) for _ in range(3))
    qn, pn, tn, gamma = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    plain = C.c_uint64()
    check(lib.kbfv_default_params(log_n, q.ctypes.data_as(_u64p), C.byref(qn), p.ctypes.data_as(_u64p),
                                  C.byref(pn), t.ctypes.data_as(_u64p), C.byref(tn), C.byref(gamma),
                                  C.byref(plain)))
    return dict(log_n=log_n, q=[int(x) for x in q[:qn.value]], p=[int(x) for x in p[:pn.value]],
                t=[int(x) for x in t[:tn.value]], gamma=gamma.value, plain=plain.value)


def _torch():
    import torch  # device memory and streams only
    return torch


def ptr(t):
    """device pointer of a contiguous torch tensor"""
    assert t.is_contiguous(), "kunpbfv needs contiguous device buffers"
    return C.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def to_device(a, device="cuda"):
    """numpy uint64 -> torch.int64 device tensor (same bits)"""
    torch = _torch()
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return torch.from_numpy(a.view(np.int64)).to(device)


def wire_peek(blob):
    """(kind, count, aux) from a wire blob's header; no context needed"""
    kind, count, aux = C.c_int(), C.c_int(), C.c_uint64()
    check(lib.kbfv_wire_peek(blob, len(blob), C.byref(kind), C.byref(count), C.byref(aux)))
    return kind.value, count.value, aux.value


def wire_checksum(words):
    """the format's payload checksum of a uint64 numpy array"""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    return int(lib.kbfv_wire_checksum(w.ctypes.data, w.size))


def to_host(t):
    """torch.int64 device tensor -> numpy uint64 (same bits)"""
    return t.detach().cpu().numpy().view(np.uint64)


def random_seed():
    """32 fresh bytes from the OS entropy source (kbfv_random_seed)."""
    buf = C.create_string_buffer(32)
    check(lib.kbfv_random_seed(buf))
    return buf.raw


def _seed(seed):
    b = bytes(seed)
    assert len(b) == 32, "seeds are 32 bytes"
    return C.create_string_buffer(b, 32)


class Context:
    """A BFV parameter set with its device tables (replaces BFVContext)."""

    INFO = ("log_n", "n", "qn", "pn", "tn", "rn", "gamma", "dnum", "ri_blocks", "batching", "nmod",
            "oP", "oQ", "oR", "oT", "oPlain")

    def __init__(self, log_n, q, p, t, gamma, plain):
        self._q, qp = _arr(q)
        self._p, pp = _arr(p)
        self._t, tp = _arr(t)
        h = C.c_void_p()
        check(lib.kbfv_context_create(log_n, qp, len(q), pp, len(p), tp, len(t), gamma, plain,
                                      C.byref(h)))
        self.h = h
        info = (C.c_int * 16)()
        check(lib.kbfv_context_info(self.h, info))
        for k, v in zip(self.INFO, info):
            setattr(self, k, v)
        self.plain = int(plain)
        mods = np.zeros(self.nmod, np.uint64)
        lib.kbfv_context_moduli(self.h, mods.ctypes.data_as(_u64p))
        self.mod = mods
        self._ws = None

    def level(self, level):
        """The context at ciphertext level `level` (Q_l = the first level + 1
        Q primes): a view owned by this context; the full level is `self`."""
        if level == self.qn - 1 and getattr(self, "parent", None) is None:
            return self
        return LevelView(self, level)

    def mod_switch(self, ct, from_level, to_level, out=None, stream=None):
        """Level from_level -> to_level (compact [batch][2][l + 1][N] batches)."""
        batch = ct.shape[0]
        out = out if out is not None else self.empty(batch, 2, to_level + 1, self.n)
        ws = self.workspace(lib.kbfv_mod_switch_workspace_bytes(self.h, from_level, batch))
        check(lib.kbfv_mod_switch(self.h, ptr(ct), from_level, ptr(out), to_level, batch, ptr(ws), ws.numel(),
                                  stream_ptr(stream)))
        return out

    def mult_relin_ragged(self, rlk, c1, c2, levels, out=None, ws=None, stream=None):
        """HomMult + relinearisation of pairs at the given levels, packed back to
        back in the flat buffers c1, c2 (entry b: 2 (levels[b] + 1) N words)."""
        need = sum(2 * (int(lvl) + 1) * self.n for lvl in levels)
        assert c1.numel() == need and c2.numel() == need, "packed batch size does not match the levels"
        lv = (C.c_int * len(levels))(*levels)
        out = out if out is not None else self.empty(*c1.shape)
        assert out.numel() == need
        ws = ws if ws is not None else self.workspace(self.mult_workspace_bytes(1))
        check(lib.kbfv_mult_relin_ragged(self.h, ptr(rlk), ptr(c1), ptr(c2), ptr(out), lv, len(levels), ptr(ws),
                                         ws.numel(), stream_ptr(stream)))
        return out

    @classmethod
    def default(cls, log_n):
        d = default_params(log_n)
        return cls(d["log_n"], d["q"], d["p"], d["t"], d["gamma"], d["plain"])

    @classmethod
    def chain60(cls, log_n, qn, pn=1, tn=4, gamma=2, plain=65537):
        q, p, t = chain_60bit(qn, pn, tn)
        return cls(log_n, q, p, t, gamma, plain)

    def close(self):
        if getattr(self, "h", None):
            lib.kbfv_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ tables
    def export(self, name):
        out = np.zeros(1 << 16, np.uint64)
        n = lib.kbfv_context_export(self.h, name.encode(), out.ctypes.data_as(_u64p), len(out))
        check(0 if n > 0 else n)
        return out[:n].copy()

    def export_f64(self, name):
        out = np.zeros(4096, np.float64)
        n = lib.kbfv_context_export_f64(self.h, name.encode(), out.ctypes.data_as(_f64p), len(out))
        check(0 if n > 0 else n)
        return out[:n].copy()

    # ------------------------------------------------------------ buffers
    def zeros(self, *shape):
        torch = _torch()
        return torch.zeros(shape, dtype=torch.int64, device="cuda")

    def empty(self, *shape):
        torch = _torch()
        return torch.empty(shape, dtype=torch.int64, device="cuda")

    def workspace(self, nbytes):
        """grow-only scratch buffer owned by this context object"""
        torch = _torch()
        nbytes = max(int(nbytes), 256)
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        return self._ws

    # ------------------------------------------------------------ transforms
    def ntt(self, src, dst, direction, mod0, limbs, polys=1, batch=1, s=None, d=None, stream=None):
        N = self.n
        s = s or (polys * limbs * N, limbs * N, N)
        d = d or s
        check(lib.kbfv_ntt(self.h, direction, ptr(src), ptr(dst), mod0, limbs, polys, batch, *s, *d,
                           stream_ptr(stream)))
        return dst

    # ------------------------------------------------------------ NTT plans / autotuner
    def ntt_plan(self, direction, bucket):
        v = C.c_uint32()
        check(lib.kbfv_ntt_plan_get(self.h, direction, bucket, C.byref(v)))
        return v.value

    def set_ntt_plan(self, direction, bucket, plan):
        check(lib.kbfv_ntt_plan_set(self.h, direction, bucket, plan))

    def autotune_ntt(self, direction, mod0, limbs, polys, reps=3, scratch=None, stream=None):
        """Time every NTT plan for `polys` polynomials of `limbs` limbs from
        modulus mod0 and keep the fastest for that transform-count bucket.
        Returns (plan, ms per transform, bucket)."""
        torch = _torch()
        words = polys * limbs * self.n
        if scratch is None:
            scratch = torch.empty(words, dtype=torch.int64, device="cuda")
        assert scratch.numel() * scratch.element_size() >= words * 8, "scratch too small"
        best, ms = C.c_uint32(), C.c_float()
        check(lib.kbfv_ntt_autotune(self.h, direction, mod0, limbs, polys, ptr(scratch), reps,
                                    stream_ptr(stream), C.byref(best), C.byref(ms)))
        return best.value, ms.value, lib.kbfv_ntt_plan_bucket(limbs * polys)

    def autotune(self, batch, scratch=None, reps=3, stream=None, cache=None):
        """Tune the NTT plan of every direction and transform-count bucket at
        the sizes HomMult+relin of `batch` ciphertexts launches: over the QR
        basis, 1 polynomial (< 64 transforms), 3 (the per-ciphertext tensor),
        a mid-size batch and the whole batch's 3 tensor polynomials.  Returns
        {(direction, bucket): (plan, ms per transform, transforms)}.

        cache: path of a JSON file of tuned plans keyed by device name, ring
        degree, modulus table, batch and plan space; a hit sets the stored
        plans without timing anything, a miss tunes and adds the entry."""
        key = self._tune_key(batch) if cache else None
        if cache:
            hit = _cache_read(cache).get(key)
            if hit is not None:
                try:
                    out = {}
                    for d, b, plan, ms, transforms in hit:
                        self.set_ntt_plan(d, b, plan)
                        out[(d, b)] = (plan, ms, transforms)
                    return out
                except (KbfvError, TypeError, ValueError):
                    pass  # a stale or malformed entry: tune again and replace it
        torch = _torch()
        limbs = self.qn + self.rn
        sizes = {}
        for polys in (1, 3, max(1, 2048 // limbs), 3 * max(1, batch)):
            b = lib.kbfv_ntt_plan_bucket(limbs * polys)
            sizes[b] = max(sizes.get(b, 0), polys)
        need = max(sizes.values()) * limbs * self.n
        if scratch is None:
            scratch = torch.empty(need, dtype=torch.int64, device="cuda")
        if scratch.numel() * scratch.element_size() < need * 8:
            raise KbfvError("autotune scratch too small")
        out = {}
        for direction in (+1, -1):
            for b, polys in sorted(sizes.items()):
                plan, ms, got = self.autotune_ntt(direction, self.oQ, limbs, polys, reps, scratch, stream)
                assert got == b
                out[(direction, b)] = (plan, ms, limbs * polys)
        if cache:
            _cache_write(cache, key, [[d, b, p, m, t] for (d, b), (p, m, t) in sorted(out.items())])
        return out

    def _tune_key(self, batch):
        import hashlib
        torch = _torch()
        mods = hashlib.sha1(np.ascontiguousarray(self.mod, dtype=np.uint64).tobytes()).hexdigest()[:16]
        space = hashlib.sha1(np.asarray(ntt_plan_candidates(self.log_n), dtype=np.uint32).tobytes()).hexdigest()[:12]
        dev = torch.cuda.get_device_name(torch.cuda.current_device())
        return f"{dev}|logN={self.log_n}|mods={mods}|batch={batch}|plans={space}|build={library_build_id()}"

    # ------------------------------------------------------------ ragged batches
    def ragged(self, items, stream=None):
        """Upload descriptors of a ragged batch: items are (src, dst, mod0, limbs)
        with src/dst device tensors holding [limbs][N] (or (src, dst, mod0, limbs,
        s_limb, d_limb) for other limb strides, in uint64 units)."""
        return RaggedBatch(self, items, stream)

    def ntt_ragged(self, direction, batch, stream=None):
        """One batched NTT (forward +1 / inverse -1) over every polynomial of a
        RaggedBatch, whatever their RNS lengths; reads only device memory, so
        it may be captured in a graph."""
        check(lib.kbfv_ntt_desc(self.h, direction, ptr(batch.desc), ptr(batch.start), batch.count, batch.total,
                                stream_ptr(stream)))

    def ragged_binop(self, items, stream=None):
        """Upload descriptors of a ragged element-wise batch: items are
        (a, b, out, mod0, limbs) with [limbs][N] device tensors."""
        return RaggedBinop(self, items, stream)

    def binop_ragged(self, op, batch, stream=None):
        """op 0 add, 1 sub, 2 mul over every descriptor of a RaggedBinop."""
        check(lib.kbfv_poly_binop_desc(self.h, op, ptr(batch.desc), ptr(batch.start), batch.count, batch.total,
                                       stream_ptr(stream)))

    def poly_op(self, op, a, b, out, mod0, limbs, polys=1, stream=None):
        fn = {"add": lib.kbfv_poly_add, "sub": lib.kbfv_poly_sub, "mul": lib.kbfv_poly_mul}[op]
        check(fn(self.h, ptr(a), ptr(b), ptr(out), mod0, limbs, polys, stream_ptr(stream)))
        return out

    def bconv(self, x, y, q_to_r=True, polys=1, stream=None):
        fn = lib.kbfv_bconv_q_to_r if q_to_r else lib.kbfv_bconv_r_to_q
        ins, outs = (self.qn, self.rn) if q_to_r else (self.rn, self.qn)
        check(fn(self.h, ptr(x), ins * self.n, ptr(y), outs * self.n, polys, stream_ptr(stream)))
        return y

    def tensor(self, c1, c2, out, mod0, limbs, batch=1, stream=None):
        check(lib.kbfv_tensor(self.h, ptr(c1), ptr(c2), ptr(out), mod0, limbs, batch, stream_ptr(stream)))
        return out

    def scale_round(self, x, y, polys=1, stream=None):
        check(lib.kbfv_scale_round(self.h, ptr(x), ptr(y), polys, stream_ptr(stream)))
        return y

    def hps_decrypt(self, x, m, batch=1, stream=None):
        check(lib.kbfv_hps_decrypt(self.h, ptr(x), ptr(m), batch, stream_ptr(stream)))
        return m

    def modup_qj_to_t(self, x, out, batch=1, stream=None):
        check(lib.kbfv_modup_qj_to_t(self.h, ptr(x), ptr(out), batch, stream_ptr(stream)))
        return out

    def modup_pq_to_t(self, x, out, batch=1, stream=None):
        check(lib.kbfv_modup_pq_to_t(self.h, ptr(x), ptr(out), batch, stream_ptr(stream)))
        return out

    def inner_product(self, cin, key, out, batch=1, stream=None):
        check(lib.kbfv_inner_product(self.h, ptr(cin), ptr(key), ptr(out), batch, stream_ptr(stream)))
        return out

    def modup_t_to_pq(self, x, out, batch=1, stream=None):
        check(lib.kbfv_modup_t_to_pq(self.h, ptr(x), ptr(out), batch, stream_ptr(stream)))
        return out

    def moddown(self, x, out, add=None, add_stride=0, batch=1, stream=None):
        a = ptr(add) if add is not None else None
        check(lib.kbfv_moddown(self.h, ptr(x), ptr(out), a, add_stride, batch, stream_ptr(stream)))
        return out

    def sample(self, kind, seed, tag, mod0, limbs, out, stream=None):
        kinds = {"uniform": 0, "ternary": 1, "zo": 2, "cbd": 3}
        sb = _seed(seed)
        check(lib.kbfv_sample(self.h, kinds[kind], sb, tag, mod0, limbs, ptr(out), stream_ptr(stream)))
        return out

    # ------------------------------------------------------------ scheme
    def keygen_sk(self, seed, stream=None):
        sk = self.empty(self.pn + self.qn, self.n)
        check(lib.kbfv_keygen_sk(self.h, _seed(seed), ptr(sk), stream_ptr(stream)))
        return sk

    def keygen_pk(self, seed, sk, stream=None):
        pk = self.empty(2, self.qn, self.n)
        check(lib.kbfv_keygen_pk(self.h, _seed(seed), ptr(sk), ptr(pk), stream_ptr(stream)))
        return pk

    def keygen_rlk(self, seed, sk, stream=None):
        rlk = self.empty(2, self.dnum, self.ri_blocks, self.tn, self.n)
        nb = lib.kbfv_keygen_rlk_workspace_bytes(self.h)
        ws = self.workspace(nb)
        check(lib.kbfv_keygen_rlk(self.h, _seed(seed), ptr(sk), ptr(rlk), ptr(ws), ws.numel(),
                                  stream_ptr(stream)))
        return rlk

    def keygen_swk(self, seed, sk, src_q, stream=None):
        """key-switching key from src_q ([qn][N], NTT over Q) to the secret"""
        key = self.empty(2, self.dnum, self.ri_blocks, self.tn, self.n)
        ws = self.workspace(lib.kbfv_keygen_rlk_workspace_bytes(self.h))
        check(lib.kbfv_keygen_swk(self.h, _seed(seed), ptr(sk), ptr(src_q), ptr(key), ptr(ws), ws.numel(),
                                  stream_ptr(stream)))
        return key

    def encrypt(self, seed, pk, m, stream=None):
        """m: [batch][N] encoded plaintexts -> [batch][2][qn][N]"""
        batch = m.shape[0]
        ct = self.empty(batch, 2, self.qn, self.n)
        ws = self.workspace(lib.kbfv_encrypt_workspace_bytes(self.h, batch))
        check(lib.kbfv_encrypt(self.h, _seed(seed), ptr(pk), ptr(m), ptr(ct), batch, ptr(ws), ws.numel(),
                               stream_ptr(stream)))
        return ct

    def decrypt(self, sk, ct, stream=None):
        batch = ct.shape[0]
        m = self.empty(batch, self.n)
        ws = self.workspace(lib.kbfv_decrypt_workspace_bytes(self.h, batch))
        check(lib.kbfv_decrypt(self.h, ptr(sk), ptr(ct), ptr(m), batch, ptr(ws), ws.numel(),
                               stream_ptr(stream)))
        return m

    def add(self, c1, c2, out=None, stream=None):
        out = out if out is not None else self.empty(*c1.shape)
        check(lib.kbfv_add(self.h, ptr(c1), ptr(c2), ptr(out), c1.shape[0], stream_ptr(stream)))
        return out

    def mult_workspace_bytes(self, batch):
        return lib.kbfv_mult_workspace_bytes(self.h, batch)

    def mult_relin(self, rlk, c1, c2, out=None, ws=None, stream=None):
        batch = c1.shape[0]
        out = out if out is not None else self.empty(*c1.shape)
        ws = ws if ws is not None else self.workspace(self.mult_workspace_bytes(batch))
        check(lib.kbfv_mult_relin(self.h, ptr(rlk), ptr(c1), ptr(c2), ptr(out), batch, ptr(ws), ws.numel(),
                                  stream_ptr(stream)))
        return out

    def mult_tensor(self, c1, c2, ws=None, stream=None):
        """the scaled tensor (a1a2, a1b2+a2b1, b1b2) of each pair, NTT form: [batch][3][qn][N]"""
        batch = c1.shape[0]
        out = self.empty(batch, 3, self.qn, self.n)
        ws = ws if ws is not None else self.workspace(self.mult_workspace_bytes(batch))
        check(lib.kbfv_mult_tensor(self.h, ptr(c1), ptr(c2), ptr(out), batch, ptr(ws), ws.numel(),
                                   stream_ptr(stream)))
        return out

    def mult_before_ip(self, c1, c2, ws=None, stream=None):
        """split-IP seam, first half: (ip_in [batch][dnum][tn][N], d01 [batch][2][qn][N])"""
        batch = c1.shape[0]
        ip_in = self.empty(batch, self.dnum, self.tn, self.n)
        d01 = self.empty(batch, 2, self.qn, self.n)
        ws = ws if ws is not None else self.workspace(self.mult_workspace_bytes(batch))
        check(lib.kbfv_mult_before_ip(self.h, ptr(c1), ptr(c2), ptr(ip_in), ptr(d01), batch, ptr(ws), ws.numel(),
                                      stream_ptr(stream)))
        return ip_in, d01

    def ip_batch(self, ip_in, key, stream=None):
        """key-stationary inner product over the batch -> [batch][2][ri_blocks][tn][N]"""
        batch = ip_in.shape[0]
        out = self.empty(batch, 2, self.ri_blocks, self.tn, self.n)
        check(lib.kbfv_inner_product(self.h, ptr(ip_in), ptr(key), ptr(out), batch, stream_ptr(stream)))
        return out

    def mult_after_ip(self, ip_out, d01, out=None, ws=None, stream=None):
        """split-IP seam, second half: the relinearised products [batch][2][qn][N]"""
        batch = ip_out.shape[0]
        out = out if out is not None else self.empty(batch, 2, self.qn, self.n)
        ws = ws if ws is not None else self.workspace(self.mult_workspace_bytes(batch))
        check(lib.kbfv_mult_after_ip(self.h, ptr(ip_out), ptr(d01), ptr(out), batch, ptr(ws), ws.numel(),
                                     stream_ptr(stream)))
        return out

    def negate(self, ct, out=None, stream=None):
        out = out if out is not None else self.empty(*ct.shape)
        check(lib.kbfv_negate(self.h, ptr(ct), ptr(out), ct.shape[0], stream_ptr(stream)))
        return out

    def mul_scalar(self, ct, k, add=None, out=None, stream=None):
        out = out if out is not None else self.empty(*ct.shape)
        a = ptr(add) if add is not None else None
        check(lib.kbfv_mul_scalar(self.h, ptr(ct), k, a, ptr(out), ct.shape[0], stream_ptr(stream)))
        return out

    def add_plain(self, ct, m, out=None, stream=None):
        out = out if out is not None else self.empty(*ct.shape)
        ws = self.workspace(lib.kbfv_add_plain_workspace_bytes(self.h, ct.shape[0]))
        check(lib.kbfv_add_plain(self.h, ptr(ct), ptr(m), ptr(out), ct.shape[0], ptr(ws), ws.numel(),
                                 stream_ptr(stream)))
        return out

    def encode_seal(self, vals, stream=None):
        """slots in SEAL's order -> plaintext coefficients, in place"""
        tmp = self.empty(*vals.shape)
        check(lib.kbfv_encode_seal(self.h, ptr(vals), ptr(tmp), vals.shape[0], stream_ptr(stream)))
        return vals

    def decode_seal(self, vals, stream=None):
        tmp = self.empty(*vals.shape)
        check(lib.kbfv_decode_seal(self.h, ptr(vals), ptr(tmp), vals.shape[0], stream_ptr(stream)))
        return vals

    # ---- Galois automorphisms / rotation
    def galois_elt(self, steps):
        return int(lib.kbfv_galois_elt(self.h, steps))

    def automorphism(self, g, x, out, mod0, limbs, polys=1, stream=None):
        check(lib.kbfv_automorphism(self.h, g, ptr(x), ptr(out), mod0, limbs, polys, stream_ptr(stream)))
        return out

    def keygen_galois(self, seed, sk, g, stream=None):
        key = self.empty(2, self.dnum, self.ri_blocks, self.tn, self.n)
        ws = self.workspace(lib.kbfv_keygen_galois_workspace_bytes(self.h))
        check(lib.kbfv_keygen_galois(self.h, _seed(seed), ptr(sk), g, ptr(key), ptr(ws), ws.numel(),
                                     stream_ptr(stream)))
        return key

    def rotate(self, gkey, g, ct, out=None, stream=None):
        batch = ct.shape[0]
        out = out if out is not None else self.empty(*ct.shape)
        ws = self.workspace(lib.kbfv_rotate_workspace_bytes(self.h, batch))
        check(lib.kbfv_rotate(self.h, ptr(gkey), g, ptr(ct), ptr(out), batch, ptr(ws), ws.numel(),
                              stream_ptr(stream)))
        return out

    def mult_relin_graph(self, rlk, c1, c2, out, ws):
        """Capture kbfv_mult_relin on these buffers as a HIP graph (MultGraph)."""
        return MultGraph(self, rlk, c1, c2, out, ws)

    def encode(self, vals, stream=None):
        check(lib.kbfv_encode(self.h, ptr(vals), vals.shape[0], stream_ptr(stream)))
        return vals

    def decode(self, vals, stream=None):
        check(lib.kbfv_decode(self.h, ptr(vals), vals.shape[0], stream_ptr(stream)))
        return vals

    def rgsw_rows(self, base_bits):
        r = lib.kbfv_rgsw_rows(self.h, base_bits)
        check(0 if r > 0 else r)
        return r

    def wire_save(self, kind, objs, aux=0, stream=None):
        """device objects (leading batch dimension) -> bytes of one wire blob"""
        count = objs.shape[0]
        nb = lib.kbfv_wire_bytes(self.h, kind, count, aux)
        if nb == 0:
            raise KbfvError(lib.kbfv_last_error().decode())
        buf = C.create_string_buffer(nb)
        check(lib.kbfv_wire_save(self.h, kind, ptr(objs), count, aux, buf, nb, stream_ptr(stream)))
        return buf.raw

    def wire_load(self, blob, kind, out, stream=None):
        """wire blob -> the device tensor `out` (leading batch dimension)"""
        check(lib.kbfv_wire_load(self.h, blob, len(blob), kind, ptr(out), out.shape[0], stream_ptr(stream)))
        return out

    def rgsw_encrypt(self, seed, sk, m_coeff, base_bits, stream=None):
        rows = self.rgsw_rows(base_bits)
        out = self.empty(rows, 2, self.qn, self.n)
        ws = self.workspace(lib.kbfv_rgsw_encrypt_workspace_bytes(self.h))
        check(lib.kbfv_rgsw_encrypt(self.h, _seed(seed), ptr(sk), ptr(m_coeff), base_bits, ptr(out), ptr(ws),
                                    ws.numel(), stream_ptr(stream)))
        return out

    def external_product(self, rgsw, ct, base_bits, out=None, stream=None):
        batch = ct.shape[0]
        out = out if out is not None else self.empty(*ct.shape)
        ws = self.workspace(lib.kbfv_external_product_workspace_bytes(self.h, batch, base_bits))
        check(lib.kbfv_external_product(self.h, ptr(rgsw), ptr(ct), ptr(out), batch, base_bits, ptr(ws),
                                        ws.numel(), stream_ptr(stream)))
        return out


def _cache_read(path):
    import json
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


_BUILD_ID = None


class LevelView(Context):
    """A Context at one ciphertext level below the full one (kbfv_context_level):
    same methods, ciphertexts [batch][2][level + 1][N]; owned by its parent."""

    def __init__(self, parent, level):
        h = C.c_void_p()
        check(lib.kbfv_context_level(parent.h, level, C.byref(h)))
        self.h = h
        self.parent = parent  # keeps the owning context alive
        info = (C.c_int * 16)()
        check(lib.kbfv_context_info(self.h, info))
        for k, v in zip(self.INFO, info):
            setattr(self, k, v)
        self.plain = parent.plain
        self.mod = parent.mod
        self.level_index = level
        self._q, self._p, self._t = parent._q, parent._p, parent._t
        self._ws = None

    def close(self):
        self.h = None  # the parent owns the view


def library_build_id():
    """Hash of the loaded libkunpbfv.so: tuned plans are kept per build, so a
    changed kernel is timed again rather than trusted from an older cache."""
    global _BUILD_ID
    if _BUILD_ID is None:
        import hashlib
        h = hashlib.sha1()
        with open(LIB_PATH, "rb") as f:
            for block in iter(lambda: f.read(1 << 20), b""):
                h.update(block)
        _BUILD_ID = h.hexdigest()[:16]
    return _BUILD_ID


def _cache_write(path, key, entry):
    """Add one entry to the JSON plan cache: read, update and rename a
    temporary file over the old one while holding an exclusive lock on a
    side file, so processes tuning at once keep each other's entries."""
    import fcntl
    import json
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(f"{path}.lock", "a") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        try:
            data = _cache_read(path)
            data[key] = entry
            tmp = f"{path}.tmp{os.getpid()}"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)


class RaggedBatch:
    """Device descriptors of polynomials of different RNS lengths (see
    kbfv_ntt_desc_upload); keeps the tensors it names alive."""

    def __init__(self, ctx, items, stream=None):
        torch = _torch()
        desc = np.zeros((len(items), 6), dtype=np.int64)
        for i, it in enumerate(items):
            src, dst, mod0, limbs = it[:4]
            s_limb, d_limb = (it[4], it[5]) if len(it) > 4 else (ctx.n, ctx.n)
            for t, st in ((src, s_limb), (dst, d_limb)):
                if limbs and t.numel() < (limbs - 1) * st + ctx.n:
                    raise KbfvError(f"descriptor {i}: tensor holds fewer than {limbs} limbs")
            desc[i] = (src.data_ptr(), dst.data_ptr(), s_limb, d_limb, mod0, limbs)
        self.desc = torch.empty(desc.size, dtype=torch.int64, device="cuda")
        self.start = torch.empty(len(items) + 1, dtype=torch.int32, device="cuda")
        total = C.c_int()
        check(lib.kbfv_ntt_desc_upload(ctx.h, desc.ctypes.data, len(items), ptr(self.desc), ptr(self.start),
                                       C.byref(total), stream_ptr(stream)))
        self.count, self.total = len(items), total.value
        self._keep = [it[:2] for it in items]


class RaggedBinop:
    """Device descriptors of an element-wise batch over polynomials of
    different RNS lengths (see kbfv_poly_binop_desc_upload)."""

    def __init__(self, ctx, items, stream=None):
        torch = _torch()
        desc = np.zeros((len(items), 8), dtype=np.int64)
        for i, (a, b, out, mod0, limbs) in enumerate(items):
            for t in (a, b, out):
                if limbs and t.numel() < limbs * ctx.n:
                    raise KbfvError(f"descriptor {i}: tensor holds fewer than {limbs} limbs")
            desc[i] = (a.data_ptr(), b.data_ptr(), out.data_ptr(), ctx.n, ctx.n, ctx.n, mod0, limbs)
        self.desc = torch.empty(desc.size, dtype=torch.int64, device="cuda")
        self.start = torch.empty(len(items) + 1, dtype=torch.int32, device="cuda")
        total = C.c_int()
        check(lib.kbfv_poly_binop_desc_upload(ctx.h, desc.ctypes.data, len(items), ptr(self.desc), ptr(self.start),
                                              C.byref(total), stream_ptr(stream)))
        self.count, self.total = len(items), total.value
        self._keep = [it[:3] for it in items]


class MultGraph:
    """A captured batched HomMult+relin; launch() replays it on a stream.  Holds
    references to its buffers so they outlive the graph."""

    def __init__(self, ctx, rlk, c1, c2, out, ws):
        self._bufs = (ctx, rlk, c1, c2, out, ws)
        h = _vp()
        check(lib.kbfv_mult_relin_graph_create(ctx.h, ptr(rlk), ptr(c1), ptr(c2), ptr(out), c1.shape[0],
                                               ptr(ws), ws.numel(), C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        check(lib.kbfv_graph_launch(self.h, stream_ptr(stream)))

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.kbfv_graph_destroy(self.h)
            self.h = None


def xsalsa20(key, nonce, length, block0=0, stream=None):
    torch = _torch()
    out = torch.empty(length, dtype=torch.uint8, device="cuda")
    kb = C.create_string_buffer(bytes(key), 32)
    nb = C.create_string_buffer(bytes(nonce), 24)
    check(lib.kbfv_xsalsa20(kb, nb, block0, ptr(out), length, stream_ptr(stream)))
    return bytes(out.cpu().numpy())
