"""ctypes wrapper over oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Limb arrays are numpy uint64 (one limb per element, reduced mod 2^k); a ring
element of GR(2^k,128) is 128 limbs in the reference layout [slot][alpha].
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_sz = C.c_size_t

VERIFY_ERRORS = {0: "Ok", 1: "MultSumcheck", 2: "LinSumcheck", 3: "LinClaimMismatch",
                 4: "LinARowMismatch", 5: "PcsW", 6: "PcsE", 7: "KernelCheck", 8: "Parse"}


def build(force=False):
    """Compile oracle/liboracle.so with gcc (make)."""
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-B" if force else "-s"], cwd=HERE)
    return LIB


def _coo_args(m):
    r, c, v = m
    return (np.ascontiguousarray(r, np.uint32), np.ascontiguousarray(c, np.uint32),
            np.ascontiguousarray(v, np.uint64), len(r))


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.L = C.CDLL(path)
        L.orc_gr_mul.argtypes = [C.c_int, _u64p, _u64p, _u64p]
        L.orc_gr_mul_gr1.argtypes = [C.c_int, _u64p, C.c_uint64, _u64p]
        L.orc_gr_mul_gr4.argtypes = [C.c_int, _u64p, _u64p, _u64p]
        L.orc_g4_mul.argtypes = [C.c_int, _u64p, _u64p, _u64p]
        L.orc_inv_2adic.argtypes = [C.c_int, C.c_uint64]
        L.orc_inv_2adic.restype = C.c_uint64
        L.orc_lift.argtypes = [C.c_int, _u64p]
        L.orc_t1.argtypes = [C.c_int, _u64p]
        L.orc_pack.argtypes = [C.c_int, _u64p, _sz, _u64p]
        L.orc_psi.argtypes = [C.c_int, _u64p, _sz, _u64p]
        L.orc_in_kernel.argtypes = [C.c_int, _u64p]
        L.orc_quotients.argtypes = [C.c_int, _u64p, _u64p, _u64p, _u64p]
        L.orc_consts.argtypes = [C.c_int, C.c_int, _u64p]
        L.orc_eq_table.argtypes = [C.c_int, _u64p, _sz, _u64p]
        L.orc_fold.argtypes = [C.c_int, _u64p, _sz, _u64p]
        L.orc_blake3.argtypes = [C.c_char_p, _sz, _u8p, _sz]
        L.orc_chacha_block.argtypes = [_u32p, C.c_uint64, C.c_uint64, C.c_int, _u32p]
        L.orc_rng_stream.argtypes = [C.c_int, C.c_uint64, C.c_int, C.c_uint64, _sz, _u64p]
        L.orc_r2_sample.argtypes = [C.c_int, C.c_uint64, _sz, _u64p]
        L.orc_transcript_target.argtypes = [C.c_int, C.c_char_p, C.c_void_p, C.c_char_p, _sz, _u64p]
        L.orc_query_size.argtypes = [C.c_double, _sz]
        L.orc_query_size.restype = _sz
        L.orc_raa_setup.argtypes = [C.c_int, _sz, _sz, C.c_uint64, _u32p, _u32p, _u64p, _u64p]
        L.orc_raa_encode.argtypes = [C.c_int, _sz, _sz, C.c_uint64, _u64p, _u64p, C.c_int]
        L.orc_pcs_shape.argtypes = [C.c_int, _sz, C.c_int, _sz, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)]
        L.orc_proof_size_bytes_for.argtypes = [_sz] * 6
        L.orc_proof_size_bytes_for.restype = _sz
        L.orc_pcs_roundtrip.argtypes = [C.c_int, _sz, _sz, _sz, C.c_uint64, C.c_int, _sz, _u64p, _u64p,
                                        C.c_void_p, C.c_int, _u64p, _u8p]
        L.orc_sumcheck.argtypes = [C.c_int, C.c_int, _sz, _u64p, _u64p, C.c_uint64, C.c_int, _u64p]
        coo = [_u32p, _u32p, _u64p, _sz]
        L.orc_simd_prove.argtypes = [C.c_int, C.c_int, C.c_int] + coo * 3 + [
            _u64p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(_sz)]
        L.orc_simd_verify.argtypes = [C.c_int, C.c_int, C.c_int] + coo * 3 + [
            C.c_uint64, C.c_uint64, C.c_char_p, _sz]
        L.orc_general_prove.argtypes = [C.c_int, C.c_int] + coo * 3 + [
            _u64p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(_sz)]
        L.orc_general_verify.argtypes = [C.c_int, C.c_int] + coo * 3 + [
            C.c_uint64, C.c_uint64, C.c_char_p, _sz]
        mset = [_u64p, _u32p, _u32p, _u64p]
        L.orc_mimd_prove.argtypes = [C.c_int, C.c_int, C.c_int] + mset * 3 + [
            _u64p, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(_sz)]
        L.orc_mimd_verify.argtypes = [C.c_int, C.c_int, C.c_int] + mset * 3 + [
            C.c_uint64, C.c_uint64, C.c_char_p, _sz]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_threads.restype = C.c_int

    def threads(self):
        """OpenMP threads the prover loops run on."""
        return int(self.L.orc_threads())

    # ---- ring ----
    def gr_mul(self, k, a, b):
        o = np.zeros(128, np.uint64)
        self.L.orc_gr_mul(k, np.ascontiguousarray(a, np.uint64), np.ascontiguousarray(b, np.uint64), o)
        return o

    def gr_mul_gr1(self, k, a, s):
        o = np.zeros(128, np.uint64)
        self.L.orc_gr_mul_gr1(k, np.ascontiguousarray(a, np.uint64), int(s), o)
        return o

    def gr_mul_gr4(self, k, a, s4):
        o = np.zeros(128, np.uint64)
        self.L.orc_gr_mul_gr4(k, np.ascontiguousarray(a, np.uint64), np.ascontiguousarray(s4, np.uint64), o)
        return o

    def g4_mul(self, k, a, b):
        o = np.zeros(4, np.uint64)
        self.L.orc_g4_mul(k, np.ascontiguousarray(a, np.uint64), np.ascontiguousarray(b, np.uint64), o)
        return o

    def inv_2adic(self, k, x):
        return int(self.L.orc_inv_2adic(k, x))

    def lift(self, k, x):
        x = np.array(x, np.uint64)
        self.L.orc_lift(k, x)
        return x

    def t1(self, k):
        o = np.zeros(64, np.uint64)
        self.L.orc_t1(k, o)
        return o.reshape(16, 4)

    def pack(self, k, slots):
        s = np.ascontiguousarray(slots, np.uint64).reshape(-1, 32)
        o = np.zeros((s.shape[0], 128), np.uint64)
        self.L.orc_pack(k, s, s.shape[0], o)
        return o

    def psi(self, k, x):
        x = np.ascontiguousarray(x, np.uint64).reshape(-1, 128)
        o = np.zeros((x.shape[0], 32), np.uint64)
        self.L.orc_psi(k, x, x.shape[0], o)
        return o

    def in_kernel(self, k, x):
        return bool(self.L.orc_in_kernel(k, np.ascontiguousarray(x, np.uint64)))

    def quotients(self, k, a, b):
        q1 = np.zeros(128, np.uint64)
        q2 = np.zeros(128, np.uint64)
        self.L.orc_quotients(k, np.ascontiguousarray(a, np.uint64), np.ascontiguousarray(b, np.uint64), q1, q2)
        return q1, q2

    def consts(self, k, which):
        size = {0: 16 * 128, 1: 128, 2: 4, 3: 32 * 128, 4: 256 * 128, 5: 1}[which]
        o = np.zeros(size, np.uint64)
        self.L.orc_consts(k, which, o)
        return o

    def eq_table(self, k, r):
        r = np.ascontiguousarray(r, np.uint64).reshape(-1, 128)
        o = np.zeros((1 << r.shape[0], 128), np.uint64)
        self.L.orc_eq_table(k, r, r.shape[0], o)
        return o

    def fold(self, k, t, r):
        t = np.array(t, np.uint64).reshape(-1, 128)
        self.L.orc_fold(k, t, t.shape[0], np.ascontiguousarray(r, np.uint64))
        return t[: t.shape[0] // 2]

    # ---- hashing / rng ----
    def blake3(self, data, olen=32):
        o = np.zeros(olen, np.uint8)
        self.L.orc_blake3(bytes(data), len(data), o, olen)
        return o.tobytes()

    def chacha_block(self, key, counter, stream=0, drounds=10):
        o = np.zeros(16, np.uint32)
        self.L.orc_chacha_block(np.ascontiguousarray(key, np.uint32), counter, stream, drounds, o)
        return o

    def rng_stream(self, k, seed, kind, cnt, arg=0):
        o = np.zeros(cnt, np.uint64)
        self.L.orc_rng_stream(k, seed, kind, arg, cnt, o)
        return o

    def r2_sample(self, k, seed, cnt):
        o = np.zeros((cnt, 128), np.uint64)
        self.L.orc_r2_sample(k, seed, cnt, o)
        return o

    def transcript_target(self, k, label, root, tlabel, cnt):
        o = np.zeros((cnt, 128), np.uint64)
        rb = C.create_string_buffer(bytes(root), 32) if root is not None else None
        self.L.orc_transcript_target(k, label.encode(), rb, tlabel.encode(), cnt, o)
        return o

    # ---- pcs ----
    def query_size(self, delta=0.19, lam=100):
        return int(self.L.orc_query_size(delta, lam))

    def raa_setup(self, k, kk, q, seed):
        n = kk * q
        p1 = np.zeros(n, np.uint32)
        p2 = np.zeros(n, np.uint32)
        d1 = np.zeros(n, np.uint64)
        d2 = np.zeros(n, np.uint64)
        self.L.orc_raa_setup(k, kk, q, seed, p1, p2, d1, d2)
        return p1, p2, d1, d2

    def raa_encode(self, k, kk, q, seed, msg, packed):
        w = 32 if packed else 128
        msg = np.ascontiguousarray(msg, np.uint64).reshape(kk, w)
        o = np.zeros((kk * q, w), np.uint64)
        self.L.orc_raa_encode(k, kk, q, seed, msg, o, 1 if packed else 0)
        return o

    def pcs_shape(self, k, nv, kind, arg=0):
        k1, k2, ell = _sz(), _sz(), _sz()
        self.L.orc_pcs_shape(k, nv, kind, arg, C.byref(k1), C.byref(k2), C.byref(ell))
        return k1.value, k2.value, ell.value

    def proof_size_bytes_for(self, *a):
        return int(self.L.orc_proof_size_bytes_for(*a))

    def pcs_roundtrip(self, k, k1, k2, ell, seed, mode, s, msgs, z, w=None, tamper=False):
        msgs = np.ascontiguousarray(msgs, np.uint64)
        z = np.ascontiguousarray(z, np.uint64)
        nv = 128 * (s if mode == 4 else 1)
        vals = np.zeros(nv, np.uint64)
        root = np.zeros(32, np.uint8)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, np.uint64)
            wp = w.ctypes.data
        ok = self.L.orc_pcs_roundtrip(k, k1, k2, ell, seed, mode, s, msgs, z, wp, 1 if tamper else 0, vals, root)
        return bool(ok), vals.reshape(-1, 128), root.tobytes()

    def sumcheck(self, k, which, nv, tabs, rt, seed, tamper=False):
        claim = np.zeros(128, np.uint64)
        rt = np.ascontiguousarray(rt if rt is not None else np.zeros((1, 128)), np.uint64)
        rc = self.L.orc_sumcheck(k, which, nv, np.ascontiguousarray(tabs, np.uint64), rt, seed,
                                 1 if tamper else 0, claim)
        return rc, claim

    # ---- snarks ----
    def _bytes(self, p, n):
        b = C.string_at(p, n)
        self.L.orc_free(p)
        return b

    def simd_prove(self, k, log_n, log_m, A, B, Cm, wit, seed_w=0xDEAD, seed_q=0xCAFE):
        out, ln = C.c_void_p(), _sz()
        rc = self.L.orc_simd_prove(k, log_n, log_m, *_coo_args(A), *_coo_args(B), *_coo_args(Cm),
                                   np.ascontiguousarray(wit, np.uint64).reshape(-1), seed_w, seed_q,
                                   C.byref(out), C.byref(ln))
        if rc != 0:
            raise RuntimeError("oracle simd prover self-check failed")
        return self._bytes(out, ln.value)

    def simd_verify(self, k, log_n, log_m, A, B, Cm, proof, seed_w=0xDEAD, seed_q=0xCAFE):
        return VERIFY_ERRORS[self.L.orc_simd_verify(k, log_n, log_m, *_coo_args(A), *_coo_args(B),
                                                    *_coo_args(Cm), seed_w, seed_q, proof, len(proof))]

    def general_prove(self, k, log_b, A, B, Cm, w, seed_w=0xDEAD, seed_e=0xCAFE):
        out, ln = C.c_void_p(), _sz()
        rc = self.L.orc_general_prove(k, log_b, *_coo_args(A), *_coo_args(B), *_coo_args(Cm),
                                      np.ascontiguousarray(w, np.uint64), seed_w, seed_e,
                                      C.byref(out), C.byref(ln))
        if rc != 0:
            raise RuntimeError("oracle general prover self-check failed")
        return self._bytes(out, ln.value)

    def general_verify(self, k, log_b, A, B, Cm, proof, seed_w=0xDEAD, seed_e=0xCAFE):
        return VERIFY_ERRORS[self.L.orc_general_verify(k, log_b, *_coo_args(A), *_coo_args(B),
                                                       *_coo_args(Cm), seed_w, seed_e, proof, len(proof))]

    # MIMD: each matrix set is a list of 32 COO triples (rows, cols, vals)
    def mimd_prove(self, k, log_n, log_m, As, Bs, Cs, wit, seed_w=0xDEAD, seed_e=0xCAFE):
        out, ln = C.c_void_p(), _sz()
        rc = self.L.orc_mimd_prove(k, log_n, log_m, *_set_args(As), *_set_args(Bs), *_set_args(Cs),
                                   np.ascontiguousarray(wit, np.uint64).reshape(-1), seed_w, seed_e,
                                   C.byref(out), C.byref(ln))
        if rc != 0:
            raise RuntimeError("oracle mimd prover self-check failed")
        return self._bytes(out, ln.value)

    def mimd_verify(self, k, log_n, log_m, As, Bs, Cs, proof, seed_w=0xDEAD, seed_e=0xCAFE):
        return VERIFY_ERRORS[self.L.orc_mimd_verify(k, log_n, log_m, *_set_args(As), *_set_args(Bs),
                                                    *_set_args(Cs), seed_w, seed_e, proof, len(proof))]


def _set_args(ms):
    if len(ms) != 32:
        raise ValueError("a MIMD matrix set holds 32 instance matrices")
    off = np.zeros(33, np.uint64)
    off[1:] = np.cumsum([len(m[0]) for m in ms])
    return (off, np.ascontiguousarray(np.concatenate([m[0] for m in ms]), np.uint32),
            np.ascontiguousarray(np.concatenate([m[1] for m in ms]), np.uint32),
            np.ascontiguousarray(np.concatenate([m[2] for m in ms]), np.uint64))


_ORACLE = None


def load():
    global _ORACLE
    if _ORACLE is None:
        build()
        _ORACLE = Oracle()
    return _ORACLE
