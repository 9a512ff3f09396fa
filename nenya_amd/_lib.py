"""ctypes binding of include/nenya.h (the library is loaded lazily)."""
import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
_CODES = {1: "shape", 2: "out of device memory", 3: "device", 4: "argument", 5: "no GPU", 6: "canceled"}


class NenyaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nenya error {code} ({_CODES.get(code, '?')}): {msg}")
        self.code = code


def lib_path():
    return os.path.join(HERE, "libnenya_amd.so")


def _lib():
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise NenyaError(3, f"{p} is not built; run `python -m nenya_amd.build`")
        L = C.CDLL(p)
        vp, sz, u64 = C.c_void_p, C.c_size_t, C.c_uint64
        coo = [vp, vp, vp, sz]
        L.nenya_params_simd_secure.argtypes = [C.c_int, C.c_int, C.c_int, u64, u64, C.POINTER(vp)]
        L.nenya_prove_simd.argtypes = [vp] + coo * 3 + [vp, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_simd_view.argtypes = [vp] + coo * 3 + [vp, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_mimd_view.argtypes = [vp, vp, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_general_view.argtypes = [vp, vp, sz, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_params_general_secure.argtypes = [C.c_int, C.c_int] + coo * 3 + [u64, u64, C.POINTER(vp)]
        mset = [vp, vp, vp, vp]
        L.nenya_params_mimd_secure.argtypes = [C.c_int, C.c_int, C.c_int] + mset * 3 + [u64, u64, C.POINTER(vp)]
        L.nenya_prove_mimd.argtypes = [vp, vp, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_submit_mimd.argtypes = [vp, vp, vp, C.POINTER(u64)]
        L.nenya_verify_simd.argtypes = [vp] + coo * 3 + [C.c_char_p, sz, C.POINTER(C.c_int)]
        L.nenya_verify_mimd.argtypes = [vp, C.c_char_p, sz, C.POINTER(C.c_int)]
        L.nenya_verify_general.argtypes = [vp, C.c_char_p, sz, C.POINTER(C.c_int)]
        L.nenya_prove_general.argtypes = [vp, vp, sz, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_general_device.argtypes = [vp, vp, sz, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_general_device_view.argtypes = [vp, vp, sz, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_prove_general_device_noalloc.argtypes = [vp, vp, sz, vp, sz, C.POINTER(sz)]
        L.nenya_set_timing.argtypes = [vp, C.c_int]
        L.nenya_last_timing.argtypes = [vp, C.POINTER(C.c_double), C.c_int, C.c_char_p, sz]
        L.nenya_setup_timing.argtypes = [vp, C.POINTER(C.c_double), C.c_int, C.c_char_p, sz]
        L.nenya_set_profile.argtypes = [vp, C.c_int]
        L.nenya_set_profile_only.argtypes = [vp, C.c_char_p]
        dp = C.POINTER(C.c_double)
        L.nenya_kernel_stats.argtypes = [vp, C.c_int, C.c_char_p, sz, dp, dp, dp, dp]
        L.nenya_server_create.argtypes = [C.c_int, sz, C.POINTER(vp)]
        L.nenya_submit_general.argtypes = [vp, vp, vp, sz, C.POINTER(u64)]
        L.nenya_submit_simd.argtypes = [vp, vp] + coo * 3 + [vp, C.POINTER(u64)]
        L.nenya_poll.argtypes = [vp, u64]
        L.nenya_wait.argtypes = [vp, u64, C.POINTER(vp), C.POINTER(sz), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.nenya_job_release.argtypes = [vp, u64]
        L.nenya_cancel.argtypes = [vp, u64]
        L.nenya_proof_container.argtypes = [vp, C.c_char_p, sz, C.POINTER(vp), C.POINTER(sz)]
        L.nenya_proof_payload.argtypes = [C.c_char_p, sz, C.POINTER(vp), C.POINTER(sz), C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.nenya_server_free.argtypes = [vp]
        L.nenya_params_free.argtypes = [vp]
        L.nenya_free.argtypes = [vp]
        L.nenya_last_error.restype = C.c_char_p
        L.nenya_device_count.restype = C.c_int
        L.nenya_device_copy.argtypes = [vp, vp, sz]
        L.nenya_autotune.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, sz]
        _LIB = L
    return _LIB


def device_count():
    return int(_lib().nenya_device_count())


def _check(rc):
    if rc != 0:
        raise NenyaError(rc, _lib().nenya_last_error().decode(errors="replace"))


def _width(k):
    if k not in _DT:
        raise NenyaError(4, "k must be 8, 16, 32 or 64")
    return _DT[k]


def _limbs(x, k):
    """x as a contiguous array of k-bit limbs (wrapping); no copy when it
    already is one (the library then borrows the caller's array)."""
    a = np.asarray(x)
    if a.dtype != _DT[k]:
        a = a.astype(_DT[k])
    return np.ascontiguousarray(a)


def _coo(m, k):
    _width(k)
    r, c, v = m
    r = np.ascontiguousarray(r, np.uint32)
    c = np.ascontiguousarray(c, np.uint32)
    v = _limbs(v, k)
    if not (len(r) == len(c) == len(v)):
        raise NenyaError(1, "COO arrays differ in length")
    return (r, c, v)


def _coo_args(m):
    r, c, v = m
    return [r.ctypes.data, c.ctypes.data, v.ctypes.data, len(r)]


def _take_proof(buf, n):
    try:
        return C.string_at(buf, n.value)
    finally:
        _lib().nenya_free(buf)


class OwnedProof:
    """An owned proof as the C ABI returns it: the pinned buffer the proof
    was assembled in, handed over without a copy and released with
    nenya_free by close() (or when the object is collected).  bytes(p) copies
    it out; views from view() must not outlive close()."""

    def __init__(self, buf, n):
        self._buf, self._n = buf, n.value

    def __len__(self):
        return self._n

    def __bytes__(self):
        if self._buf is None:
            raise ValueError("proof released")
        return C.string_at(self._buf, self._n)

    def view(self):
        if self._buf is None:
            raise ValueError("proof released")
        return _view(self._buf, C.c_size_t(self._n), self)

    def close(self):
        if self._buf is not None and _LIB is not None:
            _LIB.nenya_free(self._buf)
        self._buf = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def _result(out, n, params, view, copy):
    if view:
        return _view(out, n, params)
    if not copy:
        return OwnedProof(out, n)
    return _take_proof(out, n)


class _Params:
    kind = None

    def __init__(self, handle, k):
        self._h = handle
        self.k = k

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.nenya_params_free(h)
            self._h = None

    def set_timing(self, on=True):
        _check(_lib().nenya_set_timing(self._h, 1 if on else 0))

    def set_profile(self, on=True):
        _check(_lib().nenya_set_profile(self._h, 1 if on else 0))

    def set_profile_only(self, kernel=None):
        _check(_lib().nenya_set_profile_only(self._h, (kernel or "").encode()))

    def kernel_stats(self):
        cap = 64
        arr = [(C.c_double * cap)() for _ in range(4)]
        names = C.create_string_buffer(8192)
        n = _lib().nenya_kernel_stats(self._h, cap, names, 8192, *arr)
        labels = names.value.decode().split("\n")[:n]
        return {labels[i]: {"ms": arr[0][i], "count": int(arr[1][i]), "bytes": arr[2][i], "macs": arr[3][i]}
                for i in range(n)}

    def last_timing(self):
        return self._timing(_lib().nenya_last_timing)

    def setup_timing(self):
        """Wall-clock ms of the setup stages that made these params."""
        return self._timing(_lib().nenya_setup_timing)

    def _timing(self, fn):
        ms = (C.c_double * 64)()
        names = C.create_string_buffer(4096)
        n = fn(self._h, ms, 64, names, 4096)
        labels = names.value.decode().split("\n")[:n]
        return {labels[i]: ms[i] for i in range(n)}


class SimdSnarkParams(_Params):
    """SimdSnarkParams::secure (snark-ring/src/simd.rs:112-119)."""

    @classmethod
    def secure(cls, log_n, log_m, seed_w=0xDEAD, seed_q=0xCAFE, k=32):
        h = C.c_void_p()
        _check(_lib().nenya_params_simd_secure(k, log_n, log_m, seed_w, seed_q, C.byref(h)))
        p = cls(h, k)
        p.log_n, p.log_m = log_n, log_m
        return p


@dataclass
class SimdR1csInstance:
    """reduction/src/instance.rs:11-21: shared (A, B, C) and 32 witnesses."""
    a: tuple
    b: tuple
    c: tuple
    witnesses: np.ndarray  # shape (32, m)


def _view(out, n, params):
    """Read-only bytes view of a proof in the params' pinned buffer; the view
    keeps `params` (the buffer's owner) alive."""
    arr = (C.c_ubyte * n.value).from_address(out.value)
    arr._owner = params
    return memoryview(arr).cast("B").toreadonly()


def prove_simd(inst, params, view=False):
    """prove_simd(&SimdR1csInstance, &SimdSnarkParams) -> proof bytes.
    view=True returns a read-only memoryview of the params' pinned proof
    buffer instead (no copy; valid until the next proof with these params)."""
    k = params.k
    mats = [_coo(m, k) for m in (inst.a, inst.b, inst.c)]
    wit = _limbs(inst.witnesses, k)
    if wit.shape != (32, 1 << params.log_m):
        raise NenyaError(1, f"witnesses must have shape (32, {1 << params.log_m})")
    out, n = C.c_void_p(), C.c_size_t()
    args = [params._h]
    for m in mats:
        args += _coo_args(m)
    if view:
        _check(_lib().nenya_prove_simd_view(*args, wit.ctypes.data, C.byref(out), C.byref(n)))
        return _view(out, n, params)
    _check(_lib().nenya_prove_simd(*args, wit.ctypes.data, C.byref(out), C.byref(n)))
    return _take_proof(out, n)


def _mimd_set(ms, k):
    """32 COO matrices -> (offsets[33] as size_t, rows, cols, vals) arrays."""
    _width(k)
    if len(ms) != 32:
        raise NenyaError(4, "a MIMD matrix set holds 32 instance matrices")
    mats = [_coo(m, k) for m in ms]
    off = np.zeros(33, np.uintp)
    off[1:] = np.cumsum([len(m[0]) for m in mats])
    return (off, np.ascontiguousarray(np.concatenate([m[0] for m in mats]), np.uint32),
            np.ascontiguousarray(np.concatenate([m[1] for m in mats]), np.uint32),
            np.ascontiguousarray(np.concatenate([m[2] for m in mats]).astype(_DT[k])))


class MimdSnarkParams(_Params):
    """MimdSnarkParams::secure(log_n, log_m, pack_mimd_matrices(a, b, c), seed_w, seed_e)
    (snark-ring/src/mimd.rs:84-107, reduction/src/reduce.rs:249-273).
    `matrices` = (As, Bs, Cs), each a list of the 32 instances' COO triples."""

    @classmethod
    def secure(cls, log_n, log_m, matrices, seed_w=0xDEAD, seed_e=0xCAFE, k=32):
        sets = [_mimd_set(ms, k) for ms in matrices]
        args = []
        for st in sets:
            args += [a.ctypes.data for a in st]
        h = C.c_void_p()
        _check(_lib().nenya_params_mimd_secure(k, log_n, log_m, *args, seed_w, seed_e, C.byref(h)))
        p = cls(h, k)
        p.log_n, p.log_m = log_n, log_m
        return p


def _mimd_witnesses(witnesses, params):
    wit = _limbs(witnesses, params.k)
    if wit.shape != (32, 1 << params.log_m):
        raise NenyaError(1, f"witnesses must have shape (32, {1 << params.log_m})")
    return wit


def prove_mimd(witnesses, params, view=False):
    """prove_mimd(&[Vec<Gr1>; 32], &MimdSnarkParams) -> proof bytes (view=True:
    a read-only view of the params' pinned proof buffer, as prove_simd)."""
    wit = _mimd_witnesses(witnesses, params)
    out, n = C.c_void_p(), C.c_size_t()
    if view:
        _check(_lib().nenya_prove_mimd_view(params._h, wit.ctypes.data, C.byref(out), C.byref(n)))
        return _view(out, n, params)
    _check(_lib().nenya_prove_mimd(params._h, wit.ctypes.data, C.byref(out), C.byref(n)))
    return _take_proof(out, n)


class GeneralSnarkParams(_Params):
    """GeneralSnarkParams::secure(log_blocks, pack_general_matrices(a,b,c), seed_w, seed_e)
    (snark-ring/src/general.rs:69-95)."""

    @classmethod
    def secure(cls, log_blocks, matrices, seed_w=0xDEAD, seed_e=0xCAFE, k=32):
        mats = [_coo(m, k) for m in matrices]
        args = []
        for m in mats:
            args += _coo_args(m)
        h = C.c_void_p()
        _check(_lib().nenya_params_general_secure(k, log_blocks, *args, seed_w, seed_e, C.byref(h)))
        p = cls(h, k)
        p.log_blocks = log_blocks
        return p


def prove_general(w, params, view=False, copy=True):
    """prove_general(w, &GeneralSnarkParams) -> proof bytes (view=True: a
    read-only view of the params' pinned proof buffer, as prove_simd;
    copy=False: the owned buffer as the C ABI returns it, an OwnedProof)."""
    w = _limbs(w, params.k)
    out, n = C.c_void_p(), C.c_size_t()
    if view:
        _check(_lib().nenya_prove_general_view(params._h, w.ctypes.data, len(w), C.byref(out), C.byref(n)))
    else:
        _check(_lib().nenya_prove_general(params._h, w.ctypes.data, len(w), C.byref(out), C.byref(n)))
    return _result(out, n, params, view, copy)


def prove_general_device(dev_ptr, n_limbs, params, view=False):
    """prove_general with the witness already in device memory (e.g. a
    DeviceArray's ptr).  view=True returns a read-only memoryview of the
    proof in the params' pinned buffer (no copy; valid until the next proof
    with these params)."""
    out, n = C.c_void_p(), C.c_size_t()
    if view:
        _check(_lib().nenya_prove_general_device_view(params._h, C.c_void_p(dev_ptr), n_limbs, C.byref(out),
                                                      C.byref(n)))
        return _view(out, n, params)
    _check(_lib().nenya_prove_general_device(params._h, C.c_void_p(dev_ptr), n_limbs, C.byref(out), C.byref(n)))
    return _take_proof(out, n)


_HIP = None


def _hip():
    """The HIP runtime the library itself runs on (libamdhip64.so.7, already
    loaded with libnenya_amd.so): device buffers for callers that hand the
    prover a witness in device memory."""
    global _HIP
    if _HIP is None:
        _lib()
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _HIP = H
    return _HIP


def device_synchronize():
    if _hip().hipDeviceSynchronize() != 0:
        raise NenyaError(3, "hipDeviceSynchronize failed")


def autotune(widths=(8, 16, 32, 64), log_min=8, log_max=20):
    """Run the launch-shape autotuner (nenya_autotune) for the given widths;
    later proofs in this process use its picks.  Returns the report (dict)."""
    import json
    mask = 0
    for k in widths:
        mask |= 1 << {8: 0, 16: 1, 32: 2, 64: 3}[k]
    buf = C.create_string_buffer(1 << 16)
    _check(_lib().nenya_autotune(mask, log_min, log_max, buf, len(buf)))
    return json.loads(buf.value.decode())


def device_copy(dst, src):
    """dst <- src (DeviceArrays of equal size) with the library's copy kernel."""
    if dst.nbytes != src.nbytes:
        raise NenyaError(1, "device_copy: sizes differ")
    _check(_lib().nenya_device_copy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), src.nbytes))


class DeviceArray:
    """A copy of a host array in device memory (hipMalloc'd, freed with the
    object); `.ptr` is the device address, `.size` the element count."""

    def __init__(self, host):
        a = np.ascontiguousarray(host)
        self.size, self.nbytes, self.dtype = a.size, a.nbytes, a.dtype
        p = C.c_void_p()
        if _hip().hipMalloc(C.byref(p), max(a.nbytes, 1)) != 0:
            raise NenyaError(2, f"hipMalloc of {a.nbytes} bytes failed")
        self._p = p
        if a.nbytes and _hip().hipMemcpy(p, a.ctypes.data, a.nbytes, 1) != 0:  # hipMemcpyHostToDevice
            self.free()
            raise NenyaError(3, "hipMemcpy to the device failed")

    @property
    def ptr(self):
        return self._p.value

    def free(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value and _HIP is not None:
            _HIP.hipFree(p)
        self._p = None

    def __del__(self):
        self.free()


def prove_general_device_noalloc(dev_ptr, n_limbs, params, out=None):
    """prove_general with the witness in device memory through the entry
    point that allocates nothing (nenya_prove_general_device_noalloc): the
    proof is written into `out` (a writable buffer of at least the proof's
    size; one is made when None) and returned as bytes."""
    need = C.c_size_t()
    if out is None:
        rc = _lib().nenya_prove_general_device_noalloc(params._h, C.c_void_p(dev_ptr), n_limbs, None, 0,
                                                       C.byref(need))
        if rc != 4 or need.value == 0:
            _check(rc)
        out = bytearray(need.value)
    buf = (C.c_ubyte * len(out)).from_buffer(out)
    _check(_lib().nenya_prove_general_device_noalloc(params._h, C.c_void_p(dev_ptr), n_limbs, buf, len(out),
                                                     C.byref(need)))
    return bytes(out[:need.value])


def proof_container(params, proof):
    """The versioned container of a proof made with `params` (header + the
    serialized bytes; nenya_proof_container)."""
    out, n = C.c_void_p(), C.c_size_t()
    b = bytes(proof)
    _check(_lib().nenya_proof_container(params._h, b, len(b), C.byref(out), C.byref(n)))
    return _take_proof(out, n)


def proof_payload(container):
    """(payload bytes, model, k) of an intact container (nenya_proof_payload);
    raises NenyaError (code 4) otherwise."""
    b = bytes(container)
    pl, n, model, k = C.c_void_p(), C.c_size_t(), C.c_int(), C.c_int()
    _check(_lib().nenya_proof_payload(b, len(b), C.byref(pl), C.byref(n), C.byref(model), C.byref(k)))
    return b[72:72 + n.value], model.value, k.value


VERIFY_ERRORS = {0: "Ok", 1: "MultSumcheck", 2: "LinSumcheck", 3: "LinClaimMismatch", 4: "LinARowMismatch",
                 5: "PcsW", 6: "PcsE", 7: "KernelCheck", 8: "Parse"}


def verify_simd(inst, params, proof):
    """verify_simd(a, b, c, &params, &proof) (snark-ring/src/simd.rs:380-522)
    -> "Ok" or the VerifyError variant name ("Parse" for malformed bytes)."""
    k = params.k
    mats = [_coo(m, k) for m in (inst.a, inst.b, inst.c)]  # referenced until the call returns
    args = [params._h]
    for m in mats:
        args += _coo_args(m)
    v = C.c_int()
    _check(_lib().nenya_verify_simd(*args, bytes(proof), len(proof), C.byref(v)))
    return VERIFY_ERRORS[v.value]


def verify_mimd(params, proof):
    """verify_mimd(&params, &proof) (snark-ring/src/mimd.rs:403-551)."""
    v = C.c_int()
    _check(_lib().nenya_verify_mimd(params._h, bytes(proof), len(proof), C.byref(v)))
    return VERIFY_ERRORS[v.value]


def verify_general(params, proof):
    """verify_general(&params, &proof) (snark-ring/src/general.rs:545-698)."""
    v = C.c_int()
    _check(_lib().nenya_verify_general(params._h, bytes(proof), len(proof), C.byref(v)))
    return VERIFY_ERRORS[v.value]


class Server:
    """Batched / served proofs on one GPU (nenya_server_* in include/nenya.h):
    `workers` host threads with their own streams drain a FIFO job queue,
    admitting jobs by device-memory footprint.  Inputs are kept alive by the
    Server until the job is collected; inputs that already are contiguous
    arrays of the params' limb type are borrowed, not copied, so they must
    not change until then (the reference prover borrows its inputs too)."""

    def __init__(self, workers=4, mem_budget=0):
        h = C.c_void_p()
        _check(_lib().nenya_server_create(workers, mem_budget, C.byref(h)))
        self._h = h
        self._keep = {}

    def submit_general(self, params, w):
        w = _limbs(w, params.k)
        j = C.c_uint64()
        _check(_lib().nenya_submit_general(self._h, params._h, w.ctypes.data, len(w), C.byref(j)))
        self._keep[j.value] = (params, w)
        return j.value

    def submit_simd(self, inst, params):
        k = params.k
        mats = [_coo(m, k) for m in (inst.a, inst.b, inst.c)]
        wit = _limbs(inst.witnesses, k)
        if wit.shape != (32, 1 << params.log_m):
            raise NenyaError(1, f"witnesses must have shape (32, {1 << params.log_m})")
        args = [self._h, params._h]
        for m in mats:
            args += _coo_args(m)
        j = C.c_uint64()
        _check(_lib().nenya_submit_simd(*args, wit.ctypes.data, C.byref(j)))
        self._keep[j.value] = (params, mats, wit)
        return j.value

    def submit_mimd(self, params, witnesses):
        wit = _mimd_witnesses(witnesses, params)
        j = C.c_uint64()
        _check(_lib().nenya_submit_mimd(self._h, params._h, wit.ctypes.data, C.byref(j)))
        self._keep[j.value] = (params, wit)
        return j.value

    def cancel(self, job):
        """Cancel a job (queued: dropped; running: stops at its next
        checkpoint); wait() on it then raises NenyaError with code 6."""
        _check(_lib().nenya_cancel(self._h, job))

    def poll(self, job):
        r = _lib().nenya_poll(self._h, job)
        if r < 0:
            _check(-r)
        return r == 1

    def wait(self, job, copy=True):
        """-> (proof bytes, queued_ms, run_ms); releases the job.  With
        copy=False the proof is not copied out of its pinned buffer and its
        length is returned in place of the bytes."""
        buf, n = C.c_void_p(), C.c_size_t()
        q, r = C.c_double(), C.c_double()
        try:
            _check(_lib().nenya_wait(self._h, job, C.byref(buf), C.byref(n), C.byref(q), C.byref(r)))
            return (C.string_at(buf, n.value) if copy else n.value), q.value, r.value
        finally:
            _lib().nenya_job_release(self._h, job)
            self._keep.pop(job, None)

    def close(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.nenya_server_free(h)
            self._h = None

    def __del__(self):
        self.close()
