"""Python binding of the MI355X TFHE gate-bootstrapping engine.

A thin ctypes layer over ``libcufhe_amd.so`` (C ABI in include/cufhe_amd.h
and include/tfhe_host.h). It exists for tests, benchmarks and scripting; the
drop-in interface for cuFHEpp users is the C++ header include/cufhe_gpu.hpp.
There is no CPU fallback: every GPU entry point runs the HIP kernels, and
loading fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "CufheError", "lib", "GATES", "N0", "N1", "N2", "TLWE0_WORDS", "TLWE1_WORDS",
    "TLWE2_WORDS", "MU0", "MU1", "MU2", "secret_key_gen", "bk_gen", "ksk_gen",
    "encrypt_bits", "phase", "decrypt_bits", "DeviceBuffer", "Stream", "Plan",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUFHE_AMD_LIB") or os.path.join(_HERE, "libcufhe_amd.so")

N0, N1, N2 = 636, 1024, 2048
TLWE0_WORDS, TLWE1_WORDS, TLWE2_WORDS = N0 + 1, N1 + 1, N2 + 1
MU0 = MU1 = 1 << 29
MU2 = 1 << 61

GATES = {
    "NAND": 0, "OR": 1, "ORYN": 2, "ORNY": 3, "AND": 4, "ANDYN": 5, "ANDNY": 6,
    "NOR": 7, "XOR": 8, "XNOR": 9, "MUX": 10, "NMUX": 11, "NOT": 12, "COPY": 13,
}

# plaintext truth tables (reference test/plain.h:10-69)
TRUTH = {
    "NAND": lambda a, b: 1 - (a & b),
    "OR": lambda a, b: a | b,
    "ORYN": lambda a, b: a | (1 - b),
    "ORNY": lambda a, b: (1 - a) | b,
    "AND": lambda a, b: a & b,
    "ANDYN": lambda a, b: a & (1 - b),
    "ANDNY": lambda a, b: (1 - a) & b,
    "NOR": lambda a, b: 1 - (a | b),
    "XOR": lambda a, b: a ^ b,
    "XNOR": lambda a, b: 1 - (a ^ b),
}


class CufheError(RuntimeError):
    pass


class GateDesc(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("out", ctypes.c_uint32), ("in0", ctypes.c_uint32),
                ("in1", ctypes.c_uint32), ("inc", ctypes.c_uint32)]


class CircuitDesc(ctypes.Structure):
    _fields_ = [("n_inputs", ctypes.c_int), ("n_gates", ctypes.c_int),
                ("gates", ctypes.c_void_p), ("n_layers", ctypes.c_int),
                ("layer_end", ctypes.c_void_p), ("n_outputs", ctypes.c_int),
                ("outputs", ctypes.c_void_p)]


_lib = None


def lib():
    """Load libcufhe_amd.so (built in-tree by cufhe_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CufheError("%s is not built: run `python cufhe_amd/build.py`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u32, u64 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                             ctypes.c_uint32, ctypes.c_uint64)
    sigs = {
        "cufhe_version": ([], ctypes.c_char_p),
        "cufhe_last_error": ([], ctypes.c_char_p),
        "cufhe_set_gpu_num": ([i32], i32),
        "cufhe_initialize": ([], i32),
        "cufhe_initialize_lvl01": ([vp, sz, i32, vp, sz], i32),
        "cufhe_initialize_lvl02": ([vp, sz, i32, vp, sz], i32),
        "cufhe_initialize_lvl01_ex": ([vp, sz, i32, i32, vp, sz, i32], i32),
        "cufhe_ntt_transform": ([vp, vp, i32, i32, vp], i32),
        "cufhe_ntt_root": ([], u32),
        "cufhe_initialize_lvl02_ex": ([vp, sz, i32, i32, vp, sz], i32),
        "cufhe_ntt64_transform": ([vp, vp, i32, i32, vp], i32),
        "cufhe_ntt64_root": ([], u64),
        "cufhe_set_lvl2_jobs_per_group": ([i32], i32),
        "cufhe_cleanup": ([], i32),
        "cufhe_cleanup_lvl02": ([], i32),
        "cufhe_synchronize": ([], i32),
        "cufhe_stream_create": ([ctypes.POINTER(vp)], i32),
        "cufhe_stream_destroy": ([vp], i32),
        "cufhe_stream_query": ([vp], i32),
        "cufhe_stream_synchronize": ([vp], i32),
        "cufhe_device_alloc": ([ctypes.POINTER(vp), sz], i32),
        "cufhe_device_free": ([vp], i32),
        "cufhe_host_alloc": ([ctypes.POINTER(vp), sz], i32),
        "cufhe_host_free": ([vp], i32),
        "cufhe_memcpy_h2d": ([vp, vp, sz, vp], i32),
        "cufhe_memcpy_d2h": ([vp, vp, sz, vp], i32),
        "cufhe_memset_d": ([vp, i32, sz, vp], i32),
        "cufhe_memcpy_d2d": ([vp, vp, sz, vp], i32),
        "cufhe_plan_jobs_per_group": ([vp], i32),
        "cufhe_plan_run_timed": ([vp, vp, ctypes.POINTER(ctypes.c_float)], i32),
        "cufhe_set_key_switch_split": ([i32], i32),
        "cufhe_autotune_lvl1": ([i32, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_float), vp], i32),
        "cufhe_gate_lvl1": ([i32, vp, vp, vp, vp, vp], i32),
        "cufhe_gate_lvl02": ([i32, vp, vp, vp, vp, vp], i32),
        "cufhe_gate_lvl0": ([i32, vp, vp, vp, vp, vp], i32),
        "cufhe_plan_create_lvl01": ([vp, i32, vp, i32, i32, ctypes.POINTER(vp)], i32),
        "cufhe_sample_extract_key_switch_lvl10": ([vp, vp, i32, vp], i32),
        "cufhe_bootstrap_to_trlwe_lvl01": ([vp, vp, u32, i32, vp], i32),
        "cufhe_refresh_lvl1": ([vp, vp, i32, vp], i32),
        "cufhe_plan_create_lvl1": ([vp, i32, vp, i32, i32, ctypes.POINTER(vp)], i32),
        "cufhe_plan_run": ([vp, vp], i32),
        "cufhe_plan_create_lvl02": ([vp, i32, vp, i32, ctypes.POINTER(vp)], i32),
        "cufhe_plan_destroy": ([vp], i32),
        "cufhe_plan_bootstraps": ([vp], i32),
        "cufhe_gate_batch_lvl1": ([vp, i32, vp, i32, vp], i32),
        "cufhe_identity_key_switch_lvl10": ([vp, vp, i32, vp], i32),
        "cufhe_identity_key_switch_lvl20": ([vp, vp, i32, vp], i32),
        "cufhe_bootstrap": ([vp, vp, u32, vp], i32),
        "cufhe_bootstrap_lvl01": ([vp, vp, u32, i32, i32, vp], i32),
        "cufhe_bootstrap_lvl02": ([vp, vp, u64, i32, vp], i32),
        "cufhe_ffnt_forward": ([i32, vp, i32, vp, i32, vp], i32),
        "cufhe_ffnt_inverse": ([i32, vp, vp, i32, vp], i32),
        "cufhe_trgsw_to_fourier": ([i32, vp, vp, i32, vp], i32),
        "cufhe_cmux_accumulate": ([i32, vp, vp, u32, vp], i32),
        "cufhe_trgsw_transform_bytes_lvl1": ([], sz),
        "cufhe_trgsw_to_transform_lvl1": ([vp, vp, i32, vp], i32),
        "cufhe_cmux_lvl1": ([vp, vp, vp, vp, i32, vp], i32),
        "cufhe_cmux_batch_lvl1": ([vp, i32, vp], i32),
        "tfhe_trgsw_encrypt": ([u64, u64, vp, i32, vp], i32),
        "tfhe_trlwe_encrypt": ([u64, u64, vp, vp, ctypes.c_double, vp], i32),
        "tfhe_trlwe_encrypt_level": ([i32, u64, u64, vp, vp, ctypes.c_double, vp], i32),
        "tfhe_trlwe_phase": ([i32, vp, vp, sz, vp], i32),
        "tfhe_annihilate_key_words": ([i32], sz),
        "tfhe_annihilate_key_gen": ([i32, u64, vp, ctypes.c_double, vp, sz], i32),
        "cufhe_event_create": ([ctypes.POINTER(vp)], i32),
        "cufhe_event_destroy": ([vp], i32),
        "cufhe_event_record": ([vp, vp], i32),
        "cufhe_event_elapsed_ms": ([vp, vp, ctypes.POINTER(ctypes.c_float)], i32),
        "cufhe_annihilate_key_load": ([i32, vp, sz], i32),
        "cufhe_annihilate_key_load_ex": ([i32, vp, sz, i32], i32),
        "cufhe_annihilate_key_delete": ([i32], i32),
        "cufhe_annihilate_key_switch": ([i32, vp, vp, vp], i32),
        "cufhe_annihilate_key_switch_batch": ([i32, vp, sz, vp, sz, sz, vp], i32),
        "cufhe_bk_load": ([i32, vp, sz, i32, i32], i32),
        "cufhe_bk_delete": ([i32], i32),
        "cufhe_ksk_load": ([i32, vp, sz, i32], i32),
        "cufhe_ksk_delete": ([i32], i32),
        "cufhe_circuit_aes128_decrypt": ([vp, vp, vp, vp], i32),
        "cufhe_circuit_ascon_xof": ([i32, i32, vp, vp, vp, vp], i32),
        "cufhe_circuit_ascon_absorb": ([i32, vp, vp, vp, vp], i32),
        "cufhe_circuit_ascon_squeeze": ([i32, vp, vp, vp, vp], i32),
        "cufhe_ascon_xof_initial_state": ([vp], i32),
        "cufhe_server_create": ([sz, i32, i32, ctypes.POINTER(vp)], i32),
        "cufhe_server_submit": ([vp, vp, vp, vp, ctypes.POINTER(u64)], i32),
        "cufhe_server_wait": ([vp, u64, i32], i32),
        "cufhe_server_request_info": ([vp, u64, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(i32)], i32),
        "cufhe_server_stats": ([vp, ctypes.POINTER(ctypes.c_double)], i32),
        "cufhe_server_destroy": ([vp], i32),
        "tfhe_secret_key_gen": ([u64, i32, vp, vp, vp], i32),
        "tfhe_bk_gen": ([i32, i32, u64, vp, vp, vp, sz, ctypes.POINTER(sz)], i32),
        "tfhe_ksk_gen": ([i32, u64, vp, vp, i32, vp, sz], i32),
        "tfhe_tlwe_encrypt_bits": ([i32, u64, vp, vp, sz, vp], i32),
        "tfhe_tlwe_phase": ([i32, vp, vp, sz, vp], i32),
        "tfhe_tlwe_decrypt_bits": ([i32, vp, vp, sz, vp], i32),
        "tfhe_tlwe_words": ([i32], sz),
        "tfhe_bk_words": ([i32, i32], sz),
        "tfhe_keyfile_save": ([ctypes.c_char_p, u32, vp, vp], i32),
        "tfhe_keyfile_info": ([ctypes.c_char_p, vp, vp], i32),
        "tfhe_keyfile_load": ([ctypes.c_char_p, vp], i32),
        "tfhe_ksk_words": ([i32], sz),
        "tfhe_ksk_subset_words": ([], sz),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().cufhe_last_error().decode(errors="replace")
        raise CufheError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------- host layer

def secret_key_gen(seed, block_binary=False):
    k0 = np.zeros(N0, np.uint32)
    k1 = np.zeros(N1, np.uint32)
    k2 = np.zeros(N2, np.uint64)
    check(lib().tfhe_secret_key_gen(seed, int(block_binary), _ptr(k0), _ptr(k1), _ptr(k2)),
          "tfhe_secret_key_gen")
    return k0, k1, k2


# Key files (tfhe_keyfile_*): record names in tag order and element types.
KEYFILE_RECORDS = [("bklvl01", np.uint32), ("bklvl02", np.uint64), ("iksklvl10", np.uint32),
                   ("iksklvl20", np.uint32), ("sklvl0", np.uint32), ("sklvl1", np.uint32),
                   ("sklvl2", np.uint64)]
KEYFILE_FLAGS = [("key_bundle", 1), ("block_binary", 2), ("subset", 4), ("ntt", 8)]


def save_keyfile(path, arrays, key_bundle=True, block_binary=False, subset=False, ntt=False):
    """Write TFHEpp-layout key arrays (names of KEYFILE_RECORDS) to a key file."""
    flags = sum(bit for name, bit in KEYFILE_FLAGS
                if {"key_bundle": key_bundle, "block_binary": block_binary, "subset": subset,
                    "ntt": ntt}[name])
    known = dict(KEYFILE_RECORDS)
    unknown = set(arrays) - set(known)
    if unknown:
        raise CufheError("unknown key records: %s" % sorted(unknown))
    keep = [np.ascontiguousarray(arrays[n], known[n]) if n in arrays else None
            for n, _ in KEYFILE_RECORDS]
    ptrs = (ctypes.c_void_p * len(keep))(*[None if a is None else a.ctypes.data for a in keep])
    counts = (ctypes.c_size_t * len(keep))(*[0 if a is None else a.size for a in keep])
    rc = lib().tfhe_keyfile_save(os.fsencode(path), flags, ptrs, counts)
    if rc != 0:
        raise CufheError("tfhe_keyfile_save failed (%d): bad sizes or I/O" % rc)


def load_keyfile(path):
    """Read a key file: (flags dict, {record name: numpy array})."""
    flags = ctypes.c_uint32(0)
    counts = (ctypes.c_size_t * len(KEYFILE_RECORDS))()
    rc = lib().tfhe_keyfile_info(os.fsencode(path), ctypes.byref(flags), counts)
    if rc != 0:
        raise CufheError("tfhe_keyfile_info failed (%d)" % rc)
    arrays = {n: np.zeros(counts[i], dt) for i, (n, dt) in enumerate(KEYFILE_RECORDS) if counts[i]}
    ptrs = (ctypes.c_void_p * len(KEYFILE_RECORDS))(
        *[arrays[n].ctypes.data if n in arrays else None for n, _ in KEYFILE_RECORDS])
    rc = lib().tfhe_keyfile_load(os.fsencode(path), ptrs)
    if rc != 0:
        raise CufheError("tfhe_keyfile_load failed (%d)" % rc)
    return {name: bool(flags.value & bit) for name, bit in KEYFILE_FLAGS}, arrays


def bk_gen(level, key_bundle, seed, k0, kt):
    words = lib().tfhe_bk_words(level, int(key_bundle))
    out = np.zeros(words, np.uint32 if level == 1 else np.uint64)
    check(lib().tfhe_bk_gen(level, int(key_bundle), seed, _ptr(k0), _ptr(kt), _ptr(out), words,
                            None), "tfhe_bk_gen")
    return out


def ksk_gen(level, seed, kd, k0, subset=False):
    """KeySwitchingKey<lvl10/lvl20>; subset=True (level 1 only) keeps the rows
    j >= n0 of a subset key (the lvl1 key starts with the lvl0 key)."""
    words = lib().tfhe_ksk_subset_words() if subset else lib().tfhe_ksk_words(level)
    out = np.zeros(words, np.uint32)
    check(lib().tfhe_ksk_gen(level, seed, _ptr(kd), _ptr(k0), int(subset), _ptr(out), words),
          "tfhe_ksk_gen")
    return out


BACKEND_FFT, BACKEND_NTT = 0, 1
BR_PLAIN, BR_KEY_BUNDLE, BR_BLOCK_BINARY = 0, 1, 2


def initialize_lvl01(bk, ksk, backend=BACKEND_FFT, variant=BR_KEY_BUNDLE, subset=False):
    """Load lvl1 gate keys (polynomial-domain BK and KSK) on the GPU."""
    bk = np.ascontiguousarray(bk, np.uint32)
    ksk = np.ascontiguousarray(ksk, np.uint32)
    check(lib().cufhe_initialize_lvl01_ex(_ptr(bk), bk.size, backend, variant, _ptr(ksk), ksk.size,
                                          int(subset)), "cufhe_initialize_lvl01_ex")


def initialize_lvl02(bk, ksk, backend=BACKEND_FFT, key_bundle=True):
    """Load lvl02 gate keys (polynomial-domain lvl02 BK and lvl20 KSK) on the GPU."""
    bk = np.ascontiguousarray(bk, np.uint64)
    ksk = np.ascontiguousarray(ksk, np.uint32)
    check(lib().cufhe_initialize_lvl02_ex(_ptr(bk), bk.size, backend, int(key_bundle), _ptr(ksk),
                                          ksk.size), "cufhe_initialize_lvl02_ex")


def tlwe_words(level):
    return {0: TLWE0_WORDS, 1: TLWE1_WORDS, 2: TLWE2_WORDS}[level]


def encrypt_bits(level, seed, key, bits):
    bits = np.ascontiguousarray(bits, np.uint8)
    dt = np.uint64 if level == 2 else np.uint32
    out = np.zeros((len(bits), tlwe_words(level)), dt)
    check(lib().tfhe_tlwe_encrypt_bits(level, seed, _ptr(key), _ptr(bits), len(bits), _ptr(out)),
          "tfhe_tlwe_encrypt_bits")
    return out


def autotune_lvl1(nbootstraps, stream=None):
    """Run the library's rotation autotuner for batches of nbootstraps;
    returns (best jobs_per_group, {c: rotation ms})."""
    best = ctypes.c_int32(0)
    t = (ctypes.c_float * 9)()
    check(lib().cufhe_autotune_lvl1(int(nbootstraps), ctypes.byref(best), t, _sp(stream)),
          "cufhe_autotune_lvl1")
    return best.value, {c: float(t[c]) for c in range(9) if t[c] > 0}


def trgsw_encrypt(seed, key1, m, count=1):
    """count lvl1 TRGSW encryptions of the integers m (array or scalar), each
    4 x 2 x 1024 words, drawn from the streams (seed, 0..count-1)."""
    ms = np.broadcast_to(np.asarray(m, np.int64), (count,))
    out = np.zeros((count, 4 * 2 * N1), np.uint32)
    for i in range(count):
        check(lib().tfhe_trgsw_encrypt(seed, i, _ptr(key1), int(ms[i]), _ptr(out[i])),
              "tfhe_trgsw_encrypt")
    return out


def trlwe_encrypt(seed, key1, polys, alpha=3.42338787018369e-08):
    """lvl1 TRLWE encryptions (a then b, 2 x 1024 words) of the Torus32
    polynomials in the rows of polys."""
    polys = np.ascontiguousarray(np.atleast_2d(polys), np.uint32)
    out = np.zeros((polys.shape[0], 2 * N1), np.uint32)
    for i in range(polys.shape[0]):
        check(lib().tfhe_trlwe_encrypt(seed, i, _ptr(key1), _ptr(polys[i]), alpha, _ptr(out[i])),
              "tfhe_trlwe_encrypt")
    return out


ALPHA = {1: 3.42338787018369e-08, 2: 5.68434188608080e-14}  # the levels' TRLWE noise


def _torus(level):
    return np.uint32 if level == 1 else np.uint64


def trlwe_encrypt_level(level, seed, key, polys, alpha=None):
    """TRLWE encryptions at level 1 or 2 (a then b, 2N words of the level's
    torus type) of the polynomials in the rows of polys."""
    polys = np.ascontiguousarray(np.atleast_2d(polys), _torus(level))
    a = ALPHA[level] if alpha is None else alpha
    out = np.zeros((polys.shape[0], 2 * polys.shape[1]), _torus(level))
    for i in range(polys.shape[0]):
        check(lib().tfhe_trlwe_encrypt_level(level, seed, i, _ptr(key), _ptr(polys[i]), a,
                                             _ptr(out[i])), "tfhe_trlwe_encrypt_level")
    return out


def trlwe_phase(level, key, ct):
    """Phase polynomials b - a*s (torus words) of the TRLWEs in the rows of ct."""
    ct = np.ascontiguousarray(np.atleast_2d(ct), _torus(level))
    out = np.zeros((ct.shape[0], ct.shape[1] // 2), _torus(level))
    check(lib().tfhe_trlwe_phase(level, _ptr(key), _ptr(ct), ct.shape[0], _ptr(out)),
          "tfhe_trlwe_phase")
    return out


def annihilate_key_gen(level, seed, key, alpha=None):
    """Annihilation key of level 1 or 2 in the polynomial domain:
    (nbit, l, 2, N) torus words (half TRGSWs of tau_d(s), d = 2^(bit+1) + 1)."""
    words = lib().tfhe_annihilate_key_words(level)
    if words == 0:
        raise ValueError("annihilation level must be 1 or 2")
    out = np.zeros(words, _torus(level))
    a = ALPHA[level] if alpha is None else alpha
    check(lib().tfhe_annihilate_key_gen(level, seed, _ptr(key), a, _ptr(out), words),
          "tfhe_annihilate_key_gen")
    return out


def annihilate_key_load(level, ahk, backend=BACKEND_FFT):
    ahk = np.ascontiguousarray(ahk, _torus(level))
    check(lib().cufhe_annihilate_key_load_ex(level, _ptr(ahk), ahk.size, backend),
          "cufhe_annihilate_key_load_ex")


def phase(level, key, ct):
    ct = np.ascontiguousarray(ct)
    n = ct.shape[0] if ct.ndim == 2 else 1
    out = np.zeros(n, np.float64)
    check(lib().tfhe_tlwe_phase(level, _ptr(key), _ptr(ct), n, _ptr(out)), "tfhe_tlwe_phase")
    return out


def decrypt_bits(level, key, ct):
    return (phase(level, key, ct) > 0).astype(np.uint8)


# ---------------------------------------------------------------- device layer

class DeviceBuffer:
    """Device allocation owned by Python (hipMalloc through the C ABI)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().cufhe_device_alloc(ctypes.byref(p), max(self.nbytes, 1)),
              "cufhe_device_alloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().cufhe_memcpy_h2d(self.ptr, _ptr(a), a.nbytes, _sp(stream)), "memcpy_h2d")
        check(lib().cufhe_stream_synchronize(_sp(stream)), "stream_synchronize")

    def download(self, dtype, shape, stream=None):
        out = np.empty(shape, dtype)
        check(lib().cufhe_memcpy_d2h(_ptr(out), self.ptr, out.nbytes, _sp(stream)), "memcpy_d2h")
        check(lib().cufhe_stream_synchronize(_sp(stream)), "stream_synchronize")
        return out

    def at(self, byte_offset):
        return ctypes.c_void_p(self.ptr.value + int(byte_offset))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().cufhe_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cufhe_stream_create(ctypes.byref(p)), "cufhe_stream_create")
        self.handle = p

    def query(self):
        rc = lib().cufhe_stream_query(self.handle)
        if rc < 0:
            check(rc, "cufhe_stream_query")
        return rc == 1

    def synchronize(self):
        check(lib().cufhe_stream_synchronize(self.handle), "cufhe_stream_synchronize")

    def destroy(self):
        if self.handle is not None:
            lib().cufhe_stream_destroy(self.handle)
            self.handle = None


class Event:
    """HIP event (timing kernels on a stream)."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cufhe_event_create(ctypes.byref(p)), "cufhe_event_create")
        self.handle = p

    def record(self, stream=None):
        check(lib().cufhe_event_record(self.handle, _sp(stream)), "cufhe_event_record")

    def elapsed_ms(self, end):
        """Milliseconds from this event to `end` (waits for `end`)."""
        ms = ctypes.c_float()
        check(lib().cufhe_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)),
              "cufhe_event_elapsed_ms")
        return ms.value

    def destroy(self):
        if self.handle is not None:
            lib().cufhe_event_destroy(self.handle)
            self.handle = None


def _sp(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Plan:
    """Batched gates over a device pool of lvl1 TLWEs."""

    def __init__(self, gates, pool, jobs_per_group=0, use_graph=False, level=1):
        """level 1: lvl1 wires (IKS-BR); 2: lvl0 wires bootstrapped at lvl2
        (BR-IKS, lvl02 keys); 0: lvl0 wires bootstrapped at lvl1 (BR-IKS,
        lvl1 keys)."""
        arr = (GateDesc * len(gates))()
        for i, g in enumerate(gates):
            op, out, in0, in1, inc = g
            arr[i] = GateDesc(GATES[op] if isinstance(op, str) else int(op), out, in0, in1, inc)
        self._arr = arr
        h = ctypes.c_void_p()
        if level == 2:
            check(lib().cufhe_plan_create_lvl02(ctypes.cast(arr, ctypes.c_void_p), len(gates),
                                                pool.ptr, int(use_graph), ctypes.byref(h)),
                  "cufhe_plan_create_lvl02")
        elif level == 0:
            check(lib().cufhe_plan_create_lvl01(ctypes.cast(arr, ctypes.c_void_p), len(gates),
                                                pool.ptr, int(jobs_per_group), int(use_graph),
                                                ctypes.byref(h)), "cufhe_plan_create_lvl01")
        else:
            check(lib().cufhe_plan_create_lvl1(ctypes.cast(arr, ctypes.c_void_p), len(gates),
                                               pool.ptr, int(jobs_per_group), int(use_graph),
                                               ctypes.byref(h)), "cufhe_plan_create_lvl1")
        self.handle = h

    @property
    def bootstraps(self):
        return lib().cufhe_plan_bootstraps(self.handle)

    @property
    def jobs_per_group(self):
        return lib().cufhe_plan_jobs_per_group(self.handle)

    def run(self, stream=None):
        check(lib().cufhe_plan_run(self.handle, _sp(stream)), "cufhe_plan_run")

    def run_timed(self, stream=None):
        """Eager run with HIP events around each kernel; returns ms of
        (key switching, blind rotation, combine)."""
        ms = (ctypes.c_float * 3)()
        check(lib().cufhe_plan_run_timed(self.handle, _sp(stream), ms), "cufhe_plan_run_timed")
        return float(ms[0]), float(ms[1]), float(ms[2])

    def destroy(self):
        if self.handle is not None:
            lib().cufhe_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """Continuous-batching server of layered lvl1 circuits (cufhe_server_*)."""

    def __init__(self, pool_bytes, max_batch_bootstraps=8192, jobs_per_group=0):
        h = ctypes.c_void_p()
        check(lib().cufhe_server_create(int(pool_bytes), int(max_batch_bootstraps),
                                        int(jobs_per_group), ctypes.byref(h)),
              "cufhe_server_create")
        self.handle = h
        self._outputs = {}

    def submit(self, circuit, inputs):
        """circuit: circuits.Circuit; inputs: (n_inputs, 1025) uint32 TLWEs.
        Returns the request id; outputs via result(id) once done."""
        g = np.zeros((circuit.n_gates, 5), np.uint32)
        g[:, 0] = circuit.ops
        g[:, 2] = circuit.in0
        g[:, 3] = circuit.in1
        g[:, 4] = circuit.inc
        layer_end = np.ascontiguousarray(circuit.layer_end, np.int32)
        outs = np.ascontiguousarray(circuit.outputs, np.uint32)
        inputs = np.ascontiguousarray(inputs, np.uint32)
        assert inputs.shape == (circuit.n_inputs, TLWE1_WORDS)
        out = np.zeros((len(outs), TLWE1_WORDS), np.uint32)
        desc = CircuitDesc(circuit.n_inputs, circuit.n_gates, g.ctypes.data, len(layer_end),
                           layer_end.ctypes.data, len(outs), outs.ctypes.data)
        rid = ctypes.c_uint64()
        check(lib().cufhe_server_submit(self.handle, ctypes.byref(desc), inputs.ctypes.data,
                                        out.ctypes.data, ctypes.byref(rid)), "cufhe_server_submit")
        self._outputs[rid.value] = out
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        rc = lib().cufhe_server_wait(self.handle, rid, int(timeout_ms))
        if rc < 0:
            check(rc, "cufhe_server_wait")
        return rc == 1

    def result(self, rid):
        self.wait(rid)
        return self._outputs.pop(rid)

    def info(self, rid):
        t = (ctypes.c_double * 3)()
        rounds = ctypes.c_int()
        check(lib().cufhe_server_request_info(self.handle, rid, t, ctypes.byref(rounds)),
              "cufhe_server_request_info")
        return {"submit": t[0], "admit": t[1], "done": t[2], "rounds": rounds.value}

    def stats(self):
        st = (ctypes.c_double * 5)()
        check(lib().cufhe_server_stats(self.handle, st), "cufhe_server_stats")
        return {"rounds": int(st[0]), "bootstraps": int(st[1]), "gates": int(st[2]),
                "busy_s": st[3], "peak_pool_bytes": int(st[4])}

    def destroy(self):
        if self.handle is not None:
            rc = lib().cufhe_server_destroy(self.handle)
            self.handle = None
            check(rc, "cufhe_server_destroy")

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
