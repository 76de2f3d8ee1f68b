"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, and only as the checker or the CPU baseline. The product
library (cufhe_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

EXACT, FFT, NTT = 0, 1, 2
# blind-rotation variant passed as `kb`: plain CMUX, 2-bit key bundle,
# block-binary key (blocks of 4); SUBSET (added to the variant) selects the
# subset key switch in the lvl1 gate functions.
BR_PLAIN, BR_KB, BR_BB, SUBSET = 0, 1, 2, 4
P31 = 2147473409
GOLDILOCKS = 0xFFFFFFFF00000001

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, u32, u64 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32,
                             ctypes.c_uint32, ctypes.c_uint64)
    sigs = {
        "oracle_version": ([], ctypes.c_int),
        "oracle_negacyclic_mod": ([u64, u32, vp, vp, vp, ctypes.c_int], ctypes.c_int),
        "oracle_torus_to_ntt_mod32": ([vp, sz, u64, vp], ctypes.c_int),
        "oracle_torus_to_ntt_mod64": ([vp, sz, u64, vp], ctypes.c_int),
        "oracle_ntt_mod_to_torus32": ([vp, sz, u64, vp], ctypes.c_int),
        "oracle_ntt_mod_to_torus64": ([vp, sz, u64, vp], ctypes.c_int),
        "oracle_ffnt_forward": ([u32, vp, vp], ctypes.c_int),
        "oracle_ffnt_inverse": ([u32, vp, vp], ctypes.c_int),
        "oracle_gadget_digits": ([ctypes.c_int, vp, sz, vp], ctypes.c_int),
        "oracle_cmux_accumulate": ([ctypes.c_int, ctypes.c_int, vp, vp, u32], ctypes.c_int),
        "oracle_external_product": ([ctypes.c_int, ctypes.c_int, vp, vp, vp], ctypes.c_int),
        "oracle_blind_rotate": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, u64, vp],
                                ctypes.c_int),
        "oracle_sample_extract": ([ctypes.c_int, vp, vp], ctypes.c_int),
        "oracle_identity_key_switch": ([ctypes.c_int, ctypes.c_int, vp, vp, i32, i32, u64, vp,
                                        vp], ctypes.c_int),
        "oracle_gate_lvl1": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp],
                             ctypes.c_int),
        "oracle_gate_lvl02": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp],
                              ctypes.c_int),
        "oracle_gate_batch_lvl1": ([ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp,
                                    ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "oracle_gate_batch_lvl1_ex": ([ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp,
                                       ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "oracle_automorphism": ([ctypes.c_int, vp, u32, vp], ctypes.c_int),
        "oracle_annihilate": ([ctypes.c_int, ctypes.c_int, vp, vp, vp, sz], ctypes.c_int),
        "oracle_set_mutation": ([ctypes.c_int], ctypes.c_int),
        "oracle_cpu_baseline_prepare": ([vp, ctypes.c_int], vp),
        "oracle_cpu_baseline_release": ([vp], None),
        "oracle_cpu_baseline_gates": ([vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int],
                                      ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("oracle %s failed (%d)" % (what, rc))


def negacyclic_mod(p, a, b, brute=False):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.zeros_like(a)
    _ok(lib().oracle_negacyclic_mod(p, len(a), _p(a), _p(b), _p(out), int(brute)), "negacyclic")
    return out


def torus_to_ntt_mod(a, p):
    a = np.ascontiguousarray(a)
    out = np.zeros(a.shape, np.uint64)
    f = lib().oracle_torus_to_ntt_mod32 if a.dtype == np.uint32 else lib().oracle_torus_to_ntt_mod64
    _ok(f(_p(a), a.size, p, _p(out)), "torus_to_ntt_mod")
    return out


def ntt_mod_to_torus(v, p, bits):
    v = np.ascontiguousarray(v, np.uint64)
    out = np.zeros(v.shape, np.uint32 if bits == 32 else np.uint64)
    f = lib().oracle_ntt_mod_to_torus32 if bits == 32 else lib().oracle_ntt_mod_to_torus64
    _ok(f(_p(v), v.size, p, _p(out)), "ntt_mod_to_torus")
    return out


def ffnt_forward(poly):
    poly = np.ascontiguousarray(poly, np.float64)
    spec = np.zeros(len(poly), np.float64)
    _ok(lib().oracle_ffnt_forward(len(poly), _p(poly), _p(spec)), "ffnt_forward")
    return spec[0::2] + 1j * spec[1::2]


def ffnt_inverse(spec):
    s = np.empty(2 * len(spec), np.float64)
    s[0::2] = spec.real
    s[1::2] = spec.imag
    out = np.zeros(2 * len(spec), np.float64)
    _ok(lib().oracle_ffnt_inverse(2 * len(spec), _p(s), _p(out)), "ffnt_inverse")
    return out


def gadget_digits(level, coef):
    coef = np.ascontiguousarray(coef, np.uint32 if level == 1 else np.uint64)
    l = 2 if level == 1 else 4
    out = np.zeros(coef.size * l, np.int32)
    _ok(lib().oracle_gadget_digits(level, _p(coef), coef.size, _p(out)), "gadget_digits")
    return out.reshape(coef.size, l)


def cmux_accumulate(level, mode, trlwe, trgsw, abar):
    t = np.array(trlwe, copy=True)
    _ok(lib().oracle_cmux_accumulate(level, mode, _p(t), _p(np.ascontiguousarray(trgsw)), abar),
        "cmux_accumulate")
    return t


def external_product(level, mode, d, trgsw):
    d = np.ascontiguousarray(d)
    out = np.zeros_like(d)
    _ok(lib().oracle_external_product(level, mode, _p(d), _p(np.ascontiguousarray(trgsw)),
                                      _p(out)), "external_product")
    return out


def automorphism(level, poly, d):
    """p(X) -> p(X^d) mod X^N + 1 (d odd) on one polynomial of the level."""
    poly = np.ascontiguousarray(poly, np.uint32 if level == 1 else np.uint64)
    out = np.zeros_like(poly)
    _ok(lib().oracle_automorphism(level, _p(poly), d, _p(out)), "automorphism")
    return out


def annihilate(level, mode, trlwe, ahk):
    """Annihilation key switching of one TRLWE (a then b), or of each row of a
    2-D array of them, with the key's nbit x l x 2 polynomials."""
    trlwe = np.ascontiguousarray(trlwe, np.uint32 if level == 1 else np.uint64)
    out = np.zeros_like(trlwe)
    n = trlwe.shape[0] if trlwe.ndim == 2 else 1
    _ok(lib().oracle_annihilate(level, mode, _p(trlwe), _p(np.ascontiguousarray(ahk)), _p(out), n),
        "annihilate")
    return out


def blind_rotate(level, mode, kb, tl0, bk, mu):
    n = 1024 if level == 1 else 2048
    out = np.zeros(2 * n, np.uint32 if level == 1 else np.uint64)
    _ok(lib().oracle_blind_rotate(level, mode, int(kb), _p(np.ascontiguousarray(tl0, np.uint32)),
                                  _p(bk), mu, _p(out)), "blind_rotate")
    return out


def sample_extract(level, trlwe):
    n = 1024 if level == 1 else 2048
    out = np.zeros(n + 1, trlwe.dtype)
    _ok(lib().oracle_sample_extract(level, _p(np.ascontiguousarray(trlwe)), _p(out)),
        "sample_extract")
    return out


def identity_key_switch(level, in0, in1, ca, cb, offset, ksk, subset=False):
    out = np.zeros(637, np.uint32)
    _ok(lib().oracle_identity_key_switch(level, int(subset), _p(np.ascontiguousarray(in0)),
                                         _p(None if in1 is None else np.ascontiguousarray(in1)),
                                         ca, cb, offset, _p(ksk), _p(out)), "identity_key_switch")
    return out


def gate_lvl1(mode, kb, op, in0, in1, inc, bk, ksk):
    out = np.zeros(1025, np.uint32)
    _ok(lib().oracle_gate_lvl1(mode, int(kb), op, _p(in0), _p(in1), _p(inc), _p(bk), _p(ksk),
                               _p(out)), "gate_lvl1")
    return out


def gate_lvl02(mode, kb, op, in0, in1, inc, bk, ksk):
    out = np.zeros(637, np.uint32)
    _ok(lib().oracle_gate_lvl02(mode, int(kb), op, _p(in0), _p(in1), _p(inc), _p(bk), _p(ksk),
                                _p(out)), "gate_lvl02")
    return out


def gate_batch_lvl1(mode, kb, ops, in0, in1, bk, ksk, nthreads, inc=None):
    """Rows of lvl1 gates on nthreads host threads (inc: MUX selectors)."""
    ops = np.ascontiguousarray(ops, np.int32)
    in0 = np.ascontiguousarray(in0, np.uint32)
    in1 = np.ascontiguousarray(in1, np.uint32)
    inc = None if inc is None else np.ascontiguousarray(inc, np.uint32)
    out = np.zeros_like(in0)
    _ok(lib().oracle_gate_batch_lvl1_ex(mode, int(kb), _p(ops), _p(in0), _p(in1), _p(inc), _p(bk),
                                        _p(ksk), _p(out), len(ops), nthreads), "gate_batch_lvl1")
    return out


class CpuBaseline:
    """The benchmark's CPU-baseline engine (cpu_baseline.cpp): lvl1 binary
    gates with the bootstrapping key kept in the Fourier domain, one gate per
    host thread. Not TFHEpp; this project's own CPU implementation."""

    def __init__(self, bk, key_bundle):
        self._bk = np.ascontiguousarray(bk, np.uint32)
        self.handle = lib().oracle_cpu_baseline_prepare(_p(self._bk), 1 if key_bundle else 0)
        if not self.handle:
            raise RuntimeError("oracle_cpu_baseline_prepare failed")

    def gates(self, ops, in0, in1, ksk, nthreads):
        ops = np.ascontiguousarray(ops, np.int32)
        in0 = np.ascontiguousarray(in0, np.uint32)
        in1 = np.ascontiguousarray(in1, np.uint32)
        ksk = np.ascontiguousarray(ksk, np.uint32)
        out = np.zeros_like(in0)
        _ok(lib().oracle_cpu_baseline_gates(self.handle, _p(ops), _p(in0), _p(in1), _p(ksk),
                                            _p(out), len(ops), nthreads), "cpu_baseline_gates")
        return out

    def release(self):
        if self.handle:
            lib().oracle_cpu_baseline_release(self.handle)
            self.handle = None


MUTATIONS = {
    "ntt_torus_offset": 1, "bar_off_by_one": 2, "b_rounded": 3, "bb_sign": 4,
    "bb_not_halved": 5, "ks_no_round": 6, "kb_order": 7, "a_truncated": 8,
    "auto_sign": 9, "ah_no_halve": 10, "ah_key_row": 11,
}


class mutation:
    """Context manager planting one known mistake in the oracle (tests only)."""

    def __init__(self, name):
        self.code = MUTATIONS[name]

    def __enter__(self):
        _ok(lib().oracle_set_mutation(self.code), "set_mutation")
        return self

    def __exit__(self, *exc):
        _ok(lib().oracle_set_mutation(0), "set_mutation")
        return False
