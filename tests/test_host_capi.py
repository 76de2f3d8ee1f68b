"""Host layer (keygen/encryption) and the C ABI surface. CPU only: no GPU
compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import cufhe_amd as C
from conftest import ROOT, signed32


def _declared_functions():
    names = set()
    for h in ("cufhe_amd.h", "tfhe_host.h", "cufhe_circuits.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:cufhe|tfhe)_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = _declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_paths_without_gpu_work():
    lib = C.lib()
    assert b"gfx950" in lib.cufhe_version()
    # more than one GPU per process is refused with a clear error
    assert lib.cufhe_set_gpu_num(2) == -5
    assert b"one process" in lib.cufhe_last_error()
    assert lib.cufhe_set_gpu_num(1) == 0
    # gates before keys are loaded are refused, not run
    assert lib.cufhe_gate_lvl1(0, None, None, None, None, None) == -3
    assert lib.cufhe_plan_create_lvl1(None, 0, None, 0, 0, ctypes.byref(ctypes.c_void_p())) == -3


def test_separate_key_loading_argument_errors_without_gpu_work():
    """cufhe_bk_load / cufhe_ksk_load / *_delete (BootstrappingKeyToNTT,
    KeySwitchingKeyToDevice ...) refuse bad arguments before touching a GPU."""
    lib = C.lib()
    buf = np.zeros(16, np.uint64)
    p = buf.ctypes.data
    assert lib.cufhe_bk_load(3, p, 16, 0, 1) == -1            # level
    assert lib.cufhe_bk_load(1, None, 16, 0, 1) == -1         # null key
    assert lib.cufhe_bk_load(1, p, 16, 2, 1) == -1            # backend
    assert lib.cufhe_bk_load(1, p, 16, 0, 3) == -1            # variant
    assert lib.cufhe_bk_load(1, p, 16, 0, 2) == -5            # block-binary needs NTT
    assert lib.cufhe_bk_load(2, p, 16, 1, 2) == -5            # ... and lvl1
    assert lib.cufhe_ksk_load(0, p, 16, 0) == -1
    assert lib.cufhe_ksk_load(1, None, 16, 0) == -1
    assert lib.cufhe_ksk_load(2, p, 16, 1) == -5              # subset key is lvl10 only
    assert lib.cufhe_bk_delete(0) == -1 and lib.cufhe_ksk_delete(5) == -1


def test_annihilation_argument_errors_without_gpu_work():
    """Checked before any device call: bad level, null pointers, strides below
    the TRLWE size, an NTT key at level 1, a key of the wrong size (host
    generation), an empty batch."""
    lib = C.lib()
    buf = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    assert lib.cufhe_annihilate_key_switch(3, buf, buf, None) == -1
    assert lib.cufhe_annihilate_key_switch_batch(1, None, 2048, buf, 2048, 1, None) == -1
    assert lib.cufhe_annihilate_key_switch_batch(1, buf, 2047, buf, 2048, 1, None) == -1
    assert lib.cufhe_annihilate_key_switch_batch(2, buf, 4096, buf, 4095, 1, None) == -1
    assert lib.cufhe_annihilate_key_switch_batch(1, buf, 2048, buf, 2048, 0, None) == 0
    assert lib.cufhe_annihilate_key_load_ex(1, buf, 40960, 1) == -5
    assert lib.cufhe_annihilate_key_load_ex(2, buf, 180224, 7) == -1
    assert lib.cufhe_annihilate_key_delete(0) == -1
    assert lib.tfhe_annihilate_key_words(1) == 10 * 2 * 2 * 1024
    assert lib.tfhe_annihilate_key_words(2) == 11 * 4 * 2 * 2048
    key = np.zeros(1024, np.uint32)
    out = np.zeros(100, np.uint32)
    assert lib.tfhe_annihilate_key_gen(1, 1, key.ctypes.data, 0.0, out.ctypes.data, 100) == -2
    assert lib.tfhe_annihilate_key_gen(0, 1, key.ctypes.data, 0.0, out.ctypes.data, 100) == -1


def test_parameter_sizes():
    lib = C.lib()
    assert lib.tfhe_tlwe_words(0) == 637
    assert lib.tfhe_tlwe_words(1) == 1025
    assert lib.tfhe_tlwe_words(2) == 2049
    # BK spectra bytes quoted in the scope table: 41,680,896 (plain) and
    # 62,521,344 (KB) at lvl1 = polys x 512 bins x 16 B
    assert lib.tfhe_bk_words(1, 0) // 1024 * 512 * 16 == 41680896
    assert lib.tfhe_bk_words(1, 1) // 1024 * 512 * 16 == 62521344
    assert lib.tfhe_bk_words(2, 0) // 2048 * 1024 * 16 == 166723584
    assert lib.tfhe_bk_words(2, 1) // 2048 * 1024 * 16 == 250085376
    assert lib.tfhe_ksk_words(1) == 1024 * 7 * 2 * 637


def test_keygen_is_deterministic_and_binary():
    a = C.secret_key_gen(7)
    b = C.secret_key_gen(7)
    c = C.secret_key_gen(8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    for k in a:
        assert set(np.unique(k)) <= {0, 1}
    bb = C.secret_key_gen(7, block_binary=True)
    blocks = bb[0].reshape(-1, 4)
    assert blocks.sum(axis=1).max() <= 1          # reference test_block_binary_gpu.cc:86-114
    assert np.array_equal(bb[1][:636], bb[0])     # lvl0 key is a prefix of the lvl1 key


@pytest.mark.parametrize("level", [0, 1, 2])
def test_encrypt_decrypt_roundtrip_and_noise(level):
    k = C.secret_key_gen(3)[level]
    rng = np.random.default_rng(level)
    bits = rng.integers(0, 2, 400).astype(np.uint8)
    ct = C.encrypt_bits(level, 1234, k, bits)
    assert np.array_equal(C.decrypt_bits(level, k, ct), bits)
    ph = C.phase(level, k, ct)
    err = ph - np.where(bits == 1, 0.125, -0.125)
    alpha = {0: 0.0000925119974676756, 1: 0.0000000342338787018369, 2: 2.0 ** -44}[level]
    assert abs(np.std(err) / alpha - 1) < 0.25


def test_bootstrapping_key_rows_encrypt_gadget_of_key_bits(keys_lvl1):
    # TRGSW(m) row r decrypts (under the lvl1 key) to m * h_{r mod l} on
    # polynomial r // l (TFHEpp trgswSymEncrypt layout, src/bootstrap_gpu.cu:70-75)
    k0, k1, bk = keys_lvl1["k0"], keys_lvl1["k1"], keys_lvl1["bk"]
    n = 1024
    for i in [0, 1, 5, 635]:
        t = bk[i * 8 * n:(i + 1) * 8 * n].reshape(4, 2, n)
        for r in range(4):
            a, b = t[r, 0].astype(np.int64), t[r, 1].astype(np.int64)
            # phase = b - a*s (negacyclic)
            s = k1.astype(np.int64)
            full = np.convolve(a, s)
            prod = full[:n].copy()
            prod[: n - 1] -= full[n:]
            ph = (b - prod) % 2**32
            want = np.zeros(n, np.int64)
            if r // 2 == 1:
                want[0] = int(k0[i]) << (32 - (r % 2 + 1) * 8)
            d = signed32(((ph - want) % 2**32).astype(np.uint32))
            # row r // l == 0 carries m*h on the mask polynomial: phase picks up -m*h*s
            if r // 2 == 0:
                d = signed32(((ph + (int(k0[i]) << (32 - (r % 2 + 1) * 8)) * s) % 2**32)
                             .astype(np.uint32))
            assert np.max(np.abs(d)) < 2**12, (i, r)


def test_key_bundle_layout(keys_lvl1):
    # [3i+0] = TRGSW(s0 s1), [3i+1] = TRGSW(s0 (1-s1)), [3i+2] = TRGSW((1-s0) s1)
    k0, k1, bk = keys_lvl1["k0"], keys_lvl1["k1"], keys_lvl1["bk_kb"]
    n = 1024
    s = k1.astype(np.int64)
    for g in [0, 7, 317]:
        s0, s1 = int(k0[2 * g]), int(k0[2 * g + 1])
        for which, m in enumerate([s0 * s1, s0 * (1 - s1), (1 - s0) * s1]):
            t = bk[(3 * g + which) * 8 * n:(3 * g + which + 1) * 8 * n].reshape(4, 2, n)
            a, b = t[2, 0].astype(np.int64), t[2, 1].astype(np.int64)  # row j=1, digit 0
            full = np.convolve(a, s)
            prod = full[:n].copy()
            prod[: n - 1] -= full[n:]
            ph = signed32(((b - prod) % 2**32).astype(np.uint32))
            assert abs(ph[0] - m * (1 << 24)) < 2**12
