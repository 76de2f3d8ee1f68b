"""The benchmark's CPU-baseline engine (oracle/cpu_baseline.cpp: key spectra
precomputed once, row-wise key switch) computes the same gates as the
oracle's FFT mode: decryptions equal the truth tables and the ciphertexts
agree coefficient for coefficient up to FP64 rounding (both round the same
real-valued external products; measured equal on these seeds). Key bundle
and plain CMUX keys."""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O


@pytest.mark.parametrize("key_bundle", [True, False])
def test_cpu_baseline_matches_oracle_fft(keys_lvl1, key_bundle):
    bk = keys_lvl1["bk_kb"] if key_bundle else keys_lvl1["bk"]
    ksk, k1 = keys_lvl1["ksk"], keys_lvl1["k1"]
    rng = np.random.default_rng(21 if key_bundle else 22)
    ops = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR", "ANDNY", "ORYN"]
    n = len(ops)
    bits = rng.integers(0, 2, 2 * n).astype(np.uint8)
    ct = C.encrypt_bits(1, 23, k1, bits)
    opc = np.array([C.GATES[o] for o in ops], np.int32)
    eng = O.CpuBaseline(bk, key_bundle)
    try:
        out = eng.gates(opc, ct[:n], ct[n:], ksk, 4)
    finally:
        eng.release()
    want = [C.TRUTH[o](int(bits[i]), int(bits[n + i])) for i, o in enumerate(ops)]
    assert list(C.decrypt_bits(1, k1, out)) == want
    variant = O.BR_KB if key_bundle else O.BR_PLAIN
    ref = O.gate_batch_lvl1(O.FFT, variant, opc, ct[:n], ct[n:], bk, ksk, 4)
    assert np.mean(out == ref) >= 0.999
    d = np.abs(out.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert d.max() <= 16
