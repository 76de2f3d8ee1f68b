"""Output noise of batched lvl1 gates against the variance the parameter set
predicts (cufhe_amd/noise.py: the TFHE external-product analysis summed over
the rotation's steps). 4096 gate outputs per case: the sample variance is
within 2.2% (one standard error) of the true one, so a ±10% band separates a
correct rotation from one that adds a key-noise term or loses digit
precision (both move the variance by far more)."""
import numpy as np
import pytest

import cufhe_amd as C
from conftest import load_lvl1_keys
from cufhe_amd.noise import predicted_lvl1_variance

pytestmark = pytest.mark.gpu

B = 4096


def _gate_noise(k1, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, 2 * B).astype(np.uint8)
    W = C.TLWE1_WORDS
    pool = np.zeros((3 * B, W), np.uint32)
    pool[:2 * B] = C.encrypt_bits(1, 9000 + seed, k1, bits)
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([("NAND", 2 * B + i, i, B + i, 0) for i in range(B)], dpool)
    plan.run()
    C.check(C.lib().cufhe_synchronize(), "sync")
    plan.destroy()
    out = dpool.download(np.uint32, pool.shape)[2 * B:]
    want = 1 - bits[:B] * bits[B:]
    ph = C.phase(1, k1, out)
    assert np.array_equal((ph > 0).astype(np.uint8), want)
    err = ph - np.where(want == 1, 0.125, -0.125)
    return float(np.var(err)), float(np.mean(err))


@pytest.mark.parametrize("variant", ["kb", "plain", "bb"])
def test_output_variance_matches_prediction(keys_lvl1, keys_bb, variant):
    C.check(C.lib().cufhe_initialize(), "initialize")
    if variant == "bb":
        keys = keys_bb
        load_lvl1_keys(("ntt", "bb"), keys["bk"], keys["ksk_subset"], backend=C.BACKEND_NTT,
                       variant=C.BR_BLOCK_BINARY, subset=True)
    elif variant == "kb":
        keys = keys_lvl1
        load_lvl1_keys(("fft", True), keys["bk_kb"], keys["ksk"], backend=C.BACKEND_FFT,
                       variant=C.BR_KEY_BUNDLE)
    else:
        keys = keys_lvl1
        load_lvl1_keys(("fft", False), keys["bk"], keys["ksk"], backend=C.BACKEND_FFT,
                       variant=C.BR_PLAIN)
    var, mean = _gate_noise(keys["k1"], 61)
    want = predicted_lvl1_variance(variant)
    print(variant, "measured variance %.4g (mean %.3g), predicted %.4g, ratio %.3f"
          % (var, mean, want, var / want), flush=True)
    assert 0.9 < var / want < 1.1
    assert abs(mean) < 4 * (want / B) ** 0.5
