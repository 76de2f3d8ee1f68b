"""The library's rotation autotuner (cufhe_autotune_lvl1): it times every
valid jobs-per-workgroup choice for the loaded keys, plans created with
jobs_per_group 0 use its pick for their batch size, every choice computes
the same ciphertexts, and a key load clears the record."""
import numpy as np
import pytest

import conftest
import cufhe_amd as C
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_valid_kernel_and_plans_use_it(keys_lvl1):
    C.check(C.lib().cufhe_initialize(), "initialize")
    conftest._LOADED_LVL1["tag"] = None
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    n = 600
    best, times = C.autotune_lvl1(n)
    assert best in times and set(times) <= {1, 2, 4, 6, 8} and len(times) >= 4
    assert times[best] == min(times.values())
    rng = np.random.default_rng(31)
    ops = ["NAND", "XOR", "AND", "OR", "XNOR", "NOR"]
    names = [ops[i] for i in rng.integers(0, len(ops), n)]
    bits = rng.integers(0, 2, 2 * n).astype(np.uint8)
    pool = np.zeros((3 * n, C.TLWE1_WORDS), np.uint32)
    pool[: 2 * n] = C.encrypt_bits(1, 32, keys_lvl1["k1"], bits)
    gates = [(names[i], 2 * n + i, i, n + i, 0) for i in range(n)]
    outs = {}
    for c in (0, 1):
        d = C.DeviceBuffer.from_numpy(pool)
        plan = C.Plan(gates, d, jobs_per_group=c)
        if c == 0:
            assert plan.jobs_per_group == best
        plan.run()
        C.check(C.lib().cufhe_synchronize(), "sync")
        outs[c] = d.download(np.uint32, pool.shape)[2 * n:]
        plan.destroy()
    assert np.array_equal(outs[0], outs[1])
    want = [C.TRUTH[o](int(bits[i]), int(bits[n + i])) for i, o in enumerate(names)]
    assert list(C.decrypt_bits(1, keys_lvl1["k1"], outs[0])) == want
    # a key load clears the record: the built-in choice (8, one wave per job)
    conftest._LOADED_LVL1["tag"] = None
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    d = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan(gates, d, jobs_per_group=0)
    assert plan.jobs_per_group == 8
    plan.destroy()
