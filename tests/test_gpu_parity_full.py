"""Parity of the HIP kernels at the benchmark configurations' full sizes.

* Noiseless keys (tests/pins.py): every blind rotation's output is a closed
  form of the inputs, computed with Python integers; the headline batched
  kernel (4096 key-bundle rotations, the one-wave-per-job kernel) and the
  NTT block-binary rotation are compared with it on all 4096 jobs.
* Config 4 (NTT backend, block-binary keys, subset key switching, 4096
  NAND/XOR gates): every gate bit-exact against the oracle's NTT mode.
* Config 2 (4096 key-bundle gates through a HIP graph, default kernel):
  every gate decrypts; 64 gates equal the exact-integer oracle.
* lvl2 -> lvl0 identity key switch: bit-exact against the oracle on GPU
  rotation outputs and on fresh encryptions (reference
  test/test_keyswitch_diag.cu:31-127), closed-form phase with a noiseless key.
* A plan refuses to run after its keys were reloaded.
"""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O
import pins as Q
from conftest import load_lvl1_keys, signed32

pytestmark = pytest.mark.gpu

W0, W1, W2 = C.TLWE0_WORDS, C.TLWE1_WORDS, C.TLWE2_WORDS
NTT_TOL = 1 << 18


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "cufhe_initialize")
    yield C.lib()


def _bb_key(rng):
    s = np.zeros(Q.N0, np.int64)
    for blk in range(Q.N0 // Q.ELL):
        pos = rng.integers(0, Q.ELL + 1)
        if pos < Q.ELL:
            s[blk * Q.ELL + pos] = 1
    return s


def _gpu_rotations(tl0):
    """Batched rotations of lvl0 samples into TRLWEs (default kernel)."""
    n = tl0.shape[0]
    din = C.DeviceBuffer.from_numpy(np.ascontiguousarray(tl0, np.uint32))
    dout = C.DeviceBuffer(n * 2 * Q.N1 * 4)
    C.check(C.lib().cufhe_bootstrap_to_trlwe_lvl01(dout.ptr, din.ptr, Q.MU1, n, None),
            "bootstrap_to_trlwe")
    return dout.download(np.uint32, (n, 2 * Q.N1))


def _closed_forms(tl0, s0, variant):
    return np.stack([Q.rotated_test_vector(Q.rotation_amount(t, s0, variant)) for t in tl0])


@pytest.mark.parametrize("variant,backend", [("kb", C.BACKEND_FFT), ("bb", C.BACKEND_NTT),
                                             ("plain", C.BACKEND_FFT)])
def test_noiseless_rotations_4096_are_the_closed_form(engine, variant, backend):
    rng = np.random.default_rng(4000 + ["kb", "bb", "plain"].index(variant))
    s0 = _bb_key(rng) if variant == "bb" else rng.integers(0, 2, Q.N0)
    bk = Q.noiseless_bk(s0, variant)
    s1 = rng.integers(0, 2, Q.N1)
    if variant == "bb":
        s1[:Q.N0] = s0
    ksk = Q.noiseless_ksk(s1, s0, rng, subset=variant == "bb")
    vmap = {"kb": C.BR_KEY_BUNDLE, "bb": C.BR_BLOCK_BINARY, "plain": C.BR_PLAIN}
    load_lvl1_keys(("noiseless", variant), bk, ksk, backend=backend, variant=vmap[variant],
                   subset=variant == "bb")
    n = 4096 if variant != "plain" else 512
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    print("noiseless %s: %d rotations" % (variant, n), flush=True)
    got = _gpu_rotations(tl0)
    want = _closed_forms(tl0, s0, variant)
    diff = np.abs(signed32(got - want))
    if backend == C.BACKEND_NTT:
        assert diff.max() < NTT_TOL, int(diff.max())
        # and bit-exact to the oracle's NTT mode on a sample
        for i in rng.choice(n, 6, replace=False):
            ref = O.blind_rotate(1, O.NTT, O.BR_BB, tl0[i], bk, Q.MU1)
            assert np.array_equal(got[i], ref), int(i)
    else:
        bad = np.nonzero(diff.max(axis=1))[0]
        assert bad.size == 0, (bad[:8], int(diff.max()))


def test_config4_ntt_block_binary_4096_all_bit_exact(engine, keys_bb):
    """NTT backend, block-binary keys, subset key switching,
    4096 lvl1 NAND/XOR gates from a seed, one batched plan replayed from a HIP
    graph: every one of the 4096 output ciphertexts equals the oracle's NTT
    mode word for word."""
    bk, ksk, k1 = keys_bb["bk"], keys_bb["ksk_subset"], keys_bb["k1"]
    load_lvl1_keys("bb", bk, ksk, backend=C.BACKEND_NTT, variant=C.BR_BLOCK_BINARY, subset=True)
    rng = np.random.default_rng(44096)
    n = 4096
    ops = np.where(rng.integers(0, 2, n) == 0, "NAND", "XOR")
    bits = rng.integers(0, 2, 2 * n).astype(np.uint8)
    ct = C.encrypt_bits(1, 4444, k1, bits)
    pool = np.zeros((3 * n, W1), np.uint32)
    pool[: 2 * n] = ct
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([(str(ops[i]), 2 * n + i, i, n + i, 0) for i in range(n)], dpool, use_graph=True)
    for _ in range(2):
        plan.run()
        C.check(C.lib().cufhe_synchronize(), "sync")
    out = dpool.download(np.uint32, pool.shape)[2 * n:]
    plan.destroy()
    want = np.array([C.TRUTH[str(o)](int(bits[i]), int(bits[n + i])) for i, o in enumerate(ops)],
                    np.uint8)
    assert int(np.sum(C.decrypt_bits(1, k1, out) != want)) == 0
    print("config 4: oracle NTT mode on 4096 gates", flush=True)
    codes = np.array([C.GATES[str(o)] for o in ops], np.int32)
    ref = O.gate_batch_lvl1(O.NTT, O.BR_BB | O.SUBSET, codes, ct[:n], ct[n:], bk, ksk, 16)
    mismatched = np.nonzero(np.any(out != ref, axis=1))[0]
    assert mismatched.size == 0, mismatched[:8]


def test_config2_kb_4096_graph_vs_exact_oracle(engine, keys_lvl1):
    """4096 key-bundle lvl1 gates, op uniform over
    NAND/AND/OR/XOR/XNOR/NOR, one HIP graph with the default (headline)
    kernel, replayed twice: every gate decrypts right, and 64 seeded gates
    equal the exact-integer oracle (>= 99.9% of coefficients, |phase
    difference| < 2^-20 of the torus)."""
    k1 = keys_lvl1["k1"]
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"], backend=C.BACKEND_FFT,
                   variant=C.BR_KEY_BUNDLE)
    rng = np.random.default_rng(24096)
    n = 4096
    names = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR"]
    ops = [names[i] for i in rng.integers(0, 6, n)]
    bits = rng.integers(0, 2, 2 * n).astype(np.uint8)
    ct = C.encrypt_bits(1, 2222, k1, bits)
    pool = np.zeros((3 * n, W1), np.uint32)
    pool[: 2 * n] = ct
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([(ops[i], 2 * n + i, i, n + i, 0) for i in range(n)], dpool, use_graph=True)
    assert plan.jobs_per_group == 8
    outs = []
    for _ in range(2):
        plan.run()
        C.check(C.lib().cufhe_synchronize(), "sync")
        outs.append(dpool.download(np.uint32, pool.shape)[2 * n:])
    plan.destroy()
    assert np.array_equal(outs[0], outs[1])
    out = outs[1]
    want = np.array([C.TRUTH[o](int(bits[i]), int(bits[n + i])) for i, o in enumerate(ops)],
                    np.uint8)
    assert int(np.sum(C.decrypt_bits(1, k1, out) != want)) == 0
    sample = rng.choice(n, 64, replace=False)
    codes = np.array([C.GATES[ops[i]] for i in sample], np.int32)
    print("config 2: exact oracle on 64 gates", flush=True)
    ref = O.gate_batch_lvl1(O.EXACT, O.BR_KB, codes, ct[sample], ct[n + sample],
                            keys_lvl1["bk_kb"], keys_lvl1["ksk"], 16)
    eq = float(np.mean(out[sample] == ref))
    assert eq >= 0.999, eq
    dph = np.abs(C.phase(1, k1, out[sample]) - C.phase(1, k1, ref))
    dph = np.minimum(dph, 1 - dph)
    assert dph.max() < 2.0 ** -20, dph.max()


def test_lvl20_identity_key_switch_bit_exact(engine, keys_lvl2):
    """KeySwitchFromTLWE<lvl20param> (reference test/test_keyswitch_diag.cu
    :31-127): fresh lvl2 encryptions and lvl2 blind-rotation outputs of the
    GPU, switched on the GPU, equal the oracle word for word and decrypt."""
    k0, k2 = keys_lvl2["k0"], keys_lvl2["k2"]
    C.check(C.lib().cufhe_initialize_lvl02(C._ptr(keys_lvl2["bk_kb"]), keys_lvl2["bk_kb"].size, 1,
                                           C._ptr(keys_lvl2["ksk"]), keys_lvl2["ksk"].size),
            "initialize_lvl02")
    bits = (np.arange(20) % 2).astype(np.uint8)
    fresh = C.encrypt_bits(2, 2020, k2, bits)
    # rotation outputs: lvl0 samples of +-mu0 phases bootstrapped to lvl2
    t0 = C.encrypt_bits(0, 2021, k0, bits[:8])
    din = C.DeviceBuffer.from_numpy(t0)
    drot = C.DeviceBuffer(8 * W2 * 8)
    C.check(C.lib().cufhe_bootstrap_lvl02(drot.ptr, din.ptr, C.MU2, 8, None), "bootstrap_lvl02")
    rot = drot.download(np.uint64, (8, W2))
    inp = np.concatenate([fresh, rot])
    dks_in = C.DeviceBuffer.from_numpy(inp)
    dks_out = C.DeviceBuffer(len(inp) * W0 * 4)
    C.check(C.lib().cufhe_identity_key_switch_lvl20(dks_out.ptr, dks_in.ptr, len(inp),
                                                    None), "identity_key_switch_lvl20")
    got = dks_out.download(np.uint32, (len(inp), W0))
    for i in range(len(inp)):
        ref = O.identity_key_switch(2, inp[i], None, 1, 0, 0, keys_lvl2["ksk"])
        assert np.array_equal(got[i], ref), i
    dec = C.decrypt_bits(0, k0, got)
    assert np.array_equal(dec[:20], bits)
    assert np.array_equal(dec[20:], bits[:8])


def test_lvl20_noiseless_key_switch_closed_form(engine, keys_lvl2):
    """With a noiseless lvl20 key-switching key the GPU key switch's phase is
    the rounding closed form (tests/pins.py) for every sample."""
    rng = np.random.default_rng(21)
    s2 = rng.integers(0, 2, Q.N2)
    s0 = rng.integers(0, 2, Q.N0)
    ksk = Q.noiseless_ksk(s2, s0, rng)
    C.check(C.lib().cufhe_initialize_lvl02(C._ptr(keys_lvl2["bk_kb"]), keys_lvl2["bk_kb"].size, 1,
                                           C._ptr(ksk), ksk.size), "initialize_lvl02")
    n = 64
    inp = rng.integers(0, 2**64 - 1, (n, W2), dtype=np.uint64)
    din = C.DeviceBuffer.from_numpy(inp)
    dout = C.DeviceBuffer(n * W0 * 4)
    C.check(C.lib().cufhe_identity_key_switch_lvl20(dout.ptr, din.ptr, n, None), "ks20")
    got = dout.download(np.uint32, (n, W0))
    for i in range(n):
        assert Q.lwe_phase32(got[i], s0) == Q.ks_phase_closed_form(inp[i], s2, 64), i
    # restore the real keys for later tests
    C.check(C.lib().cufhe_initialize_lvl02(C._ptr(keys_lvl2["bk_kb"]), keys_lvl2["bk_kb"].size, 1,
                                           C._ptr(keys_lvl2["ksk"]), keys_lvl2["ksk"].size),
            "initialize_lvl02")


def test_plan_refuses_to_run_after_a_key_reload(engine, keys_lvl1):
    """A plan is bound to the keys it was created with: after a reload
    cufhe_plan_run returns CUFHE_ERR_STATE before launching anything."""
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"], backend=C.BACKEND_FFT,
                   variant=C.BR_KEY_BUNDLE)
    bits = np.array([1, 0], np.uint8)
    pool = np.zeros((3, W1), np.uint32)
    pool[:2] = C.encrypt_bits(1, 5, keys_lvl1["k1"], bits)
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([("NAND", 2, 0, 1, 0)], dpool, use_graph=True)
    plan.run()
    C.check(C.lib().cufhe_synchronize(), "sync")
    load_lvl1_keys(None, keys_lvl1["bk_kb"], keys_lvl1["ksk"], backend=C.BACKEND_FFT,
                   variant=C.BR_KEY_BUNDLE)  # the same keys, loaded again
    assert C.lib().cufhe_plan_run(plan.handle, None) == -3
    assert b"reloaded" in C.lib().cufhe_last_error()
    plan.destroy()
    plan = C.Plan([("NAND", 2, 0, 1, 0)], dpool)
    plan.run()
    C.check(C.lib().cufhe_synchronize(), "sync")
    out = dpool.download(np.uint32, pool.shape)[2:]
    assert list(C.decrypt_bits(1, keys_lvl1["k1"], out)) == [1]
    plan.destroy()
