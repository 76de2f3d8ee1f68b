"""The batched key switch as an int8 MFMA product (every batched plan) is
bit-identical to the gather form used by the plain per-gate kernels: full
lvl10 KSK (FFT key bundle), subset KSK (NTT block-binary) and lvl20 (lvl02
gates), eagerly and as HIP-graph replays interleaved with eager work, with
the digit dimension unsplit (plain read-add-write epilogue) and split over
waves (atomic epilogue), at job counts that are not multiples of the
32-job MFMA block."""
import numpy as np
import pytest

import cufhe_amd as C
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu

W0, W1 = C.TLWE0_WORDS, C.TLWE1_WORDS
N = 320  # several GEMM row tiles
OPS = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR", "ANDNY", "ORYN", "MUX", "NMUX"]


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "initialize")


def _gates(rng, n):
    return [OPS[i] for i in rng.integers(0, len(OPS), n)]


def _plain(ops, pool_ct, n, words, gate_fn):
    dp = C.DeviceBuffer.from_numpy(pool_ct)
    do = C.DeviceBuffer(n * words * 4)
    st = C.Stream()
    B = words * 4
    for i, op in enumerate(ops):
        C.check(gate_fn(C.GATES[op], do.at(i * B), dp.at(i * B), dp.at((n + i) * B),
                        dp.at((2 * n + i) * B), st.handle), "gate")
    st.synchronize()
    st.destroy()
    return do.download(np.uint32, (n, words))


def _batched(ops, pool_ct, n, words, level, c=0, graph=False, replays=1):
    pool = np.zeros((4 * n, words), np.uint32)
    pool[: 3 * n] = pool_ct
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([(ops[i], 3 * n + i, i, n + i, 2 * n + i) for i in range(n)], dpool,
                  jobs_per_group=c, use_graph=graph, level=level)
    outs = []
    for _ in range(replays):
        plan.run()
        C.check(C.lib().cufhe_synchronize(), "sync")
        outs.append(dpool.download(np.uint32, pool.shape)[3 * n:])
        # eager work between replays (a plain gate on the same pool)
        C.check(C.lib().cufhe_synchronize(), "sync")
    plan.destroy()
    return outs


@pytest.mark.parametrize("name", ["fft_kb", "ntt_bb"])
def test_lvl1_gemm_key_switch_equals_gather(engine, keys_lvl1, keys_bb, name):
    if name == "fft_kb":
        load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                       backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
        k1 = keys_lvl1["k1"]
    else:
        load_lvl1_keys(("ntt", "bb"), keys_bb["bk"], keys_bb["ksk_subset"],
                       backend=C.BACKEND_NTT, variant=C.BR_BLOCK_BINARY, subset=True)
        k1 = keys_bb["k1"]
    rng = np.random.default_rng(90)
    ops = _gates(rng, N)
    bits = rng.integers(0, 2, 3 * N).astype(np.uint8)
    pool_ct = C.encrypt_bits(1, 91, k1, bits)
    plain = _plain(ops, pool_ct, N, W1, C.lib().cufhe_gate_lvl1)
    eager = _batched(ops, pool_ct, N, W1, 1)[0]
    assert np.array_equal(eager, plain)
    for out in _batched(ops, pool_ct, N, W1, 1, graph=True, replays=3):
        assert np.array_equal(out, plain)


def test_lvl20_gemm_key_switch_equals_gather(engine, keys_lvl2):
    bk, ksk = keys_lvl2["bk_kb"], keys_lvl2["ksk"]
    C.check(C.lib().cufhe_initialize_lvl02(C._ptr(bk), bk.size, 1, C._ptr(ksk), ksk.size),
            "initialize_lvl02")
    rng = np.random.default_rng(92)
    ops = _gates(rng, N)
    bits = rng.integers(0, 2, 3 * N).astype(np.uint8)
    pool_ct = C.encrypt_bits(0, 93, keys_lvl2["k0"], bits)
    plain = _plain(ops, pool_ct, N, W0, C.lib().cufhe_gate_lvl02)
    outs = _batched(ops, pool_ct, N, W0, 2) + _batched(ops, pool_ct, N, W0, 2, graph=True, replays=2)
    for out in outs:
        assert np.array_equal(out, plain)


@pytest.mark.parametrize("split", [1, 5])
def test_lvl1_mfma_key_switch_split_and_ragged(engine, keys_lvl1, split):
    """333 gates (10 full 32-job blocks and a partial one) with the split
    forced to 1 and to 5; the default choice restored afterwards."""
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    n = 333
    rng = np.random.default_rng(94 + split)
    ops = _gates(rng, n)
    bits = rng.integers(0, 2, 3 * n).astype(np.uint8)
    pool_ct = C.encrypt_bits(1, 95, keys_lvl1["k1"], bits)
    plain = _plain(ops, pool_ct, n, W1, C.lib().cufhe_gate_lvl1)
    C.check(C.lib().cufhe_set_key_switch_split(split), "set split")
    try:
        out = _batched(ops, pool_ct, n, W1, 1)[0]
    finally:
        C.check(C.lib().cufhe_set_key_switch_split(0), "set split")
    assert np.array_equal(out, plain)
