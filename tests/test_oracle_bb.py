"""Block-binary keys, subset key switching and the block-binary blind
rotation of the oracle (reference include/gatebootstrapping_gpu.cuh:661-884,
include/keyswitch_gpu.cuh:214-222, test/test_block_binary_gpu.cc). CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import signed32

N0, N1, ELL = 636, 1024, 4
MU1 = 1 << 29


def trlwe_phase(trlwe, key):
    """b - a*s mod (X^N + 1, 2^32) as signed int64 coefficients."""
    a = trlwe[:N1].astype(np.int64)
    b = trlwe[N1:].astype(np.int64)
    s = key.astype(np.int64)
    full = np.convolve(a, s)  # |values| < 2^42
    prod = full[:N1].copy()
    prod[: N1 - 1] -= full[N1:]
    ph = (b - prod) % (1 << 32)
    return np.where(ph >= 1 << 31, ph - (1 << 32), ph)


def test_block_binary_key_structure(keys_bb):
    # mirrors test/test_block_binary_gpu.cc:84-114 (ValidateBlockKey)
    k0, k1 = keys_bb["k0"], keys_bb["k1"]
    w = k0.reshape(-1, ELL).sum(axis=1)
    assert w.max() <= 1
    assert np.array_equal(k0, k1[:N0])
    # all ell+1 block patterns occur
    pat = [int(np.argmax(b)) + 1 if b.any() else 0 for b in k0.reshape(-1, ELL)]
    assert set(pat) == set(range(ELL + 1))


def test_subset_ksk_is_the_full_key_from_row_n0(keys_bb):
    import cufhe_amd as C
    k0, k1 = keys_bb["k0"], keys_bb["k1"]
    full = C.ksk_gen(1, 3, k1, k0)
    sub = keys_bb["ksk_subset"]
    row = 7 * 2 * (N0 + 1)
    assert sub.size == (N1 - N0) * row
    assert np.array_equal(sub, full[N0 * row:])


def test_subset_key_switch_decrypts(keys_bb):
    import cufhe_amd as C
    k0, k1 = keys_bb["k0"], keys_bb["k1"]
    bits = np.array([0, 1, 1, 0], np.uint8)
    ct = C.encrypt_bits(1, 21, k1, bits)
    ph1 = C.phase(1, k1, ct)
    for i in range(len(bits)):
        out = O.identity_key_switch(1, ct[i], None, 1, 0, 0, keys_bb["ksk_subset"], subset=True)
        ph0 = C.phase(0, k0, out[None, :])[0]
        assert abs(ph0 - ph1[i]) < 0.03


def test_subset_key_switch_copies_the_lvl0_part(keys_bb):
    # coefficients i < n0 start from a_i (include/keyswitch_gpu.cuh:214-222)
    # and only j >= n0 are switched: an input with a_j = 0 for j >= n0 is
    # already a lvl0 sample and comes out unchanged
    rng = np.random.default_rng(9)
    ct = rng.integers(0, 2**32, N1 + 1, dtype=np.uint64).astype(np.uint32)
    ct[N0:N1] = 0
    out = O.identity_key_switch(1, ct, None, 1, 0, 0, keys_bb["ksk_subset"], subset=True)
    assert np.array_equal(out[:N0], ct[:N0])
    assert out[N0] == ct[N1]


def _lvl0_sample(rng, k0, phase_units):
    """lvl0 TLWE with a_i multiples of 2^21 (no mod-switch rounding, so the
    block-binary correction is 0) and b = <a, s> + phase."""
    a = (rng.integers(0, 2048, N0).astype(np.uint64) << np.uint64(21)).astype(np.uint32)
    b = (int(np.dot(a.astype(np.uint64), k0.astype(np.uint64))) + phase_units) % (1 << 32)
    return np.concatenate([a, np.array([b], np.uint32)])


def test_block_binary_rotation_equals_plain_cmux_message(keys_bb):
    """With at most one key bit per block, sum_e BK_e (X^a_e - 1) is the CMUX
    product of the block's bits, so the block-binary rotation must carry the
    same message as the per-bit CMUX rotation on the same key."""
    rng = np.random.default_rng(5)
    k1 = keys_bb["k1"]
    for ph in (3 << 28, -(3 << 28) % (1 << 32), 1 << 27):
        tl0 = _lvl0_sample(rng, keys_bb["k0"], ph)
        plain = O.blind_rotate(1, O.EXACT, O.BR_PLAIN, tl0, keys_bb["bk"], MU1)
        bb = O.blind_rotate(1, O.EXACT, O.BR_BB, tl0, keys_bb["bk"], MU1)
        p_plain, p_bb = trlwe_phase(plain, k1), trlwe_phase(bb, k1)
        # both are +-mu test vectors up to BR noise (sigma ~ 2^24 per coefficient)
        assert np.all(np.abs(np.abs(p_plain) - MU1) < 1 << 27)
        assert np.array_equal(np.sign(p_plain), np.sign(p_bb))
        assert np.max(np.abs(p_plain - p_bb)) < 1 << 28


def test_block_binary_rotation_sign_follows_phase(keys_bb):
    # constant coefficient of the rotation = +mu iff the phase is in (0, 1/2)
    rng = np.random.default_rng(6)
    for ph in (1 << 29, 3 << 29, 5 << 29, 7 << 29):
        tl0 = _lvl0_sample(rng, keys_bb["k0"], ph)
        bb = O.blind_rotate(1, O.NTT, O.BR_BB, tl0, keys_bb["bk"], MU1)
        c0 = trlwe_phase(bb, keys_bb["k1"])[0]
        assert (c0 > 0) == (ph < 1 << 31)
        assert abs(abs(c0) - MU1) < 1 << 27


def test_block_binary_ntt_carries_the_exact_message(keys_bb):
    # Once one digit differs the masks diverge (key rows are uniform), so the
    # NTT and exact rotations are different ciphertexts of the same message:
    # compare decrypted phases, not words.
    rng = np.random.default_rng(7)
    tl0 = rng.integers(0, 2**32, N0 + 1, dtype=np.uint64).astype(np.uint32)
    e = O.blind_rotate(1, O.EXACT, O.BR_BB, tl0, keys_bb["bk"], MU1)
    t = O.blind_rotate(1, O.NTT, O.BR_BB, tl0, keys_bb["bk"], MU1)
    pe, pt = trlwe_phase(e, keys_bb["k1"]), trlwe_phase(t, keys_bb["k1"])
    assert np.array_equal(np.sign(pe), np.sign(pt))
    assert np.max(np.abs(pe - pt)) < 1 << 28
    assert np.any(signed32(t - e) != 0)


def test_block_binary_fft_is_rejected(keys_bb):
    tl0 = np.zeros(N0 + 1, np.uint32)
    with pytest.raises(RuntimeError):
        O.blind_rotate(1, O.FFT, O.BR_BB, tl0, keys_bb["bk"], MU1)


@pytest.mark.parametrize("op,fn", [(0, lambda x, y: 1 - (x & y)), (8, lambda x, y: x ^ y),
                                   (4, lambda x, y: x & y)])
def test_block_binary_subset_gates_decrypt(keys_bb, op, fn):
    import cufhe_amd as C
    k1 = keys_bb["k1"]
    x = np.array([0, 0, 1, 1], np.uint8)
    y = np.array([0, 1, 0, 1], np.uint8)
    cx, cy = C.encrypt_bits(1, 31 + op, k1, x), C.encrypt_bits(1, 41 + op, k1, y)
    outs = np.stack([O.gate_lvl1(O.NTT, O.BR_BB | O.SUBSET, op, cx[i], cy[i], None,
                                 keys_bb["bk"], keys_bb["ksk_subset"]) for i in range(4)])
    assert list(C.decrypt_bits(1, k1, outs)) == [fn(int(a), int(b)) for a, b in zip(x, y)]


def test_block_binary_subset_mux_decrypts(keys_bb):
    import cufhe_amd as C
    k1 = keys_bb["k1"]
    sel = np.array([0, 0, 1, 1, 0, 1], np.uint8)
    t = np.array([0, 1, 0, 1, 1, 1], np.uint8)
    f = np.array([1, 0, 1, 0, 0, 0], np.uint8)
    cs, ct, cf = (C.encrypt_bits(1, s, k1, v) for s, v in ((51, sel), (52, t), (53, f)))
    outs = np.stack([O.gate_lvl1(O.NTT, O.BR_BB | O.SUBSET, 10, cf[i], ct[i], cs[i],
                                 keys_bb["bk"], keys_bb["ksk_subset"]) for i in range(len(sel))])
    want = [int(tt if s else ff) for s, tt, ff in zip(sel, t, f)]
    assert list(C.decrypt_bits(1, k1, outs)) == want
