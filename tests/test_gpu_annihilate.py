"""Annihilation key switching on the GPU (FFT backend, levels 1 and 2)
against the oracle (oracle.cpp annihilate, pinned in
tests/test_oracle_annihilate.py), through the C ABI:

* a trivial TRLWE comes back exactly (reference test/test_annihilate_gpu.cu
  case 0);
* encrypted TRLWEs: the GPU output's phase equals the exact oracle's within
  the FP64 rounding of the external products (level 1: 2^-12 of the torus,
  level 2: 2^-24), and at level 1 almost every output word is identical;
* with a noiseless key the output phase is (m_0, 0, ..., 0) within the
  decomposition error (the same bound the oracle meets);
* strided and in-place batches equal one-TRLWE calls bit for bit, and a batch
  larger than the resident grid keeps every output.
The reference test compares the GPU phase with TFHEpp's CPU phase within
2^26 (level 1) and 2^56 (level 2) torus units.
"""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

LEVELS = {1: (1024, np.uint32, 32), 2: (2048, np.uint64, 64)}
PHASE_TOL = {1: 1 << 20, 2: 1 << 40}
ERR_ARG, ERR_SIZE, ERR_STATE = -1, -2, -3
NOISELESS_BOUND = {1: 1 << 23, 2: 1 << 37}


def _signed(x, level):
    return x.view(np.int32).astype(np.int64) if level == 1 else x.view(np.int64)


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "initialize")


@pytest.fixture(scope="module")
def secret(keys_lvl1):
    return {1: keys_lvl1["k1"], 2: keys_lvl1["k2"]}


def _run(level, ct, out_stride=None, in_stride=None, in_place=False):
    """Annihilate the rows of ct on the GPU (one batched launch)."""
    N, T, _ = LEVELS[level]
    n = ct.shape[0]
    si = in_stride or 2 * N
    so = out_stride or 2 * N
    buf = np.zeros((n, si), T)
    buf[:, :2 * N] = ct
    din = C.DeviceBuffer.from_numpy(buf)
    dout = din if in_place else C.DeviceBuffer(n * so * np.dtype(T).itemsize)
    if in_place:
        so = si
    C.check(C.lib().cufhe_annihilate_key_switch_batch(level, dout.ptr, so, din.ptr, si, n, None),
            "annihilate")
    C.check(C.lib().cufhe_synchronize(), "sync")
    return dout.download(T, (n, so))[:, :2 * N]


def _messages(level, rng, n):
    N, T, w = LEVELS[level]
    m = rng.integers(0, 2**63, (n, N), dtype=np.uint64)
    return m.astype(T) if level == 2 else (m >> np.uint64(31)).astype(T)


@pytest.mark.parametrize("level", [1, 2])
def test_trivial_body_is_returned_exactly(engine, secret, level):
    N, T, w = LEVELS[level]
    C.annihilate_key_load(level, C.annihilate_key_gen(level, 31, secret[level]))
    ct = np.zeros((2, 2 * N), T)
    ct[0, N] = T(1) << T(w - 3)
    ct[1, N] = T(0) - (T(1) << T(w - 3))
    print("level", level, "trivial bodies", flush=True)
    assert np.array_equal(_run(level, ct), ct)


@pytest.mark.parametrize("level", [1, 2])
def test_phase_matches_exact_oracle(engine, secret, level):
    N, T, w = LEVELS[level]
    ahk = C.annihilate_key_gen(level, 32, secret[level])
    C.annihilate_key_load(level, ahk)
    rng = np.random.default_rng(40 + level)
    n = 4
    m = _messages(level, rng, n)
    m[0] = 0
    m[0, 0] = T(1) << T(w - 3)  # the reference's encrypted case: mu at X^0
    ct = C.trlwe_encrypt_level(level, 77, secret[level], m)
    got = _run(level, ct)
    for i in range(n):
        want = O.annihilate(level, O.EXACT, ct[i], ahk)
        dph = _signed(C.trlwe_phase(level, secret[level], got[i])[0] -
                      C.trlwe_phase(level, secret[level], want)[0], level)
        same = float(np.mean(got[i] == want))
        print("level", level, "case", i, "max phase diff 2^%.1f" % np.log2(np.abs(dph).max() + 1),
              "words equal %.4f" % same, flush=True)
        assert np.abs(dph).max() <= PHASE_TOL[level], i
        if level == 1:
            assert same > 0.99, i


@pytest.mark.parametrize("level", [1, 2])
def test_noiseless_key_keeps_the_constant_coefficient(engine, secret, level):
    N, T, w = LEVELS[level]
    C.annihilate_key_load(level, C.annihilate_key_gen(level, 33, secret[level], alpha=0.0))
    rng = np.random.default_rng(50 + level)
    m = _messages(level, rng, 3)
    ct = C.trlwe_encrypt_level(level, 78, secret[level], m, alpha=0.0)
    ph = C.trlwe_phase(level, secret[level], _run(level, ct))
    want = np.zeros_like(ph)
    want[:, 0] = m[:, 0]
    e = np.abs(_signed(ph - want, level)).max()
    print("level", level, "noiseless-key error 2^%.1f" % np.log2(e + 1), flush=True)
    assert e < NOISELESS_BOUND[level]


@pytest.mark.parametrize("level", [1, 2])
def test_strided_and_in_place_batches_equal_single_calls(engine, secret, level):
    N, T, w = LEVELS[level]
    C.annihilate_key_load(level, C.annihilate_key_gen(level, 34, secret[level]))
    rng = np.random.default_rng(60 + level)
    ct = C.trlwe_encrypt_level(level, 79, secret[level], _messages(level, rng, 5))
    single = np.stack([_run(level, ct[i:i + 1])[0] for i in range(5)])
    assert np.array_equal(_run(level, ct, out_stride=2 * N + 96, in_stride=2 * N + 32), single)
    assert np.array_equal(_run(level, ct, in_stride=2 * N + 64, in_place=True), single)
    one = C.DeviceBuffer.from_numpy(np.ascontiguousarray(ct[2]))
    C.check(C.lib().cufhe_annihilate_key_switch(level, one.ptr, one.ptr, None), "annihilate one")
    C.check(C.lib().cufhe_synchronize(), "sync")
    assert np.array_equal(one.download(T, (2 * N,)), single[2])


def test_large_batch_level1(engine, secret):
    """2048 level-1 TRLWEs in one launch (8 per CU): sampled outputs equal a
    small launch's, every output's constant coefficient decrypts."""
    N, T, w = LEVELS[1]
    C.annihilate_key_load(1, C.annihilate_key_gen(1, 35, secret[1]))
    rng = np.random.default_rng(70)
    n = 2048
    bits = rng.integers(0, 2, n)
    m = np.zeros((n, N), T)
    m[:, 0] = np.where(bits == 1, T(1 << 29), T((1 << 32) - (1 << 29)))
    m[:, 1:] = _messages(1, rng, n)[:, 1:]  # junk elsewhere: annihilated
    ct = C.trlwe_encrypt_level(1, 80, secret[1], m)
    big = _run(1, ct)
    pick = [0, 1, 1023, 2047]
    assert np.array_equal(big[pick], _run(1, np.ascontiguousarray(ct[pick])))
    ph0 = _signed(C.trlwe_phase(1, secret[1], big)[:, 0], 1)
    assert np.array_equal((ph0 > 0).astype(np.int64), bits)
    assert np.abs(np.abs(ph0) - (1 << 29)).max() < (1 << 26)


def test_errors_without_key_and_bad_arguments(engine, secret):
    N = 1024
    d = C.DeviceBuffer(2 * N * 4)
    C.check(C.lib().cufhe_annihilate_key_delete(1), "delete")
    assert C.lib().cufhe_annihilate_key_switch(1, d.ptr, d.ptr, None) == ERR_STATE
    assert C.lib().cufhe_annihilate_key_switch(3, d.ptr, d.ptr, None) == ERR_ARG
    assert C.lib().cufhe_annihilate_key_switch_batch(1, d.ptr, N, d.ptr, 2 * N, 1, None) == ERR_ARG
    short = np.zeros(5, np.uint32)
    assert C.lib().cufhe_annihilate_key_load(1, short.ctypes.data, 5) == ERR_SIZE


def test_lvl2_ntt_backend_bit_exact_vs_oracle(engine, secret):
    """Level 2 on the Goldilocks NTT backend (the reference's USE_FFT=OFF
    arithmetic): every output word equals the oracle's NTT mode, trivial
    bodies come back exactly, the noiseless-key identity holds; loading an
    FFT key afterwards switches the level back."""
    N, T, w = LEVELS[2]
    ahk = C.annihilate_key_gen(2, 36, secret[2])
    C.annihilate_key_load(2, ahk, backend=C.BACKEND_NTT)
    rng = np.random.default_rng(90)
    m = _messages(2, rng, 3)
    m[0] = 0
    m[0, 0] = T(1) << T(w - 3)
    ct = C.trlwe_encrypt_level(2, 81, secret[2], m)
    got = _run(2, ct)
    for i in range(3):
        want = O.annihilate(2, O.NTT, ct[i], ahk)
        print("ntt case", i, "mismatching words", int(np.sum(got[i] != want)), flush=True)
        assert np.array_equal(got[i], want), i
    triv = np.zeros((1, 2 * N), T)
    triv[0, N] = T(1) << T(w - 3)
    assert np.array_equal(_run(2, triv), triv)
    C.annihilate_key_load(2, C.annihilate_key_gen(2, 37, secret[2], alpha=0.0),
                          backend=C.BACKEND_NTT)
    ph = C.trlwe_phase(2, secret[2], _run(2, ct))
    want = np.zeros_like(ph)
    want[:, 0] = C.trlwe_phase(2, secret[2], ct)[:, 0]
    e = np.abs(_signed(ph - want, 2)).max()
    print("ntt noiseless-key error 2^%.1f" % np.log2(e + 1), flush=True)
    assert e < (1 << 33)  # key noise of the encryptions (alpha 2^-44) plus decomposition
    C.annihilate_key_load(2, ahk)  # FFT again
    fft = _run(2, ct[:1])
    assert not np.array_equal(fft[0], got[0])
    dph = _signed(C.trlwe_phase(2, secret[2], fft)[0] - C.trlwe_phase(2, secret[2], got[:1])[0], 2)
    assert np.abs(dph).max() <= PHASE_TOL[2]


def test_lvl1_ntt_backend_is_refused(engine, secret):
    ahk = C.annihilate_key_gen(1, 38, secret[1])
    assert C.lib().cufhe_annihilate_key_load_ex(1, ahk.ctypes.data, ahk.size, C.BACKEND_NTT) == -5
