"""Pin the oracle's annihilation key switching (oracle.cpp: automorphism,
annihilate) to what the mathematics fixes, and the host's annihilation key
generation to its definition. CPU only.

* automorphism: tau_d(p) evaluated at the odd powers of exp(i pi / N) equals
  p evaluated at their d-th powers (p(X) -> p(X^d) mod X^N + 1), and
  tau_d1 tau_d2 = tau_(d1 d2 mod 2N);
* annihilating a trivial TRLWE (a = 0, b = mu at X^0) returns it exactly
  (reference test/test_annihilate_gpu.cu:68-71, its case 0), in every mode;
* with a noiseless key the output phase keeps the constant coefficient of the
  input phase and the other coefficients vanish, within the gadget
  decomposition error (random message at every coefficient, random masks),
  in every mode and at both levels;
* each planted mistake (the automorphism's sign, the halving, a digit paired
  with the wrong key row) breaks that check;
* the key's rows decrypt to tau_d(s) h_q exactly for a noiseless key.
"""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O

MODES = {"exact": O.EXACT, "fft": O.FFT, "ntt": O.NTT}
LEVELS = {1: (1024, np.uint32, 32, 2, 8), 2: (2048, np.uint64, 64, 4, 10)}  # N, T, w, l, Bgbit
# Phase error bound (torus units) of a noiseless-key annihilation: the l-digit
# decomposition leaves 2^(w - l Bgbit - 1) per coefficient, summed over ~N/2
# key bits (x sqrt(N/2)) and nbit rounds (x sqrt(nbit)): ~2^20.4 at level 1,
# ~2^29.6 at level 2, measured 2^21.0 and 2^29.9. FP64 at level 2 also rounds
# the 64-bit key words to 53 bits (measured 2^34.6). An index or sign mistake
# leaves errors near 2^(w-1).
BOUND = {(1, "exact"): 1 << 23, (1, "fft"): 1 << 23, (1, "ntt"): 1 << 23,
         (2, "exact"): 1 << 32, (2, "fft"): 1 << 37, (2, "ntt"): 1 << 32}


def _signed(x, level):
    return x.view(np.int32).astype(np.int64) if level == 1 else x.view(np.int64)


@pytest.fixture(scope="module")
def keys():
    k0, k1, k2 = C.secret_key_gen(21)
    return {1: k1, 2: k2}


@pytest.fixture(scope="module")
def noiseless_ahk(keys):
    return {lv: C.annihilate_key_gen(lv, 7, keys[lv], alpha=0.0) for lv in (1, 2)}


@pytest.mark.parametrize("level", [1, 2])
def test_automorphism_is_substitution_x_to_x_d(level):
    N, T, w = LEVELS[level][:3]
    rng = np.random.default_rng(level)
    p = rng.integers(-1000, 1000, N)
    words = p.astype(np.int64).astype(T) if level == 2 else p.astype(np.int32).view(np.uint32)
    ks = np.array([0, 1, 5, N // 3, N - 1])
    zeta = np.exp(1j * np.pi * (2 * ks + 1) / N)
    for d in (3, 5, 2 * N - 1, 2 ** (LEVELS[level][0].bit_length() - 1) + 1):
        got = _signed(O.automorphism(level, words, d), level).astype(np.float64)
        lhs = np.array([np.polyval(got[::-1], z) for z in zeta])
        rhs = np.array([np.polyval(p[::-1].astype(np.float64), z ** d) for z in zeta])
        assert np.allclose(lhs, rhs, rtol=0, atol=1e-6 * np.abs(p).sum()), d
    d1, d2 = 5, 2 * N - 3
    twice = O.automorphism(level, O.automorphism(level, words, d2), d1)
    assert np.array_equal(twice, O.automorphism(level, words, (d1 * d2) % (2 * N)))
    assert np.array_equal(O.automorphism(level, words, 1), words)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("level", [1, 2])
def test_trivial_body_is_returned_exactly(level, mode, noiseless_ahk, keys):
    N, T, w = LEVELS[level][:3]
    ahk = C.annihilate_key_gen(level, 8, keys[level])  # noisy key: a = 0 never reaches it
    ct = np.zeros(2 * N, T)
    ct[N] = T(1) << T(w - 3)
    assert np.array_equal(O.annihilate(level, MODES[mode], ct, ahk), ct)


def annihilation_error(level, mode, ahk, key, seed):
    """max |phase(out) - (m_0, 0, ..., 0)| over the coefficients (torus units)."""
    N, T, w = LEVELS[level][:3]
    rng = np.random.default_rng(seed)
    m = rng.integers(0, 2**63, N, dtype=np.uint64)
    m = m.astype(T) if level == 2 else (m >> np.uint64(31)).astype(T)
    ct = C.trlwe_encrypt_level(level, seed, key, m, alpha=0.0)[0]
    ph = C.trlwe_phase(level, key, O.annihilate(level, MODES[mode], ct, ahk))[0]
    want = np.zeros(N, T)
    want[0] = m[0]
    return int(np.max(np.abs(_signed(ph - want, level))))


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("level", [1, 2])
def test_noiseless_key_keeps_the_constant_coefficient(level, mode, noiseless_ahk, keys):
    for seed in (1, 2):
        e = annihilation_error(level, mode, noiseless_ahk[level], keys[level], seed)
        assert e < BOUND[(level, mode)], (seed, np.log2(e + 1))


@pytest.mark.parametrize("name", ["auto_sign", "ah_no_halve", "ah_key_row"])
@pytest.mark.parametrize("level", [1, 2])
def test_planted_annihilation_mistakes_are_caught(name, level, noiseless_ahk, keys):
    with O.mutation(name):
        e = annihilation_error(level, "exact", noiseless_ahk[level], keys[level], 3)
    assert e > 64 * BOUND[(level, "exact")], (name, np.log2(e + 1))


@pytest.mark.parametrize("level", [1, 2])
def test_key_rows_decrypt_to_automorphism_of_the_key(level, noiseless_ahk, keys):
    N, T, w, l, bg = LEVELS[level]
    nbit = N.bit_length() - 1
    rows = noiseless_ahk[level].reshape(nbit, l, 2 * N)
    s = keys[level]
    for bit in range(nbit):
        d = (2 << bit) + 1
        ts = O.automorphism(level, s, d)
        for q in range(l):
            h = T(1) << T(w - (q + 1) * bg)
            assert np.array_equal(C.trlwe_phase(level, s, rows[bit, q])[0], ts * h), (bit, q)
    noisy = C.annihilate_key_gen(level, 7, s).reshape(nbit, l, 2 * N)
    err = _signed(C.trlwe_phase(level, s, noisy[0, 0])[0] - O.automorphism(level, s, 3) *
                  (T(1) << T(w - bg)), level)
    assert 0 < np.max(np.abs(err)) < (1 << (w - 18))
