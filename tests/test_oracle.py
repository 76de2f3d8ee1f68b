"""The CPU oracle pinned against the reference's known-answer tests and the
mathematics (closed forms, brute force, invariants). CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import signed32, signed64

MU1 = 1 << 29


def _nega_brute_int(a, b):
    """Exact negacyclic product of integer vectors (Python ints)."""
    n = len(a)
    full = np.convolve(np.asarray(a, object), np.asarray(b, object))
    out = full[:n].copy()
    out[: n - 1] -= full[n:]
    return out


@pytest.mark.parametrize("p,n", [(O.P31, 1024), (O.GOLDILOCKS, 1024), (O.GOLDILOCKS, 2048)])
def test_negacyclic_ntt_matches_brute_force(p, n):
    rng = np.random.default_rng(5)
    a = rng.integers(0, min(p, 2**63), n, dtype=np.uint64) % np.uint64(p)
    b = rng.integers(0, min(p, 2**63), n, dtype=np.uint64) % np.uint64(p)
    fast = O.negacyclic_mod(p, a, b)
    brute = O.negacyclic_mod(p, a, b, brute=True)
    assert np.array_equal(fast, brute)
    # independent check of one coefficient with Python integers
    k = 7
    ref = sum(int(a[i]) * int(b[(k - i) % n]) * (1 if i <= k else -1) for i in range(n)) % p
    assert int(fast[k]) == ref


@pytest.mark.parametrize("n", [1024, 2048])
def test_ffnt_roundtrip(n):
    # reference test/test_fft_roundtrip.cu:141-184: U(-100, 100), error < 1e-6
    rng = np.random.default_rng(42)
    x = rng.uniform(-100, 100, n)
    y = O.ffnt_inverse(O.ffnt_forward(x))
    assert np.max(np.abs(x - y)) < 1e-6


@pytest.mark.parametrize("n", [1024, 2048])
def test_ffnt_square_of_trinomial(n):
    # reference test/test_fft_roundtrip.cu:186-243: (1+X+X^2)^2 = 1+2X+3X^2+2X^3+X^4
    a = np.zeros(n)
    a[:3] = 1
    s = O.ffnt_forward(a)
    c = O.ffnt_inverse(s * s)
    want = np.zeros(n)
    want[:5] = [1, 2, 3, 2, 1]
    assert np.max(np.abs(c - want)) < 1e-6


@pytest.mark.parametrize("n", [1024, 2048])
def test_ffnt_random_small_ints_negacyclic(n):
    # reference test/test_fft_roundtrip.cu:245-295: ints in [-5, 5], within 0.5
    rng = np.random.default_rng(123)
    a = rng.integers(-5, 6, n)
    b = rng.integers(-5, 6, n)
    c = O.ffnt_inverse(O.ffnt_forward(a.astype(float)) * O.ffnt_forward(b.astype(float)))
    want = np.array(_nega_brute_int(a, b), dtype=np.float64)
    assert np.max(np.abs(c - want)) < 0.5


@pytest.mark.parametrize("n", [1024, 2048])
def test_ffnt_is_evaluation_at_odd_roots(n):
    # closed form: Z[k] = p(zeta_k), zeta_k = exp(i*pi*(1-4k)/N)
    rng = np.random.default_rng(3)
    p = rng.integers(-3, 4, n).astype(float)
    z = O.ffnt_forward(p)
    for k in [0, 1, 5, n // 2 - 1]:
        zeta = np.exp(1j * np.pi * (1 - 4 * k) / n)
        ev = np.polyval(p[::-1], zeta)
        assert abs(z[k] - ev) < 1e-8 * max(1.0, abs(ev))


def test_torus_to_ntt_mod_is_rounding():
    # torus_to_ntt_mod(a) = floor((a P + 2^(w-1)) / 2^w) = round-half-up(a P / 2^w)
    rng = np.random.default_rng(9)
    a = np.concatenate([rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32),
                        np.array([0, 1, 2**31, 2**32 - 1], np.uint32)])
    got = O.torus_to_ntt_mod(a, O.P31)
    want = [(int(x) * O.P31 * 2 + 2**32) // 2**33 for x in a]
    assert [int(g) for g in got] == want
    # the top of the torus maps to P itself (reference lemma: values in [0, P])
    assert int(got[-1]) == O.P31
    a64 = rng.integers(0, 2**63, 100, dtype=np.uint64)
    got64 = O.torus_to_ntt_mod(a64, O.GOLDILOCKS)
    assert [int(g) for g in got64] == [(int(x) * O.GOLDILOCKS + 2**63) // 2**64 for x in a64]


def test_ntt_mod_to_torus32_is_near_rounding():
    p = O.P31
    rng = np.random.default_rng(10)
    v = np.concatenate([rng.integers(0, p, 2000, dtype=np.uint64),
                        np.array([0, 1, p // 2, p // 2 + 1, p - 1, p], np.uint64)])
    got = O.ntt_mod_to_torus(v, p, 32).astype(np.int64)
    for x, g in zip(v, got):
        x = int(x)
        c = x - p if x > p // 2 else x  # centred residue
        exact = c * 2**32 / p
        d = (int(g) - round(exact)) % 2**32
        assert d in (0, 1, 2**32 - 1), (x, g, exact)
    assert int(got[-6]) == 0 and int(got[-1]) == 0  # 0 and P both map to 0


@pytest.mark.parametrize("level", [1, 2])
def test_gadget_decomposition_identity(level):
    rng = np.random.default_rng(4)
    if level == 1:
        c = rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)
        l, bg, w = 2, 8, 32
    else:
        c = rng.integers(0, 2**64, 5000, dtype=np.uint64)
        l, bg, w = 4, 10, 64
    d = O.gadget_digits(level, c)
    assert d.min() >= -(1 << (bg - 1)) and d.max() < (1 << (bg - 1))
    for ci, di in zip(c[:500], d[:500]):
        rec = sum(int(di[q]) << (w - (q + 1) * bg) for q in range(l)) % 2**w
        err = (int(ci) - rec) % 2**w
        err = err - 2**w if err >= 2**(w - 1) else err
        assert abs(err) <= 2 ** (w - l * bg - 1)


def _noiseless_trgsw(level, m):
    n, l, bg, w = (1024, 2, 8, 32) if level == 1 else (2048, 4, 10, 64)
    dt = np.uint32 if level == 1 else np.uint64
    t = np.zeros((2 * l, 2, n), dt)
    for r in range(2 * l):
        t[r, r // l, 0] = dt(m << (w - (r % l + 1) * bg)) if m else 0
    return t.reshape(-1)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("mode", [O.EXACT, O.FFT, O.NTT])
def test_accumulate_kat_identity(level, mode):
    # reference test/test_accumulate_diag.cu:266-336: TRGSW = 0, abar = 0 -> unchanged
    n = 1024 if level == 1 else 2048
    dt = np.uint32 if level == 1 else np.uint64
    rng = np.random.default_rng(1)
    acc = rng.integers(0, np.iinfo(dt).max, 2 * n, dtype=np.uint64).astype(dt)
    out = O.cmux_accumulate(level, mode, acc, _noiseless_trgsw(level, 0), 0)
    assert np.array_equal(out, acc)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("mode", [O.EXACT, O.FFT, O.NTT])
def test_accumulate_kat_rotation(level, mode):
    # reference test/test_accumulate_diag.cu:137-262: noiseless TRGSW(1),
    # abar = 1, acc = (0, mu sum X^i) -> b = (-mu, mu, ..., mu), a = 0;
    # tolerance 1 (Torus32) / 2^20 (Torus64) as in the reference (:242)
    n = 1024 if level == 1 else 2048
    dt = np.uint32 if level == 1 else np.uint64
    mu = (1 << 29) if level == 1 else (1 << 61)
    acc = np.zeros(2 * n, dt)
    acc[n:] = mu
    out = O.cmux_accumulate(level, mode, acc, _noiseless_trgsw(level, 1), 1)
    want = np.zeros(2 * n, dt)
    want[n:] = mu
    want[n] = dt((-mu) % (2 ** (32 if level == 1 else 64)))
    diff = (out - want).view(np.int32 if level == 1 else np.int64).astype(np.float64)
    tol = 1 if level == 1 else 2 ** 20
    if mode == O.NTT:
        tol = 4 if level == 1 else 2 ** 20  # torus<->Z_P rounding of the gadget entries
    assert np.max(np.abs(diff)) <= tol


def _exact_extprod_numpy(d, trgsw, l, bg, w):
    """ExtProd(d, C) with numpy big ints: digits x key, negacyclic, mod 2^w."""
    n = d.size // 2
    digits = O.gadget_digits(1 if w == 32 else 2, d).reshape(2, n, l)
    C = trgsw.reshape(2 * l, 2, n)
    out = []
    for o in range(2):
        s = np.zeros(n, object)
        for r in range(2 * l):
            j, q = divmod(r, l)
            key = signed32(C[r, o]) if w == 32 else signed64(C[r, o])
            s = s + _nega_brute_int(digits[j, :, q].astype(np.int64), key)
        out.append(np.array([int(v) % 2**w for v in s], dtype=object))
    return np.concatenate(out)


def test_exact_external_product_matches_brute_force(keys_lvl1):
    rng = np.random.default_rng(8)
    d = rng.integers(0, 2**32, 2048, dtype=np.uint64).astype(np.uint32)
    trgsw = keys_lvl1["bk"][: 4 * 2 * 1024]
    got = O.external_product(1, O.EXACT, d, trgsw)
    want = _exact_extprod_numpy(d, trgsw, 2, 8, 32)
    assert [int(x) for x in got] == list(want)


def test_fft_external_product_equals_exact_lvl1(keys_lvl1):
    # FP64 error << 1/2 unit at lvl1, so the FFT mirror rounds to the exact result
    rng = np.random.default_rng(11)
    d = rng.integers(0, 2**32, 2048, dtype=np.uint64).astype(np.uint32)
    trgsw = keys_lvl1["bk"][: 4 * 2 * 1024]
    e = O.external_product(1, O.EXACT, d, trgsw)
    f = O.external_product(1, O.FFT, d, trgsw)
    assert np.array_equal(e, f)


def test_ntt_external_product_close_to_exact(keys_lvl1):
    # Z_P arithmetic approximates the torus: each key word is rounded to
    # Z_P (<= 1 torus unit), so |NTT - exact| <= sum |digits| * 1 ~ 2^19
    rng = np.random.default_rng(12)
    d = rng.integers(0, 2**32, 2048, dtype=np.uint64).astype(np.uint32)
    trgsw = keys_lvl1["bk"][: 4 * 2 * 1024]
    e = O.external_product(1, O.EXACT, d, trgsw)
    t = O.external_product(1, O.NTT, d, trgsw)
    diff = (t - e).view(np.int32).astype(np.int64)
    assert np.max(np.abs(diff)) < 2 ** 19
    assert np.std(diff) > 0  # the NTT path is not the exact one


def test_identity_key_switch_decrypts(keys_lvl1):
    import cufhe_amd as C
    k0, k1, ksk = keys_lvl1["k0"], keys_lvl1["k1"], keys_lvl1["ksk"]
    bits = np.array([0, 1, 1, 0, 1, 0, 0, 1], np.uint8)
    ct = C.encrypt_bits(1, 77, k1, bits)
    ph1 = C.phase(1, k1, ct)
    for i in range(len(bits)):
        out = O.identity_key_switch(1, ct[i], None, 1, 0, 0, ksk)
        ph0 = C.phase(0, k0, out[None, :])[0]
        assert abs(ph0 - ph1[i]) < 0.03  # key-switch noise, sigma ~ 7e-3


def test_gate_linear_combination_matches_reference_constants(keys_lvl1):
    # pre-add then key switch equals key switch of the combined input
    import cufhe_amd as C
    k0, k1, ksk = keys_lvl1["k0"], keys_lvl1["k1"], keys_lvl1["ksk"]
    ct = C.encrypt_bits(1, 5, k1, np.array([1, 0], np.uint8))
    comb = (np.uint32(2**32 - 1) * ct[0] + np.uint32(2**32 - 1) * ct[1]).astype(np.uint32)
    comb[-1] = np.uint32((int(comb[-1]) + MU1) % 2**32)
    a = O.identity_key_switch(1, ct[0], ct[1], -1, -1, MU1, ksk)
    b = O.identity_key_switch(1, comb, None, 1, 0, 0, ksk)
    assert np.array_equal(a, b)
