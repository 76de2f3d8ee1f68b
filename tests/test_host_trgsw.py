"""Host-side TRGSW / TRLWE encryption (tfhe_trgsw_encrypt, tfhe_trlwe_encrypt):
each TRGSW row r is a TRLWE whose phase is m * 2^(32 - (r % l + 1) Bgbit) on
component r / l (constant coefficient) plus lvl1 noise, as TFHEpp's
trgswSymEncrypt builds it; a TRLWE's phase is its message. CPU only."""
import numpy as np

import cufhe_amd as C

N = C.N1


def _negacyclic_mul(a, s):
    """a * s mod (X^N + 1, 2^32) for a binary s (int64 arithmetic)."""
    out = np.zeros(N, np.int64)
    for j in np.nonzero(s)[0]:
        out += np.concatenate([-a[N - j:], a[:N - j]]) if j else a
    return out


def _phase(trlwe, s):
    a = trlwe[:N].astype(np.int64)
    b = trlwe[N:].astype(np.int64)
    return ((b - _negacyclic_mul(a, s)) % (1 << 32)).astype(np.uint32)


def _signed(x):
    return x.view(np.int32).astype(np.int64)


def test_trgsw_rows_phase(keys_lvl1):
    k1 = keys_lvl1["k1"].astype(np.int64)
    for m in (0, 1):
        t = C.trgsw_encrypt(5, keys_lvl1["k1"], m)[0].reshape(4, 2 * N)
        for r in range(4):
            j, q = r // 2, r % 2
            h = 1 << (32 - 8 * (q + 1))
            ph = _phase(t[r], k1)
            want = np.zeros(N, np.int64)
            if j == 1:
                want[0] = m * h  # b-component message
            else:
                # the a-component message m*h shifts the phase by -m*h*s
                want = -m * h * _negacyclic_mul(np.eye(1, N, 0, np.int64)[0], k1)
            err = _signed((ph - (want % (1 << 32)).astype(np.uint32)).astype(np.uint32))
            assert np.abs(err).max() < 1 << 20, (m, r)


def test_trlwe_phase_is_message(keys_lvl1):
    rng = np.random.default_rng(6)
    p = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    ct = C.trlwe_encrypt(7, keys_lvl1["k1"], p)[0]
    err = _signed((_phase(ct, keys_lvl1["k1"].astype(np.int64)) - p).astype(np.uint32))
    assert np.abs(err).max() < 1 << 12
