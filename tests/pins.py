"""Closed forms that pin the oracle (and, through the GPU tests that reuse
them, the HIP kernels) to the mathematics of TFHE gate bootstrapping.

With NOISELESS keys every step of a blind rotation is exact: a TRGSW(m) with
zero mask and zero noise is m times the gadget matrix, the test vector holds
only 0 and +-mu (exact in the l*Bgbit = 16 top bits the gadget keeps), so the
external product multiplies by X^(a_i s_i) exactly and the rotation ends at

    acc = (0, X^total * mu * (1 + X + ... + X^(N-1)))  in Z_2^32[X]/(X^N + 1),
    total = -ms(b) + sum_i ms_r(a_i) s_i  (mod 2N),

where ms(x) = floor(x 2N / 2^32) (the body is truncated) and ms_r rounds
half up (the mask coefficients carry a rounding offset). The block-binary
variant first subtracts half of the total mask rounding error from the body
and then rounds it. A noiseless key-switching key (random masks, zero noise)
makes the key switch's phase a closed form too:

    phase_out = phase_in + sum_j s_j (a_j - R(a_j)),  R = rounding to the
    t * basebit = 14 top bits of the torus.

Everything here is computed with Python integers from these definitions, not
by calling the code under test (reference semantics:
include/gatebootstrapping_gpu.cuh:10-53, :801-884, :1335-1380;
include/keyswitch_gpu.cuh:13-50, :189-256).
"""
import numpy as np

N0, N1, N2 = 636, 1024, 2048
MU1 = 1 << 29
MU2 = 1 << 61
ELL = 4


def ms_trunc(x, n=N1, w=32):
    """floor(x * 2N / 2^w): the body's mod switch (no rounding)."""
    return (int(x) % (1 << w)) * 2 * n // (1 << w)


def ms_round(x, n=N1, w=32):
    """round-half-up(x * 2N / 2^w) mod 2N: the mask's mod switch."""
    x = int(x) % (1 << w)
    return ((2 * x * 2 * n + (1 << w)) // (2 << w)) % (2 * n)


def _trunc_div2(c):
    """c / 2 rounded toward zero (C integer division)."""
    return c // 2 if c >= 0 else -((-c) // 2)


def bb_rounded_body(tl0, n=N1, w=32):
    """Block-binary body: b - (sum of the masks' rounding errors) / 2, then
    rounded to the 2N grid. The rounding error of a_i is a_i minus its
    rounded grid point, taken in the centred interval around 0."""
    unit = (1 << w) // (2 * n)
    corr = 0
    for a in tl0[:N0]:
        a = int(a)
        e = a - ms_round(a, n, w) * unit
        e = (e + (1 << (w - 1))) % (1 << w) - (1 << (w - 1))  # centred
        corr += e
    return ms_round((int(tl0[N0]) - _trunc_div2(corr)) % (1 << w), n, w), corr


def rotation_amount(tl0, s0, variant, n=N1):
    """total = -ms(b) + sum_i ms_r(a_i) s_i mod 2N (variant 'plain', 'kb',
    'bb'; 'bb' rounds the corrected body)."""
    if variant == "bb":
        b_idx, _ = bb_rounded_body(tl0, n)
    else:
        b_idx = ms_trunc(tl0[N0], n)
    total = -b_idx + sum(ms_round(tl0[i], n) * int(s0[i]) for i in range(N0))
    return total % (2 * n)


def rotated_test_vector(total, n=N1, mu=MU1, w=32):
    """TRLWE (a = 0, b = X^total * mu * sum_{j<N} X^j) as a 2N array."""
    dt = np.uint32 if w == 32 else np.uint64
    b = [0] * n
    for j in range(n):
        e = (j + total) % (2 * n)
        if e < n:
            b[e] = mu
        else:
            b[e - n] = (-mu) % (1 << w)
    out = np.zeros(2 * n, dt)
    out[n:] = np.array(b, dtype=object).astype(dt)
    return out


def noiseless_trgsw_rows(m, n=N1, l=2, bgbit=8, w=32):
    """TRGSW(m) for a vector m of 0/1 values: zero mask, zero noise."""
    dt = np.uint32 if w == 32 else np.uint64
    m = np.asarray(m, dt)
    t = np.zeros((len(m), 2 * l, 2, n), dt)
    for r in range(2 * l):
        t[:, r, r // l, 0] = m * dt(1 << (w - (r % l + 1) * bgbit))
    return t.reshape(-1)


def noiseless_bk(s0, variant, n=N1, l=2, bgbit=8, w=32):
    """Bootstrapping key of noiseless TRGSWs in the TFHEpp layout: one per
    key bit (plain, block-binary) or the key-bundle triples
    (s0 s1, s0 (1 - s1), (1 - s0) s1) per pair."""
    s0 = np.asarray(s0, np.int64)
    if variant == "kb":
        e, o = s0[0::2], s0[1::2]
        m = np.stack([e * o, e * (1 - o), (1 - e) * o], axis=1).reshape(-1)
    else:
        m = s0
    return noiseless_trgsw_rows(m, n, l, bgbit, w)


def noiseless_ksk(s_dom, s0, rng, t=7, basebit=2, subset=False):
    """Key-switching key [j][digit][|v|-1][n0+1] with uniform masks and zero
    noise: row (j, k, h) encrypts (h + 1) s_dom[j] 2^(32 - (k+1) basebit)."""
    first = N0 if subset else 0
    nj = len(s_dom) - first
    nh = 1 << (basebit - 1)
    a = rng.integers(0, 2**32, (nj, t, nh, N0), dtype=np.uint64)
    dot = (a * np.asarray(s0, np.uint64)).sum(axis=3) % (1 << 32)
    msg = np.zeros((nj, t, nh), np.uint64)
    for k in range(t):
        for h in range(nh):
            msg[:, k, h] = (h + 1) * np.asarray(s_dom[first:], np.uint64) * (1 << (32 - (k + 1) * basebit))
    b = (dot + msg) % (1 << 32)
    return np.concatenate([a, b[..., None]], axis=3).astype(np.uint32).reshape(-1)


def ks_round(a, w, t=7, basebit=2):
    """a rounded to the t * basebit top bits of the torus, as a Torus32
    value (w = 32 for lvl1 inputs, 64 for lvl2)."""
    keep = t * basebit
    sh = w - keep
    q = ((int(a) % (1 << w)) + (1 << (sh - 1))) >> sh  # round half up
    return (q << (32 - keep)) % (1 << 32)


def ks_phase_closed_form(c, s_dom, w=32, subset_s0=None):
    """Phase (Torus32, under the lvl0 key) of the key switch of a sample c
    (a_0..a_{n-1}, b) of width w under the noiseless key-switching key."""
    n = len(s_dom)
    if w == 32:
        ph = int(c[n])
    else:
        ph = ((int(c[n]) + (1 << 31)) >> 32) % (1 << 32)  # Torus64 -> Torus32
    first = 0
    if subset_s0 is not None:
        first = N0
        ph -= sum(int(c[i]) * int(subset_s0[i]) for i in range(N0))
    ph -= sum(ks_round(c[j], w) * int(s_dom[j]) for j in range(first, n))
    return ph % (1 << 32)


def lwe_phase32(ct, s):
    return (int(ct[-1]) - int(np.dot(ct[:-1].astype(np.uint64) % (1 << 32),
                                     np.asarray(s, np.uint64)) % (1 << 32))) % (1 << 32)
