"""Pin the oracle's integer semantics to closed forms (tests/pins.py) and show
that each pin catches a planted mistake. CPU only.

Pinned here: the blind-rotation index arithmetic (body truncation, rounded
masks, bar = 2N - ms(b), the test vector, the key-bundle pairing, the
block-binary body correction with correction != 0) in the exact, FFT and NTT
modes, at lvl1 and (with masks that round to a = N and a0 + a1 = N) at lvl2; the identity key switch's rounding (lvl10, subset lvl10, lvl20); and
ntt_mod_to_torus32 over every residue.
"""
import numpy as np
import pytest

import oracle as O
import pins as Q

MODES = {"exact": O.EXACT, "fft": O.FFT, "ntt": O.NTT}
VARIANTS = {"plain": O.BR_PLAIN, "kb": O.BR_KB, "bb": O.BR_BB}
# NTT mode: each step rounds the gadget entries to Z_P and back (a few
# hundred torus units), far below the 2 mu = 2^30 step of an index error.
NTT_TOL = 1 << 18


def _key(rng, variant):
    if variant == "bb":
        s = np.zeros(Q.N0, np.int64)
        for blk in range(Q.N0 // Q.ELL):
            pos = rng.integers(0, Q.ELL + 1)
            if pos < Q.ELL:
                s[blk * Q.ELL + pos] = 1
        return s
    return rng.integers(0, 2, Q.N0)


def _sample(rng):
    return rng.integers(0, 2**32, Q.N0 + 1, dtype=np.uint64).astype(np.uint32)


def rotation_mismatch(mode, variant, tl0, s0, bk):
    """Largest |oracle - closed form| over the rotated TRLWE (torus units)."""
    got = O.blind_rotate(1, MODES[mode], VARIANTS[variant], tl0, bk, Q.MU1)
    want = Q.rotated_test_vector(Q.rotation_amount(tl0, s0, variant))
    return int(np.max(np.abs((got - want).view(np.int32).astype(np.int64))))


CASES = [("plain", "exact"), ("plain", "fft"), ("plain", "ntt"), ("kb", "exact"), ("kb", "fft"),
         ("kb", "ntt"), ("bb", "exact"), ("bb", "ntt")]


@pytest.mark.parametrize("variant,mode", CASES)
def test_noiseless_rotation_is_the_closed_form(variant, mode):
    rng = np.random.default_rng(100 + list(VARIANTS).index(variant))
    s0 = _key(rng, variant)
    bk = Q.noiseless_bk(s0, variant)
    for _ in range(2 if mode == "exact" else 3):
        tl0 = _sample(rng)
        d = rotation_mismatch(mode, variant, tl0, s0, bk)
        if mode == "ntt":
            assert d < NTT_TOL, d
        else:
            assert d == 0, d


# lvl2 (N = 2048, Torus64): Z_P rounding of the gadget entries stays far below
# the 2 mu = 2^62 step of an index error
NTT2_TOL = 1 << 40
FFT2_TOL = 1 << 40


def _sample_lvl2_edges(rng):
    """A lvl0 sample whose first mask pair rounds to a0 + a1 = N (X^N - 1 = -2
    in the key bundle) and whose third mask rounds to a = N."""
    tl0 = _sample(rng)
    unit = (1 << 32) // (2 * Q.N2)
    tl0[0], tl0[1], tl0[2] = 1000 * unit, 1048 * unit, Q.N2 * unit
    assert Q.ms_round(tl0[0], Q.N2) + Q.ms_round(tl0[1], Q.N2) == Q.N2
    return tl0


@pytest.mark.parametrize("variant,mode", [("plain", "ntt"), ("kb", "ntt"), ("kb", "fft"),
                                          ("kb", "exact")])
def test_noiseless_rotation_is_the_closed_form_lvl2(variant, mode):
    rng = np.random.default_rng(150 + list(VARIANTS).index(variant))
    s0 = _key(rng, variant)
    s0[0] = s0[1] = s0[2] = 1  # the edge masks contribute
    bk = Q.noiseless_bk(s0, variant, n=Q.N2, l=4, bgbit=10, w=64)
    tl0 = _sample_lvl2_edges(rng)
    got = O.blind_rotate(2, MODES[mode], VARIANTS[variant], tl0, bk, Q.MU2)
    want = Q.rotated_test_vector(Q.rotation_amount(tl0, s0, variant, n=Q.N2), n=Q.N2, mu=Q.MU2,
                                 w=64)
    d = int(np.max(np.abs((got - want).view(np.int64).astype(np.float64))))
    tol = {"ntt": NTT2_TOL, "fft": FFT2_TOL, "exact": 0}[mode]
    assert d <= tol, d


def test_block_binary_samples_have_nonzero_correction():
    # the block-binary case above is only a pin of the correction if the
    # correction moves the rotation index for its samples
    rng = np.random.default_rng(100 + 2)
    _key(rng, "bb")
    moved = 0
    for _ in range(3):
        tl0 = _sample(rng)
        idx, corr = Q.bb_rounded_body(tl0)
        assert corr != 0
        moved += idx != Q.ms_round(tl0[Q.N0])
    assert moved >= 2


def test_closed_form_test_vector_edges():
    # total = 0: (mu, ..., mu); total = N: (-mu, ..., -mu); total = 1: (-mu, mu, ..., mu)
    # (reference test/test_accumulate_diag.cu:137-262 expects the last one after X^1)
    z = Q.rotated_test_vector(0)
    assert np.all(z[:Q.N1] == 0) and np.all(z[Q.N1:] == Q.MU1)
    m = Q.rotated_test_vector(Q.N1)
    assert np.all(m[Q.N1:] == np.uint32((-Q.MU1) % 2**32))
    o = Q.rotated_test_vector(1)
    assert o[Q.N1] == np.uint32((-Q.MU1) % 2**32) and np.all(o[Q.N1 + 1:] == Q.MU1)


# ------------------------------------------------------------------ key switch

def _ks_setup(rng, level, subset=False):
    s_dom = rng.integers(0, 2, Q.N1 if level == 1 else Q.N2)
    s0 = rng.integers(0, 2, Q.N0)
    if subset:
        s_dom[:Q.N0] = s0
    ksk = Q.noiseless_ksk(s_dom, s0, rng, subset=subset)
    return s_dom, s0, ksk


def ks_mismatch(level, rng, s_dom, s0, ksk, subset=False, count=6):
    w = 32 if level == 1 else 64
    dt = np.uint32 if level == 1 else np.uint64
    n = len(s_dom)
    bad = 0
    for _ in range(count):
        c = rng.integers(0, 2**w - 1, n + 1, dtype=np.uint64).astype(dt)
        out = O.identity_key_switch(level, c, None, 1, 0, 0, ksk, subset=subset)
        got = Q.lwe_phase32(out, s0)
        want = Q.ks_phase_closed_form(c, s_dom, w, subset_s0=s0 if subset else None)
        bad += got != want
    return bad


@pytest.mark.parametrize("level,subset", [(1, False), (1, True), (2, False)])
def test_noiseless_key_switch_phase_is_the_rounding_closed_form(level, subset):
    rng = np.random.default_rng(200 + level + 10 * subset)
    s_dom, s0, ksk = _ks_setup(rng, level, subset)
    assert ks_mismatch(level, rng, s_dom, s0, ksk, subset) == 0


def test_key_switch_pre_add_is_linear():
    # KS(ca in0 + cb in1 + offset) phase = closed form of the combined sample
    rng = np.random.default_rng(210)
    s_dom, s0, ksk = _ks_setup(rng, 1)
    for ca, cb, off in [(-1, -1, Q.MU1), (2, 2, 2 * Q.MU1), (1, -1, (-Q.MU1) % 2**32)]:
        x = rng.integers(0, 2**32, Q.N1 + 1, dtype=np.uint64)
        y = rng.integers(0, 2**32, Q.N1 + 1, dtype=np.uint64)
        comb = (ca * x.astype(object) + cb * y.astype(object)) % 2**32
        comb[-1] = (comb[-1] + off) % 2**32
        out = O.identity_key_switch(1, x.astype(np.uint32), y.astype(np.uint32), ca, cb, off, ksk)
        assert Q.lwe_phase32(out, s0) == Q.ks_phase_closed_form(comb, s_dom, 32)


# ------------------------------------------------------------------ ntt_mod_to_torus32

def ntt_to_torus32_violations(chunk=1 << 24, stride=1):
    """Check ntt_mod_to_torus32 for residues v = mag (mag <= (P-1)/2, every
    `stride`-th) and their negations P - mag against the error analysis of
    the reference's reciprocal floor(2^63 / P) (ntt_small_modulus.cuh:317-341):
    with x = mag 2^32 / P, the result is round-half-up(x - d) for some
    d in [0, mag / 2^31), so it is floor(x + 1/2) or one less, and exactly
    floor(x + 1/2) whenever frac(x + 1/2) >= mag / 2^31."""
    p = O.P31
    half = p // 2
    bad = 0
    for lo in range(0, half + 1, chunk * stride):
        mag = np.arange(lo, min(half + 1, lo + chunk * stride), stride, dtype=np.uint64)
        num = (mag << np.uint64(33)) + np.uint64(p)          # < 2^63 + 2^31
        rhu = num // np.uint64(2 * p)                        # floor(x + 1/2)
        rem = num % np.uint64(2 * p)                         # frac(x + 1/2) * 2P
        forced = rem * np.uint64(1 << 31) >= mag * np.uint64(2 * p)
        got = O.ntt_mod_to_torus(mag, p, 32).astype(np.uint64)
        ok = (got == rhu) | (~forced & (got + np.uint64(1) == rhu))
        bad += int(np.count_nonzero(~ok))
        neg_in = (np.uint64(p) - mag[mag > 0]) % np.uint64(p)
        neg = O.ntt_mod_to_torus(neg_in, p, 32).astype(np.uint64)
        bad += int(np.count_nonzero(neg != (np.uint64(1 << 32) - got[mag > 0]) % np.uint64(1 << 32)))
    return bad


def test_ntt_mod_to_torus32_every_residue():
    assert ntt_to_torus32_violations() == 0


# ------------------------------------------------------------------ the pins bite

def _mutated(name, fn):
    with O.mutation(name):
        return fn()


def test_planted_ntt_offset_is_caught():
    assert _mutated("ntt_torus_offset", lambda: ntt_to_torus32_violations(stride=997)) > 0
    assert ntt_to_torus32_violations(stride=997) == 0


@pytest.mark.parametrize("name,variant,mode", [
    ("bar_off_by_one", "plain", "fft"), ("b_rounded", "kb", "fft"), ("a_truncated", "kb", "fft"),
    ("kb_order", "kb", "fft"), ("bb_sign", "bb", "ntt"), ("bb_not_halved", "bb", "ntt"),
])
def test_planted_rotation_mistakes_are_caught(name, variant, mode):
    rng = np.random.default_rng(300)
    s0 = _key(rng, variant)
    bk = Q.noiseless_bk(s0, variant)
    tol = NTT_TOL if mode == "ntt" else 0
    worst = 0
    for _ in range(3):
        tl0 = _sample(rng)
        # body bits below the 2N grid at 3/4 of a step: truncation and
        # rounding of the body differ
        tl0[Q.N0] = (int(tl0[Q.N0]) & ~((1 << 21) - 1)) | (3 << 19)
        assert rotation_mismatch(mode, variant, tl0, s0, bk) <= tol
        worst = max(worst, _mutated(name, lambda: rotation_mismatch(mode, variant, tl0, s0, bk)))
    assert worst > tol


def test_planted_key_switch_rounding_mistake_is_caught():
    rng = np.random.default_rng(220)
    s_dom, s0, ksk = _ks_setup(rng, 1)
    assert _mutated("ks_no_round", lambda: ks_mismatch(1, rng, s_dom, s0, ksk)) > 0
