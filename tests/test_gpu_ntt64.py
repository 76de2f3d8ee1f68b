"""lvl2 NTT backend (Goldilocks prime P = 2^64 - 2^32 + 1, N = 2048): the
transform against direct evaluation, and whole lvl2 rotations and lvl02
gates bit-exact against the oracle's NTT mode (exact arithmetic mod P with
the reference's torus <-> Z_P conversions, so every output word must match).
"""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

P = O.GOLDILOCKS
N = 2048
W0 = C.TLWE0_WORDS
MU2 = 1 << 61


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "initialize")


def ntt(x, inverse=False):
    x = np.ascontiguousarray(x, np.uint64)
    din = C.DeviceBuffer.from_numpy(x)
    dout = C.DeviceBuffer(x.nbytes)
    C.check(C.lib().cufhe_ntt64_transform(din.ptr, dout.ptr, x.shape[0], int(inverse), None),
            "ntt64")
    C.check(C.lib().cufhe_synchronize(), "sync")
    return dout.download(np.uint64, x.shape)


def rand_residues(rng, shape):
    # uniform in [0, P): rejection of the 2^32 - 1 values >= P
    x = rng.integers(0, 2**64, shape, dtype=np.uint64)
    bad = x >= np.uint64(P)
    while bad.any():
        x[bad] = rng.integers(0, 2**64, int(bad.sum()), dtype=np.uint64)
        bad = x >= np.uint64(P)
    return x


def evaluate_sparse(x, psi):
    """Z[k] = sum_i x_i psi^(i (2k+1)) mod P by direct evaluation (Python ints)."""
    nz = [i for i in range(N) if int(x[i])]
    out = []
    for k in range(N):
        z = 0
        for i in nz:
            z += int(x[i]) * pow(psi, (i * (2 * k + 1)) % (2 * N), P)
        out.append(z % P)
    return out


def test_root_is_primitive_2n_th(engine):
    psi = int(C.lib().cufhe_ntt64_root())
    assert pow(psi, N, P) == P - 1
    assert pow(psi, 2 * N, P) == 1


def test_forward_is_evaluation_at_odd_powers(engine):
    rng = np.random.default_rng(11)
    x = np.zeros((3, N), np.uint64)
    x[0, 5] = 1  # monomial: Z[k] = psi^(5 (2k+1))
    x[1, [0, 1, 777, 2047]] = rand_residues(rng, 4)
    x[2, rng.choice(N, 6, replace=False)] = rand_residues(rng, 6)
    x[2, 3] = P - 1  # the largest residue
    z = ntt(x)
    psi = int(C.lib().cufhe_ntt64_root())
    for i in range(3):
        assert [int(v) for v in z[i]] == evaluate_sparse(x[i], psi), i


def test_inverse_roundtrip(engine):
    rng = np.random.default_rng(12)
    x = rand_residues(rng, (5, N))
    x[0] = P - 1
    x[1] = 0
    x[2, :] = np.uint64(P - 2)
    assert np.array_equal(ntt(ntt(x), inverse=True), x)


def test_pointwise_product_is_negacyclic_product(engine):
    rng = np.random.default_rng(13)
    a = rand_residues(rng, (2, N))
    b = rand_residues(rng, (2, N))
    b[1] = rng.integers(0, 1024, N).astype(np.uint64)  # small digits, as in the rotation
    za, zb = ntt(a), ntt(b)
    prod = np.array([[int(u) * int(v) % P for u, v in zip(za[i], zb[i])] for i in range(2)],
                    np.uint64)
    got = ntt(prod, inverse=True)
    for i in range(2):
        assert np.array_equal(got[i], O.negacyclic_mod(P, a[i], b[i])), i


# ------------------------------------------------------------ rotations and gates
@pytest.fixture(scope="module")
def keys2_plain(keys_lvl2):
    k0, k2 = keys_lvl2["k0"], keys_lvl2["k2"]
    return C.bk_gen(2, False, 0, k0, k2)


def _bootstrap(tl0, mu=MU2):
    n = tl0.shape[0]
    din = C.DeviceBuffer.from_numpy(np.ascontiguousarray(tl0))
    dout = C.DeviceBuffer(n * (N + 1) * 8)
    C.check(C.lib().cufhe_bootstrap_lvl02(dout.ptr, din.ptr, mu, n, None), "bootstrap_lvl02")
    return dout.download(np.uint64, (n, N + 1))


@pytest.mark.parametrize("kb", [True, False], ids=["key_bundle", "plain"])
def test_rotation_bit_exact_vs_oracle(engine, keys_lvl2, keys2_plain, kb):
    """lvl2 rotation + SampleExtract on the NTT backend == the oracle's NTT
    mode word for word; the outputs also carry the test vector's +-1/8."""
    bk = keys_lvl2["bk_kb"] if kb else keys2_plain
    C.initialize_lvl02(bk, keys_lvl2["ksk"], backend=C.BACKEND_NTT, key_bundle=kb)
    rng = np.random.default_rng(31 + int(kb))
    n = 3
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    tl0[2, :-1] = 0  # all masks zero: the rotation is X^-bar of the test vector alone
    got = _bootstrap(tl0)
    for i in range(n):
        want = O.sample_extract(2, O.blind_rotate(2, O.NTT, kb, tl0[i], bk, MU2))
        print("case", i, "kb", kb, "mismatching words", int(np.sum(got[i] != want)), flush=True)
        assert np.array_equal(got[i], want), i
    ph = np.abs(C.phase(2, keys_lvl2["k2"], got))
    assert np.all(np.abs(ph - 0.125) < 2 ** -16)


def test_rotation_batch_larger_than_resident_grid(engine, keys_lvl2):
    """More lvl2 rotations than workgroups resident at once (grid stride):
    each output equals the same job run in a small launch, bit for bit."""
    C.initialize_lvl02(keys_lvl2["bk_kb"], keys_lvl2["ksk"], backend=C.BACKEND_NTT,
                       key_bundle=True)
    rng = np.random.default_rng(33)
    n = 600
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    big = _bootstrap(tl0)
    pick = np.r_[0:4, 300:304, n - 4:n]
    small = _bootstrap(np.ascontiguousarray(tl0[pick]))
    assert np.array_equal(big[pick], small)
    ph = np.abs(C.phase(2, keys_lvl2["k2"], big))
    assert np.all(np.abs(ph - 0.125) < 2 ** -16)


def _gate_cases():
    # (op, a, b, c)
    return [("NAND", 1, 1, 0), ("XOR", 1, 0, 0), ("AND", 0, 1, 0), ("MUX", 0, 1, 1),
            ("NMUX", 1, 0, 0), ("XNOR", 0, 0, 0)]


def test_lvl02_gates_bit_exact_vs_oracle(engine, keys_lvl2):
    """lvl02 gates (BR at lvl2 on the NTT backend, then the lvl2 -> lvl0 key
    switch): the plain per-gate form and a batched plan give the oracle's
    NTT-mode gate word for word, and decrypt to the truth table."""
    k0 = keys_lvl2["k0"]
    C.initialize_lvl02(keys_lvl2["bk_kb"], keys_lvl2["ksk"], backend=C.BACKEND_NTT,
                       key_bundle=True)
    cases = _gate_cases()
    n = len(cases)
    pool = np.zeros((4 * n, W0), np.uint32)
    pool[:n] = C.encrypt_bits(0, 401, k0, np.array([c[1] for c in cases], np.uint8))
    pool[n:2 * n] = C.encrypt_bits(0, 402, k0, np.array([c[2] for c in cases], np.uint8))
    pool[2 * n:3 * n] = C.encrypt_bits(0, 403, k0, np.array([c[3] for c in cases], np.uint8))
    dpool = C.DeviceBuffer.from_numpy(pool)
    st = C.Stream()
    for i, (op, _, _, _) in enumerate(cases):
        C.check(C.lib().cufhe_gate_lvl02(C.GATES[op], dpool.at((3 * n + i) * W0 * 4),
                                         dpool.at(i * W0 * 4), dpool.at((n + i) * W0 * 4),
                                         dpool.at((2 * n + i) * W0 * 4), st.handle), "gate_lvl02")
    st.synchronize()
    st.destroy()
    plain = dpool.download(np.uint32, pool.shape)[3 * n:]
    dpool2 = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([(cases[i][0], 3 * n + i, i, n + i, 2 * n + i) for i in range(n)], dpool2,
                  level=2)
    plan.run()
    C.check(C.lib().cufhe_synchronize(), "sync")
    batched = dpool2.download(np.uint32, pool.shape)[3 * n:]
    plan.destroy()
    assert np.array_equal(batched, plain)
    for i, (op, a, b, c) in enumerate(cases):
        want_bit = (b if c else a) if op == "MUX" else (1 - (b if c else a)) if op == "NMUX" \
            else C.TRUTH[op](a, b)
        assert C.decrypt_bits(0, k0, plain[i][None, :])[0] == want_bit, op
        if i < 3:
            e = O.gate_lvl02(O.NTT, True, C.GATES[op], pool[i], pool[n + i], pool[2 * n + i],
                             keys_lvl2["bk_kb"], keys_lvl2["ksk"])
            print(op, "mismatching words vs oracle", int(np.sum(e != plain[i])), flush=True)
            assert np.array_equal(e, plain[i]), op
    # restore the FFT backend for the other lvl02 tests
    C.initialize_lvl02(keys_lvl2["bk_kb"], keys_lvl2["ksk"], backend=C.BACKEND_FFT,
                       key_bundle=True)


def test_lvl02_mix_full_size_on_ntt_backend(engine, keys_lvl2):
    """The lvl02 benchmark mix at its full size (1536 binary gates uniform
    over NAND/AND/OR/XOR/XNOR/NOR + 512 MUX, interleaved) as one batched plan
    on the Goldilocks backend: every output decrypts to the plaintext
    evaluation, and a sample of gates equals the oracle's NTT mode word for
    word."""
    k0 = keys_lvl2["k0"]
    C.initialize_lvl02(keys_lvl2["bk_kb"], keys_lvl2["ksk"], backend=C.BACKEND_NTT,
                       key_bundle=True)
    rng = np.random.default_rng(77)
    mix = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR"]
    ops = [mix[i] for i in rng.integers(0, len(mix), 1536)] + ["MUX"] * 512
    ops = [ops[i] for i in rng.permutation(len(ops))]
    n = len(ops)
    bits = rng.integers(0, 2, 3 * n).astype(np.uint8)
    pool = np.zeros((4 * n, W0), np.uint32)
    pool[:3 * n] = C.encrypt_bits(0, 4100, k0, bits)
    dpool = C.DeviceBuffer.from_numpy(pool)
    plan = C.Plan([(ops[i], 3 * n + i, i, n + i, 2 * n + i) for i in range(n)], dpool, level=2)
    plan.run()
    C.check(C.lib().cufhe_synchronize(), "sync")
    plan.destroy()
    out = dpool.download(np.uint32, pool.shape)[3 * n:]
    x, y, s = bits[:n], bits[n:2 * n], bits[2 * n:]
    want = np.array([(y[i] if s[i] else x[i]) if o == "MUX" else C.TRUTH[o](int(x[i]), int(y[i]))
                     for i, o in enumerate(ops)], np.uint8)
    assert np.array_equal(C.decrypt_bits(0, k0, out), want)
    for i in (0, n // 2, n - 1):
        e = O.gate_lvl02(O.NTT, True, C.GATES[ops[i]], pool[i], pool[n + i], pool[2 * n + i],
                         keys_lvl2["bk_kb"], keys_lvl2["ksk"])
        assert np.array_equal(e, out[i]), (i, ops[i])
    C.initialize_lvl02(keys_lvl2["bk_kb"], keys_lvl2["ksk"], backend=C.BACKEND_FFT,
                       key_bundle=True)
