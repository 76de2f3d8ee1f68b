"""GPU parity of the small-modulus NTT backend (P = 2147473409) against the
CPU oracle, through the C ABI. Every result of this backend is an exact
residue computation, so the bar is bit-exact ciphertext equality with the
oracle's NTT mode (reference include/gatebootstrapping_gpu.cuh:522-797,
:1187-1332; include/ntt_small_modulus.cuh:284-509)."""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu

P = O.P31
N = 1024
MU = C.MU1
W1 = C.TLWE1_WORDS


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "cufhe_initialize")
    yield C.lib()


def ntt(x, inverse=False):
    x = np.ascontiguousarray(x, np.uint32)
    cnt = x.shape[0]
    din = C.DeviceBuffer.from_numpy(x)
    dout = C.DeviceBuffer(x.nbytes)
    C.check(C.lib().cufhe_ntt_transform(din.ptr, dout.ptr, cnt, int(inverse), None), "ntt")
    C.check(C.lib().cufhe_synchronize(), "sync")
    return dout.download(np.uint32, x.shape)


def evaluate(x, psi):
    """Z[k] = sum_i x_i psi^(i (2k+1)) mod P by direct evaluation."""
    pw = np.array([pow(psi, t, P) for t in range(2 * N)], np.int64)
    k = np.arange(N, dtype=np.int64)
    z = np.zeros(N, np.int64)
    for i in range(N):
        if x[i]:
            z = (z + int(x[i]) * pw[(i * (2 * k + 1)) % (2 * N)]) % P
    return z


def test_root_is_primitive_2n_th(engine):
    psi = int(C.lib().cufhe_ntt_root())
    assert pow(psi, N, P) == P - 1
    assert pow(psi, 2 * N, P) == 1


def test_forward_is_evaluation_at_odd_powers(engine):
    rng = np.random.default_rng(1)
    x = rng.integers(0, P, (3, N), dtype=np.int64).astype(np.uint32)
    x[2] = 0
    x[2, 5] = 1  # monomial: Z[k] = psi^(5 (2k+1))
    z = ntt(x)
    psi = int(C.lib().cufhe_ntt_root())
    for i in range(3):
        assert np.array_equal(z[i].astype(np.int64), evaluate(x[i], psi)), i


def test_inverse_roundtrip(engine):
    rng = np.random.default_rng(2)
    x = rng.integers(0, P, (5, N), dtype=np.int64).astype(np.uint32)
    x[0] = P - 1
    x[1] = 0
    assert np.array_equal(ntt(ntt(x), inverse=True), x)


def test_pointwise_product_is_negacyclic_product(engine):
    rng = np.random.default_rng(3)
    a = rng.integers(0, P, (2, N), dtype=np.int64).astype(np.uint32)
    b = rng.integers(0, P, (2, N), dtype=np.int64).astype(np.uint32)
    za, zb = ntt(a).astype(np.uint64), ntt(b).astype(np.uint64)
    prod = ntt((za * zb % P).astype(np.uint32), inverse=True)
    for i in range(2):
        want = O.negacyclic_mod(P, a[i].astype(np.uint64), b[i].astype(np.uint64))
        assert np.array_equal(prod[i].astype(np.uint64), want)


# ------------------------------------------------------------ key sets
def keyset(name, keys_lvl1, keys_bb):
    """(bk, ksk, variant, subset, oracle variant flags, lvl1 key) of a key set."""
    if name == "plain":
        return (keys_lvl1["bk"], keys_lvl1["ksk"], C.BR_PLAIN, False, O.BR_PLAIN, keys_lvl1["k1"])
    if name == "kb":
        return (keys_lvl1["bk_kb"], keys_lvl1["ksk"], C.BR_KEY_BUNDLE, False, O.BR_KB,
                keys_lvl1["k1"])
    return (keys_bb["bk"], keys_bb["ksk_subset"], C.BR_BLOCK_BINARY, True,
            O.BR_BB | O.SUBSET, keys_bb["k1"])


def load(name, keys_lvl1, keys_bb):
    bk, ksk, variant, subset, okb, k1 = keyset(name, keys_lvl1, keys_bb)
    load_lvl1_keys(("ntt", name), bk, ksk, backend=C.BACKEND_NTT, variant=variant, subset=subset)
    return bk, ksk, okb, k1


VARIANTS = ["plain", "kb", "bb"]


@pytest.mark.parametrize("name", VARIANTS)
@pytest.mark.parametrize("jpg", [1, 2, 6])
def test_bootstrap_bit_exact(engine, keys_lvl1, keys_bb, name, jpg):
    bk, _, okb, _ = load(name, keys_lvl1, keys_bb)
    rng = np.random.default_rng(60 + jpg)
    n = 5
    tl0 = rng.integers(0, 2**32, (n, 637), dtype=np.uint64).astype(np.uint32)
    din = C.DeviceBuffer.from_numpy(tl0)
    dout = C.DeviceBuffer(n * W1 * 4)
    C.check(C.lib().cufhe_bootstrap_lvl01(dout.ptr, din.ptr, MU, n, jpg, None), "bootstrap")
    g = dout.download(np.uint32, (n, W1))
    for i in range(n):
        tr = O.blind_rotate(1, O.NTT, okb & 3, tl0[i], bk, MU)
        e = O.sample_extract(1, tr)
        assert np.array_equal(g[i], e), (name, i, int(np.sum(g[i] != e)))


def _encrypt(k1, seed, bits):
    return C.encrypt_bits(1, seed, k1, np.asarray(bits, np.uint8))


@pytest.mark.parametrize("name", VARIANTS)
def test_plain_gates_bit_exact(engine, keys_lvl1, keys_bb, name):
    bk, ksk, okb, k1 = load(name, keys_lvl1, keys_bb)
    ops = ["NAND", "XOR", "AND", "OR", "XNOR", "MUX", "NMUX", "NOR"]
    a = [0, 1, 1, 0, 1, 0, 1, 1]
    b = [1, 1, 0, 0, 1, 1, 0, 1]
    s = [0, 0, 0, 0, 0, 1, 0, 1]
    ca, cb, cc = _encrypt(k1, 300, a), _encrypt(k1, 301, b), _encrypt(k1, 302, s)
    n = len(ops)
    da, db, dc = (C.DeviceBuffer.from_numpy(x) for x in (ca, cb, cc))
    do = C.DeviceBuffer(n * W1 * 4)
    st = C.Stream()
    for i, op in enumerate(ops):
        off = i * W1 * 4
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[op], do.at(off), da.at(off), db.at(off),
                                        dc.at(off), st.handle), "gate")
    st.synchronize()
    out = do.download(np.uint32, (n, W1))
    st.destroy()
    want = []
    for op, x, y, sel in zip(ops, a, b, s):
        if op == "MUX":
            want.append(y if sel else x)
        elif op == "NMUX":
            want.append(1 - (y if sel else x))
        else:
            want.append(C.TRUTH[op](x, y))
    assert list(C.decrypt_bits(1, k1, out)) == want
    for i, op in enumerate(ops):
        e = O.gate_lvl1(O.NTT, okb, C.GATES[op], ca[i], cb[i], cc[i], bk, ksk)
        assert np.array_equal(out[i], e), (name, op, int(np.sum(out[i] != e)))


def test_block_binary_batch_4096_nand_xor(engine, keys_lvl1, keys_bb):
    """NTT backend, block-binary keys, subset key switching: 4096 lvl1
    NAND/XOR gates from a seed as one batched plan (graph replay), every
    gate decrypting correctly and a seeded sample bit-exact against the
    oracle."""
    bk, ksk, okb, k1 = load("bb", keys_lvl1, keys_bb)
    rng = np.random.default_rng(4096)
    n = 4096
    ops = np.where(rng.integers(0, 2, n) == 0, "NAND", "XOR")
    bits = rng.integers(0, 2, 2 * n).astype(np.uint8)
    pool_ct = C.encrypt_bits(1, 777, k1, bits)
    pool = np.zeros((3 * n, W1), np.uint32)
    pool[: 2 * n] = pool_ct
    gates = [(str(ops[i]), 2 * n + i, i, n + i, 0) for i in range(n)]
    res = []
    for c in (1, 6):  # per-job key loads / key shared by six jobs through LDS
        dpool = C.DeviceBuffer.from_numpy(pool)
        plan = C.Plan(gates, dpool, jobs_per_group=c, use_graph=True)
        for _ in range(2):
            plan.run()
            C.check(C.lib().cufhe_synchronize(), "sync")
        res.append(dpool.download(np.uint32, pool.shape)[2 * n:])
        plan.destroy()
    assert np.array_equal(res[0], res[1])
    out = res[1]
    want = np.array([C.TRUTH[str(o)](int(bits[i]), int(bits[n + i])) for i, o in enumerate(ops)],
                    np.uint8)
    dec = C.decrypt_bits(1, k1, out)
    assert int(np.sum(dec != want)) == 0
    sample = rng.choice(n, 48, replace=False)
    codes = np.array([C.GATES[str(ops[i])] for i in sample], np.int32)
    ref = O.gate_batch_lvl1(O.NTT, okb, codes, pool_ct[sample], pool_ct[n + sample], bk, ksk, 16)
    for j, i in enumerate(sample):
        assert np.array_equal(out[i], ref[j]), int(i)
