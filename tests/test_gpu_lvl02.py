"""lvl02 gates (lvl0 in/out, blind rotation at N=2048 over Torus64, lvl2->lvl0
key switch): plain per-gate form, batched plans, and oracle parity."""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

W0 = C.TLWE0_WORDS
BINARY = ["NAND", "OR", "ORYN", "ORNY", "AND", "ANDYN", "ANDNY", "NOR", "XOR", "XNOR"]


@pytest.fixture(scope="module")
def lvl02(keys_lvl2):
    C.check(C.lib().cufhe_initialize(), "initialize")
    bk, ksk = keys_lvl2["bk_kb"], keys_lvl2["ksk"]
    C.check(C.lib().cufhe_initialize_lvl02(C._ptr(bk), bk.size, 1, C._ptr(ksk), ksk.size),
            "initialize_lvl02")
    return keys_lvl2


def _want(op, x, y, s):
    if op == "MUX":
        return y if s else x
    if op == "NMUX":
        return 1 - (y if s else x)
    if op == "NOT":
        return 1 - x
    if op == "COPY":
        return x
    return C.TRUTH[op](x, y)


def _cases():
    ops, a, b, c = [], [], [], []
    for nm in BINARY:
        for x in (0, 1):
            for y in (0, 1):
                ops.append(nm); a.append(x); b.append(y); c.append(0)
    for nm in ("MUX", "NMUX"):
        for s in (0, 1):
            for x in (0, 1):
                ops.append(nm); a.append(x); b.append(1 - x); c.append(s)
    for nm in ("NOT", "COPY"):
        for x in (0, 1):
            ops.append(nm); a.append(x); b.append(0); c.append(0)
    return ops, np.array(a, np.uint8), np.array(b, np.uint8), np.array(c, np.uint8)


def test_lvl02_plain_and_batched(lvl02):
    k0 = lvl02["k0"]
    ops, a, b, c = _cases()
    n = len(ops)
    pool = np.zeros((4 * n, W0), np.uint32)
    pool[:n] = C.encrypt_bits(0, 300, k0, a)
    pool[n:2 * n] = C.encrypt_bits(0, 301, k0, b)
    pool[2 * n:3 * n] = C.encrypt_bits(0, 302, k0, c)
    want = np.array([_want(o, int(x), int(y), int(s)) for o, x, y, s in zip(ops, a, b, c)],
                    np.uint8)
    # plain per-gate form
    dpool = C.DeviceBuffer.from_numpy(pool)
    st = C.Stream()
    for i, op in enumerate(ops):
        C.check(C.lib().cufhe_gate_lvl02(C.GATES[op], dpool.at((3 * n + i) * W0 * 4),
                                         dpool.at(i * W0 * 4), dpool.at((n + i) * W0 * 4),
                                         dpool.at((2 * n + i) * W0 * 4), st.handle), "gate_lvl02")
    st.synchronize()
    st.destroy()
    plain = dpool.download(np.uint32, pool.shape)[3 * n:]
    assert np.array_equal(C.decrypt_bits(0, k0, plain), want)
    # batched plan (and graph replay) == plain bit for bit
    for graph in (False, True):
        dpool2 = C.DeviceBuffer.from_numpy(pool)
        plan = C.Plan([(ops[i], 3 * n + i, i, n + i, 2 * n + i) for i in range(n)], dpool2,
                      use_graph=graph, level=2)
        for _ in range(2 if graph else 1):
            plan.run()
            C.check(C.lib().cufhe_synchronize(), "sync")
        batched = dpool2.download(np.uint32, pool.shape)[3 * n:]
        plan.destroy()
        assert np.array_equal(batched, plain), graph
    # Oracle parity of whole gates (exact-integer lvl2 external products).
    # FP64 cannot hold Torus64 exactly: once one gadget digit of a rotation
    # differs the masks diverge, and the lvl2 -> lvl0 key switch then draws
    # different KSK rows, so the outputs are different lvl0 samples of the
    # same message. The bar is the message and a phase difference within the
    # key-switch noise (sigma ~ 0.01 for 2 x 2048 x 7 lvl0-noise rows).
    for i in [0, 42]:
        e = O.gate_lvl02(O.EXACT, True, C.GATES[ops[i]], pool[i], pool[n + i], pool[2 * n + i],
                         lvl02["bk_kb"], lvl02["ksk"])
        dph = abs(C.phase(0, k0, e[None, :])[0] - C.phase(0, k0, plain[i][None, :])[0])
        print(ops[i], "lvl0 phase diff vs oracle", dph)
        assert C.decrypt_bits(0, k0, e[None, :])[0] == want[i]
        assert dph < 0.08


def test_lvl02_rotation_phase_matches_exact_oracle(lvl02):
    """lvl2 rotation + SampleExtract against the exact-integer oracle before
    the key switch: both are Torus64 samples of the same test-vector
    coefficient; their phases differ by the rotation noise (sigma ~ 2^-24),
    far below the 2^-20 tolerance, while a wrong rotation index, sign or
    key-bundle term moves the phase by mu = 1/8."""
    k0, k2 = lvl02["k0"], lvl02["k2"]
    rng = np.random.default_rng(202)
    n = 4
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    din = C.DeviceBuffer.from_numpy(tl0)
    dout = C.DeviceBuffer(n * 2049 * 8)
    mu2 = 1 << 61
    C.check(C.lib().cufhe_bootstrap_lvl02(dout.ptr, din.ptr, mu2, n, None), "bootstrap_lvl02")
    g = dout.download(np.uint64, (n, 2049))
    for i in range(n):
        e = O.sample_extract(2, O.blind_rotate(2, O.EXACT, True, tl0[i], lvl02["bk_kb"], mu2))
        pg = C.phase(2, k2, g[i][None, :])[0]
        pe = C.phase(2, k2, e[None, :])[0]
        print("lvl2 phase gpu", pg, "exact", pe)
        assert abs(abs(pe) - 0.125) < 2 ** -16
        assert abs(pg - pe) < 2 ** -20


def test_lvl02_batch_larger_than_resident_grid(lvl02):
    """1100 lvl2 rotations in one launch: more jobs than the workgroups that
    fit on the device at once, so the grid walks them with a stride. Every
    output must equal the same job's output from a launch small enough for
    one pass, bit for bit, and carry a test-vector phase of +-1/8."""
    k2 = lvl02["k2"]
    rng = np.random.default_rng(203)
    n = 1100
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    mu2 = 1 << 61
    din = C.DeviceBuffer.from_numpy(tl0)
    dout = C.DeviceBuffer(n * 2049 * 8)
    C.check(C.lib().cufhe_bootstrap_lvl02(dout.ptr, din.ptr, mu2, n, None), "bootstrap_lvl02")
    big = dout.download(np.uint64, (n, 2049))
    pick = np.r_[0:8, 520:528, n - 8:n]
    sub = C.DeviceBuffer.from_numpy(np.ascontiguousarray(tl0[pick]))
    sout = C.DeviceBuffer(len(pick) * 2049 * 8)
    C.check(C.lib().cufhe_bootstrap_lvl02(sout.ptr, sub.ptr, mu2, len(pick), None),
            "bootstrap_lvl02")
    small = sout.download(np.uint64, (len(pick), 2049))
    assert np.array_equal(big[pick], small)
    ph = np.abs(C.phase(2, k2, big))
    print("lvl2 |phase| range", ph.min(), ph.max())
    assert np.all(np.abs(ph - 0.125) < 2 ** -16)


def test_lvl02_two_jobs_per_workgroup_bit_identical(lvl02):
    """The two-jobs-per-workgroup lvl2 rotation (cufhe_set_lvl2_jobs_per_group)
    gives the one-job kernel's outputs bit for bit, including an odd count
    (a dead slot in the last workgroup)."""
    rng = np.random.default_rng(204)
    n = 701
    tl0 = rng.integers(0, 2**32, (n, W0), dtype=np.uint64).astype(np.uint32)
    mu2 = 1 << 61
    din = C.DeviceBuffer.from_numpy(tl0)
    outs = []
    try:
        for c in (1, 2):
            C.check(C.lib().cufhe_set_lvl2_jobs_per_group(c), "lvl2 jobs per group")
            dout = C.DeviceBuffer(n * 2049 * 8)
            C.check(C.lib().cufhe_bootstrap_lvl02(dout.ptr, din.ptr, mu2, n, None), "bootstrap")
            outs.append(dout.download(np.uint64, (n, 2049)))
    finally:
        C.check(C.lib().cufhe_set_lvl2_jobs_per_group(1), "lvl2 jobs per group")
    assert np.array_equal(outs[0], outs[1])
    assert C.lib().cufhe_set_lvl2_jobs_per_group(3) != 0
