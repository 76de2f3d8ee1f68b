"""lvl0 gate API with the lvl1 keys (reference Nand<lvl0param> ... NMux<lvl0param>:
BR-IKS at lvl1) and the TRLWE extras (SampleExtractAndKeySwitch,
GateBootstrappingTLWE2TRLWE, Refresh), on both backends. NTT: bit-exact to
the oracle composed from its primitives; FFT: decryptions and, before the key
switch, coefficient equality with the exact-integer oracle."""
import numpy as np
import pytest

import cufhe_amd as C
import oracle as O
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu

W0, W1 = C.TLWE0_WORDS, C.TLWE1_WORDS
MU = C.MU1
OPS = ["NAND", "OR", "AND", "XOR", "XNOR", "NOR", "ANDYN", "ORNY", "MUX", "NMUX", "NOT", "COPY"]


@pytest.fixture(scope="module")
def engine():
    C.check(C.lib().cufhe_initialize(), "initialize")


def _setups(keys_lvl1, keys_bb):
    return {
        "fft_kb": dict(keys=keys_lvl1, bk=keys_lvl1["bk_kb"], ksk=keys_lvl1["ksk"],
                       backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE, subset=False,
                       mode=O.EXACT, okb=O.BR_KB),
        "fft_plain": dict(keys=keys_lvl1, bk=keys_lvl1["bk"], ksk=keys_lvl1["ksk"],
                          backend=C.BACKEND_FFT, variant=C.BR_PLAIN, subset=False,
                          mode=O.EXACT, okb=O.BR_PLAIN),
        "ntt_kb": dict(keys=keys_lvl1, bk=keys_lvl1["bk_kb"], ksk=keys_lvl1["ksk"],
                       backend=C.BACKEND_NTT, variant=C.BR_KEY_BUNDLE, subset=False,
                       mode=O.NTT, okb=O.BR_KB),
        "ntt_bb": dict(keys=keys_bb, bk=keys_bb["bk"], ksk=keys_bb["ksk_subset"],
                       backend=C.BACKEND_NTT, variant=C.BR_BLOCK_BINARY, subset=True,
                       mode=O.NTT, okb=O.BR_BB),
    }


def _load(name, cfg):
    load_lvl1_keys(("lvl0api", name), cfg["bk"], cfg["ksk"], backend=cfg["backend"],
                   variant=cfg["variant"], subset=cfg["subset"])


def oracle_gate_lvl01(cfg, op, x, y, s):
    """BR-IKS at lvl1 composed from oracle primitives (reference
    src/bootstrap_gpu.cu:1248-1273 HomGate, :1557-1672 Mux/NMux)."""
    mu0 = MU
    sub = cfg["subset"]

    def rot(a, b, sa, sb, off):
        t = (np.uint32(sa & 0xFFFFFFFF) * a + np.uint32(sb & 0xFFFFFFFF) * b).astype(np.uint32)
        t[-1] = np.uint32((int(t[-1]) + off) % 2**32)
        tr = O.blind_rotate(1, cfg["mode"], cfg["okb"], t, cfg["bk"], MU)
        return O.sample_extract(1, tr)
    if op in ("MUX", "NMUX"):
        r1 = rot(s, y, 1, 1, -mu0)
        r0 = rot(s, x, -1, 1, -mu0)
        sg = 1 if op == "MUX" else -1
        return O.identity_key_switch(1, r1, r0, sg, sg, (sg * MU) % 2**32, cfg["ksk"], subset=sub)
    if op in ("NOT", "COPY"):
        return (np.uint32(0) - x).astype(np.uint32) if op == "NOT" else x.copy()
    ca, cb, off = {"NAND": (-1, -1, 1), "OR": (1, 1, 1), "ORYN": (1, -1, 1), "ORNY": (-1, 1, 1),
                   "AND": (1, 1, -1), "ANDYN": (1, -1, -1), "ANDNY": (-1, 1, -1),
                   "NOR": (-1, -1, -1), "XOR": (2, 2, 2), "XNOR": (-2, -2, -2)}[op]
    r = rot(x, y, ca, cb, (off * mu0) % 2**32)
    return O.identity_key_switch(1, r, None, 1, 0, 0, cfg["ksk"], subset=sub)


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


@pytest.mark.parametrize("name", ["fft_kb", "fft_plain", "ntt_kb", "ntt_bb"])
def test_lvl0_gates_plain_and_batched(engine, keys_lvl1, keys_bb, name):
    cfg = _setups(keys_lvl1, keys_bb)[name]
    _load(name, cfg)
    k0 = cfg["keys"]["k0"]
    rng = np.random.default_rng(70)
    n = 2 * len(OPS)
    ops = OPS * 2
    bits = rng.integers(0, 2, 3 * n).astype(np.uint8)
    pool = np.zeros((4 * n, W0), np.uint32)
    pool[: 3 * n] = C.encrypt_bits(0, 71, k0, bits)
    want = np.array([_want(o, int(bits[i]), int(bits[n + i]), int(bits[2 * n + i]))
                     for i, o in enumerate(ops)], np.uint8)
    dpool = C.DeviceBuffer.from_numpy(pool)
    st = C.Stream()
    B = W0 * 4
    for i, op in enumerate(ops):
        C.check(C.lib().cufhe_gate_lvl0(C.GATES[op], dpool.at((3 * n + i) * B), dpool.at(i * B),
                                        dpool.at((n + i) * B), dpool.at((2 * n + i) * B),
                                        st.handle), "gate_lvl0")
    st.synchronize()
    st.destroy()
    plain = dpool.download(np.uint32, pool.shape)[3 * n:]
    assert np.array_equal(C.decrypt_bits(0, k0, plain), want)
    # batched plan (C = 1 and graph replay) == plain bit for bit
    for c, graph in ((1, False), (2, True)):
        dp2 = C.DeviceBuffer.from_numpy(pool)
        plan = C.Plan([(ops[i], 3 * n + i, i, n + i, 2 * n + i) for i in range(n)], dp2,
                      jobs_per_group=c, use_graph=graph, level=0)
        for _ in range(2 if graph else 1):
            plan.run()
            C.check(C.lib().cufhe_synchronize(), "sync")
        got = dp2.download(np.uint32, pool.shape)[3 * n:]
        plan.destroy()
        assert np.array_equal(got, plain), (c, graph)
    # oracle parity (NTT: every word; FFT: the key switch follows a rotation
    # that equals the exact oracle, so outputs match wherever the rotation did)
    for i in [0, 3, 8, 9]:
        e = oracle_gate_lvl01(cfg, ops[i], pool[i], pool[n + i], pool[2 * n + i])
        if cfg["backend"] == C.BACKEND_NTT:
            assert np.array_equal(plain[i], e), ops[i]
        else:
            assert np.mean(plain[i] == e) >= 0.99, ops[i]


@pytest.mark.parametrize("name", ["fft_kb", "ntt_bb"])
def test_trlwe_extras(engine, keys_lvl1, keys_bb, name):
    cfg = _setups(keys_lvl1, keys_bb)[name]
    _load(name, cfg)
    k0 = cfg["keys"]["k0"]
    rng = np.random.default_rng(80)
    n = 3
    bits = np.array([1, 0, 1], np.uint8)
    tl0 = C.encrypt_bits(0, 81, k0, bits)
    d_in = C.DeviceBuffer.from_numpy(tl0)
    d_tr = C.DeviceBuffer(n * 2 * 1024 * 4)
    C.check(C.lib().cufhe_bootstrap_to_trlwe_lvl01(d_tr.ptr, d_in.ptr, MU, n, None), "to_trlwe")
    tr = d_tr.download(np.uint32, (n, 2048))
    for i in range(n):
        e = O.blind_rotate(1, cfg["mode"], cfg["okb"], tl0[i], cfg["bk"], MU)
        eq = np.mean(tr[i] == e)
        assert eq == 1.0 if cfg["backend"] == C.BACKEND_NTT else eq >= 0.999
    # SampleExtractAndKeySwitch of those TRLWEs: the test vector's constant
    # coefficient is +mu for bit 1 -> decrypts to the input bit
    d_out = C.DeviceBuffer(n * W0 * 4)
    C.check(C.lib().cufhe_sample_extract_key_switch_lvl10(d_out.ptr, d_tr.ptr, n, None), "seiks")
    C.check(C.lib().cufhe_synchronize(), "sync")
    out = d_out.download(np.uint32, (n, W0))
    assert list(C.decrypt_bits(0, k0, out)) == list(bits)
    for i in range(n):
        e = O.identity_key_switch(1, O.sample_extract(1, tr[i]), None, 1, 0, 0, cfg["ksk"],
                                  subset=cfg["subset"])
        assert np.array_equal(out[i], e)
    # Refresh: TRLWE -> SE + KS -> rotation -> TRLWE carrying the same bit
    d_ref = C.DeviceBuffer(n * 2 * 1024 * 4)
    C.check(C.lib().cufhe_refresh_lvl1(d_ref.ptr, d_tr.ptr, n, None), "refresh")
    ref = d_ref.download(np.uint32, (n, 2048))
    k1 = cfg["keys"]["k1"]
    for i in range(n):
        se = O.sample_extract(1, ref[i])
        assert C.decrypt_bits(1, k1, se[None, :])[0] == bits[i]
