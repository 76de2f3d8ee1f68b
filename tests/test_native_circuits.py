"""The library's C++ circuit builders (include/cufhe_circuits.h, behind the
C++ AESDec / ASCONXOF drop-ins) on the CPU: the AES-128 decryption circuit
decrypts the FIPS-197 Appendix C.1 block and random blocks to the oracle's
plaintexts (oracle/aes_ref.py); the AES and ASCON-XOF circuits equal the
Python builders' gate for gate; the ASCON ones match the oracle's Ascon-Xof
and the published empty-message digest; argument and buffer errors. No GPU."""
import ctypes

import numpy as np
import pytest

import cufhe_amd as C
from cufhe_amd import aes as A
from cufhe_amd import ascon as S
from cufhe_amd import circuits as K
from oracle import aes_ref as R
from oracle import ascon_ref as AR

KEY = bytes(range(16))
PT = bytes.fromhex("00112233445566778899aabbccddeeff")
CT = bytes.fromhex("69c4e0d86a7b0430d8cdb78070b4c55a")  # FIPS-197 Appendix C.1


def _layered(c):
    layer_of = np.zeros(c.n_inputs + c.n_gates, np.int64)
    start = 0
    for layer, end in enumerate(c.layer_end):
        layer_of[c.n_inputs + start:c.n_inputs + end] = layer + 1
        start = end
    g = np.arange(c.n_gates)
    me = layer_of[c.n_inputs + g]
    mux = np.isin(c.ops, [A.GATES["MUX"], A.GATES["NMUX"]])
    unary = np.isin(c.ops, [A.GATES["NOT"], A.GATES["COPY"]])
    return (start == c.n_gates and np.all(layer_of[c.in0] < me) and
            np.all(layer_of[c.in1[~unary]] < me[~unary]) and
            np.all(layer_of[c.inc[mux]] < me[mux]))


def test_native_aes_circuit_decrypts():
    c = K.native_aes128_decrypt()
    assert c.n_inputs == A.N_INPUTS and len(c.outputs) == 128 and _layered(c)
    assert A.from_bits(c.evaluate(A.input_bits(CT, R.expand_key(KEY)))) == PT
    rng = np.random.default_rng(11)
    for _ in range(6):
        k = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
        p = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
        ct = R.encrypt_block(k, p)
        assert A.from_bits(c.evaluate(A.input_bits(ct, R.expand_key(k)))) == p
    # the same construction as the Python builder, gate for gate
    py = A.decrypt_circuit()
    for f in ("ops", "in0", "in1", "inc", "layer_end", "outputs"):
        assert np.array_equal(getattr(c, f), getattr(py, f)), f


@pytest.mark.parametrize("in_bytes,out_bytes", [(0, 32), (1, 8), (8, 16), (13, 40)])
def test_native_ascon_circuit(in_bytes, out_bytes):
    c = K.native_ascon_xof(in_bytes, out_bytes)
    py = S.xof_circuit(in_bytes, out_bytes)
    for f in ("ops", "in0", "in1", "inc", "layer_end", "outputs"):
        assert np.array_equal(getattr(c, f), getattr(py, f)), f
    rng = np.random.default_rng(in_bytes * 7 + out_bytes)
    msg = bytes(rng.integers(0, 256, in_bytes, dtype=np.uint8).tolist())
    inp = np.concatenate([A.bits(msg), np.array([0, 1], np.uint8)])
    assert A.from_bits(c.evaluate(inp)) == AR.ascon_xof(msg, out_bytes)
    if in_bytes == 0:
        assert A.from_bits(c.evaluate(inp)).hex() == (
            "5d4cbde6350ea4c174bd65b5b332f8408f99740b81aa02735eaefbcf0ba0339e")


def test_native_circuit_argument_errors():
    L = C.lib()
    sz = K._Sizes()
    assert L.cufhe_circuit_ascon_xof(-1, 8, ctypes.byref(sz), None, None, None) == -1
    assert L.cufhe_circuit_ascon_xof(8, 0, ctypes.byref(sz), None, None, None) == -1
    assert L.cufhe_circuit_ascon_xof(1025, 8, ctypes.byref(sz), None, None, None) == -1
    assert L.cufhe_circuit_aes128_decrypt(None, None, None, None) == -1
    # buffers smaller than the circuit are refused
    assert L.cufhe_circuit_ascon_xof(1, 8, ctypes.byref(sz), None, None, None) == 0
    sz.n_gates -= 1
    g = np.zeros((sz.n_gates + 1, 5), np.uint32)
    le = np.zeros(sz.n_layers, np.int32)
    out = np.zeros(sz.n_outputs, np.uint32)
    assert L.cufhe_circuit_ascon_xof(1, 8, ctypes.byref(sz), g.ctypes.data, le.ctypes.data,
                                     out.ctypes.data) == -2
    # only some of the buffers is an argument error
    assert L.cufhe_circuit_ascon_xof(1, 8, ctypes.byref(sz), g.ctypes.data, None, None) == -1
    # ... except the gate buffer of a circuit without gates (a one-block squeeze)
    assert L.cufhe_circuit_ascon_squeeze(8, ctypes.byref(sz), None, None, None) == 0
    assert sz.n_gates == 0
    le = np.zeros(sz.n_layers, np.int32)
    out = np.zeros(sz.n_outputs, np.uint32)
    assert L.cufhe_circuit_ascon_squeeze(8, ctypes.byref(sz), None, le.ctypes.data,
                                         out.ctypes.data) == 0
    assert list(out[:64]) == [8 * (7 - j) + b for j in range(8) for b in range(8)]


@pytest.mark.parametrize("in_bytes,out_bytes", [(0, 32), (3, 8), (16, 24)])
def test_native_ascon_absorb_squeeze_compose_to_the_xof(in_bytes, out_bytes):
    """Initialize (the library's public state) -> absorb -> squeeze equals
    the oracle's XOF; the squeeze's state output is the state after its
    permutations (the reference permutes between blocks only)."""
    words = (ctypes.c_uint64 * 5)()
    assert C.lib().cufhe_ascon_xof_initial_state(words) == 0
    state = np.array([(words[i] >> k) & 1 for i in range(5) for k in range(64)], np.uint8)
    assert list(words) == AR.permutation([AR.XOF_IV, 0, 0, 0, 0])
    rng = np.random.default_rng(in_bytes + 1)
    msg = bytes(rng.integers(0, 256, in_bytes, dtype=np.uint8).tolist())
    ab = K.native_ascon_absorb(in_bytes)
    assert ab.n_inputs == 320 + 8 * in_bytes + 2 and len(ab.outputs) == 320 and _layered(ab)
    state = ab.evaluate(np.concatenate([state, A.bits(msg), [0, 1]]))
    sq = K.native_ascon_squeeze(out_bytes)
    assert sq.n_inputs == 322 and len(sq.outputs) == 8 * out_bytes + 320
    res = sq.evaluate(np.concatenate([state, [0, 1]])) if sq.n_gates else \
        np.concatenate([state, [0, 1]])[sq.outputs]
    assert A.from_bits(res[:8 * out_bytes]) == AR.ascon_xof(msg, out_bytes)
    if out_bytes <= 8:
        assert sq.n_gates == 0 and np.array_equal(res[8 * out_bytes:], state)
