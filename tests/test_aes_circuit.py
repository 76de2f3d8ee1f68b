"""AES-128 decryption as a gate circuit (cufhe_amd/aes.py) on the CPU: the
oracle's AES (oracle/aes_ref.py) against FIPS-197 known answers, the
circuit's table against the oracle's, and the circuit's plaintext evaluation
against the oracle's decryption. No GPU."""
import numpy as np

from cufhe_amd import aes as A
from oracle import aes_ref as R

KEY = bytes(range(16))
PT = bytes.fromhex("00112233445566778899aabbccddeeff")
CT = bytes.fromhex("69c4e0d86a7b0430d8cdb78070b4c55a")  # FIPS-197 Appendix C.1


def test_oracle_aes_fips197_known_answers():
    assert R.encrypt_block(KEY, PT) == CT
    assert R.decrypt_block(KEY, CT) == PT
    # FIPS-197 Appendix A.1: the last round key of 2b7e1516...
    rk = R.expand_key(bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c"))
    assert rk[10].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"
    # FIPS-197 5.1.1 example and Figures 7 / 14 entries
    assert R.SBOX[0x53] == 0xED and R.SBOX[0x00] == 0x63 and R.SBOX[0xFF] == 0x16
    assert R.INV_SBOX[0x00] == 0x52 and R.INV_SBOX[0xED] == 0x53
    assert sorted(R.SBOX) == list(range(256))


def test_circuit_table_is_the_inverse_sbox():
    assert A.INV_SBOX == R.INV_SBOX


def test_tower_field_isomorphism_and_karatsuba():
    """The map into GF((2^4)^2) is a field isomorphism (all 65,536 products),
    and the Karatsuba products give every GF(2^4) product."""
    to = A._to_tower
    assert sorted(to(x) for x in range(256)) == list(range(256))
    for x in range(256):
        tx = to(x)
        for y in range(256):
            assert to(R.gf_mul(x, y)) == A._mul8(tx, to(y))
    for a in range(16):
        for b in range(16):
            prods = [A._parity(a & m) & A._parity(b & m) for m in A._KARA]
            got = sum((sum(prods[t] for t in A._KARA_SUB[k]) & 1) << k for k in range(4))
            assert got == A._mul4(a, b)


def test_inv_sbox_circuit_exhaustive():
    """The S-box sub-circuit of the decryption over all 256 inputs."""
    c = A.inv_sbox_circuit()
    for y in range(256):
        out = c.evaluate(np.concatenate([A.bits([y]), [0, 1]]))
        assert A.from_bits(out)[0] == R.INV_SBOX[y], y


def test_inv_sbox_circuit_catches_a_wrong_isomorphism(monkeypatch):
    """A basis image off by one bit (a plausible slip) breaks the S-box."""
    monkeypatch.setattr(A, "_PLAN", None)
    bad = list(A._BASIS)
    bad[3] ^= 1
    monkeypatch.setattr(A, "_BASIS", bad)
    monkeypatch.setattr(A, "_FROM_TOWER", {A._to_tower(x): x for x in range(256)})
    c = A.inv_sbox_circuit()
    got = [A.from_bits(c.evaluate(np.concatenate([A.bits([y]), [0, 1]])))[0] for y in range(256)]
    assert got != R.INV_SBOX


def test_host_key_expansion_matches_the_oracle():
    k = bytes.fromhex("2b7e151628aed2a6abf7158809cf4f3c")
    assert A.round_keys(k) == R.expand_key(k)
    assert A.round_keys(k)[10].hex() == "d014f9a8c9ee2589e13f0cc8b6630ca6"


def test_circuit_decrypts_fips_vector_and_random_blocks():
    c = A.decrypt_circuit()
    assert c.n_inputs == A.N_INPUTS and len(c.outputs) == 128
    out = c.evaluate(A.input_bits(CT, R.expand_key(KEY)))
    assert A.from_bits(out) == PT
    rng = np.random.default_rng(5)
    for _ in range(3):
        k = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
        p = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
        ct = R.encrypt_block(k, p)
        assert A.from_bits(c.evaluate(A.input_bits(ct, R.expand_key(k)))) == p


def test_circuit_is_layered():
    """Every gate reads only inputs and wires of earlier layers."""
    c = A.decrypt_circuit()
    layer_of = np.zeros(c.n_inputs + c.n_gates, np.int64)
    start = 0
    for l, end in enumerate(c.layer_end):
        layer_of[c.n_inputs + start:c.n_inputs + end] = l + 1
        start = end
    assert start == c.n_gates
    g = np.arange(c.n_gates)
    me = layer_of[c.n_inputs + g]
    assert np.all(layer_of[c.in0] < me) and np.all(layer_of[c.in1] < me)
    mux = np.isin(c.ops, [A.GATES["MUX"], A.GATES["NMUX"]])
    assert np.all(layer_of[c.inc[mux]] < me[mux])


def test_trivial_constants_decrypt():
    import cufhe_amd as C
    k0, k1, _ = C.secret_key_gen(3)
    assert list(C.decrypt_bits(1, k1, A.trivial_bits([0, 1, 1, 0]))) == [0, 1, 1, 0]
