"""AES-128 decryption circuit (cufhe_amd/aes.py, 29,604 gates, 189 layers)
run by the serving loop on encrypted inputs: the FIPS-197 Appendix C.1
block and a random block under a random key, submitted together (one
batch per round mixes both), decrypt to the plaintext (oracle/aes_ref.py)
and both outputs equal the same circuits run one at a time."""
import numpy as np
import pytest

import cufhe_amd as C
from cufhe_amd import aes as A
from conftest import load_lvl1_keys
from oracle import aes_ref as R

pytestmark = pytest.mark.gpu

KEY = bytes(range(16))
CT = bytes.fromhex("69c4e0d86a7b0430d8cdb78070b4c55a")
PT = bytes.fromhex("00112233445566778899aabbccddeeff")


def _inputs(k1, seed, ct, key):
    bits = A.input_bits(ct, A.round_keys(key))
    enc = C.encrypt_bits(1, seed, k1, bits[:-2])
    return np.concatenate([enc, A.trivial_bits(bits[-2:])])


def test_aes_decryption_circuit_on_the_server(keys_lvl1):
    C.check(C.lib().cufhe_initialize(), "initialize")
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    k1 = keys_lvl1["k1"]
    circ = A.decrypt_circuit()
    rng = np.random.default_rng(9)
    key2 = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
    pt2 = bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist())
    ct2 = R.encrypt_block(key2, pt2)
    in1, in2 = _inputs(k1, 71, CT, KEY), _inputs(k1, 72, ct2, key2)
    wires = circ.n_inputs + circ.n_gates
    srv = C.Server(pool_bytes=3 * wires * C.TLWE1_WORDS * 4, max_batch_bootstraps=16384)
    try:
        print("aes: submitting two blocks", flush=True)
        r1, r2 = srv.submit(circ, in1), srv.submit(circ, in2)
        mixed1, mixed2 = srv.result(r1), srv.result(r2)
        print("aes: mixed done, rounds", srv.info(r1)["rounds"], flush=True)
        alone1 = srv.result(srv.submit(circ, in1))
        print("aes: alone done", flush=True)
    finally:
        srv.destroy()
    assert A.from_bits(C.decrypt_bits(1, k1, mixed1)) == PT
    assert A.from_bits(C.decrypt_bits(1, k1, mixed2)) == pt2
    assert np.array_equal(mixed1, alone1)
