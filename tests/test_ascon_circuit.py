"""ASCON-XOF as a gate circuit (cufhe_amd/ascon.py) on the CPU: the oracle's
Ascon (oracle/ascon_ref.py) against the published empty-message digests, and
the circuit's plaintext evaluation against the oracle's XOF for empty,
partial-block, whole-block and multi-block messages and outputs. No GPU."""
import numpy as np
import pytest

from cufhe_amd import aes as A
from cufhe_amd import ascon as S
from oracle import ascon_ref as R


def test_oracle_ascon_published_digests():
    # Ascon v1.2 known-answer tests, empty message
    assert R.ascon_hash(b"").hex() == (
        "7346bc14f036e87ae03d0997913088f5f68411434b3cf8b54fa796a80d251f91")
    assert R.ascon_xof(b"", 32).hex() == (
        "5d4cbde6350ea4c174bd65b5b332f8408f99740b81aa02735eaefbcf0ba0339e")


@pytest.mark.parametrize("in_bytes,out_bytes", [(0, 32), (1, 8), (7, 9), (8, 16), (13, 40)])
def test_circuit_matches_the_oracle_xof(in_bytes, out_bytes):
    c = S.xof_circuit(in_bytes, out_bytes)
    assert c.n_inputs == 8 * in_bytes + 2 and len(c.outputs) == 8 * out_bytes
    rng = np.random.default_rng(in_bytes * 100 + out_bytes)
    for _ in range(2):
        msg = bytes(rng.integers(0, 256, in_bytes, dtype=np.uint8).tolist())
        inp = np.concatenate([A.bits(msg), np.array([0, 1], np.uint8)])
        assert A.from_bits(c.evaluate(inp)) == R.ascon_xof(msg, out_bytes)


def test_circuit_has_no_inner_not_gates():
    c = S.xof_circuit(8, 16)
    nots = np.flatnonzero(c.ops == A.GATES["NOT"])
    # negations are folded into the consumers: a NOT only feeds an output
    assert set((c.n_inputs + nots).tolist()) <= set(c.outputs.tolist())
