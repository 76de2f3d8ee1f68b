"""ASCON-XOF circuit (cufhe_amd/ascon.py) run by the serving loop on
encrypted message bits: two 8-byte messages with 16-byte outputs (three
p^12, 210 layers), submitted together, decrypt to the oracle's XOF
(oracle/ascon_ref.py, pinned by the published digests), and the mixed run
equals a lone run bit for bit."""
import numpy as np
import pytest

import cufhe_amd as C
from cufhe_amd import aes as A
from cufhe_amd import ascon as S
from conftest import load_lvl1_keys
from oracle import ascon_ref as R

pytestmark = pytest.mark.gpu


def test_ascon_xof_circuit_on_the_server(keys_lvl1):
    C.check(C.lib().cufhe_initialize(), "initialize")
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    k1 = keys_lvl1["k1"]
    circ = S.xof_circuit(8, 16)
    msgs = [b"TFHE-gpu", bytes(range(200, 208))]
    ins = [np.concatenate([C.encrypt_bits(1, 80 + i, k1, A.bits(m)), A.trivial_bits([0, 1])])
           for i, m in enumerate(msgs)]
    wires = circ.n_inputs + circ.n_gates
    srv = C.Server(pool_bytes=3 * wires * C.TLWE1_WORDS * 4, max_batch_bootstraps=8192)
    try:
        print("ascon: submitting two messages", flush=True)
        rids = [srv.submit(circ, x) for x in ins]
        mixed = [srv.result(r) for r in rids]
        print("ascon: mixed done", flush=True)
        alone = srv.result(srv.submit(circ, ins[0]))
    finally:
        srv.destroy()
    for m, out in zip(msgs, mixed):
        assert A.from_bits(C.decrypt_bits(1, k1, out)) == R.ascon_xof(m, 16)
    assert np.array_equal(mixed[0], alone)
