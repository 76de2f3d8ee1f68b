"""Serving loop (cufhe_server_*): continuous batching of layered lvl1 circuits
with device-memory admission. Each circuit's result must be bit-identical
alone and in the mix, and decrypt to the plaintext evaluation."""
import threading
import time

import numpy as np
import pytest

import cufhe_amd as C
from cufhe_amd.circuits import random_circuit
from conftest import load_lvl1_keys

pytestmark = pytest.mark.gpu

W1 = C.TLWE1_WORDS


@pytest.fixture(scope="module")
def served(keys_lvl1):
    C.check(C.lib().cufhe_initialize(), "initialize")
    load_lvl1_keys(("fft", True), keys_lvl1["bk_kb"], keys_lvl1["ksk"],
                   backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    rng = np.random.default_rng(2024)
    circuits, inputs, bits = [], [], []
    for i in range(6):
        c = random_circuit(rng, int(rng.integers(16, 49)), int(rng.integers(2, 6)))
        b = rng.integers(0, 2, c.n_inputs).astype(np.uint8)
        circuits.append(c)
        bits.append(b)
        inputs.append(C.encrypt_bits(1, 500 + i, keys_lvl1["k1"], b))
    return keys_lvl1, circuits, inputs, bits


def test_alone_and_mixed_are_bit_identical(served):
    keys, circuits, inputs, bits = served
    k1 = keys["k1"]
    # alone: one circuit in the server at a time
    srv = C.Server(pool_bytes=64 << 20)
    alone = []
    for c, x in zip(circuits, inputs):
        rid = srv.submit(c, x)
        alone.append(srv.result(rid))
    alone_rounds = srv.stats()["rounds"]
    srv.destroy()
    for c, b, out in zip(circuits, bits, alone):
        assert np.array_equal(C.decrypt_bits(1, k1, out), c.evaluate(b))
    assert alone_rounds == sum(c.depth for c in circuits)

    # mixed: staggered arrivals from a client thread, a pool that holds only
    # some circuits at once (admission waits for memory), layers co-batched
    need = max((c.n_inputs + c.n_gates) for c in circuits) * W1 * 4
    srv = C.Server(pool_bytes=int(2.5 * need))
    ids = [None] * len(circuits)

    def client():
        for i, (c, x) in enumerate(zip(circuits, inputs)):
            ids[i] = srv.submit(c, x)
            time.sleep(0.003)

    t = threading.Thread(target=client)
    t.start()
    t.join()
    mixed = [srv.result(r) for r in ids]
    st = srv.stats()
    infos = [srv.info(r) for r in ids]
    srv.destroy()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m)
    # layers of different circuits shared rounds
    assert st["rounds"] < sum(c.depth for c in circuits)
    assert st["peak_pool_bytes"] <= int(2.5 * need)
    for inf in infos:
        assert inf["submit"] <= inf["admit"] <= inf["done"]


def test_submit_rejects_bad_circuits(served):
    _, circuits, inputs, _ = served
    c = circuits[0]
    srv = C.Server(pool_bytes=8 * W1 * 4)  # eight wires
    with pytest.raises(C.CufheError, match="larger than the wire pool"):
        srv.submit(c, inputs[0])
    bad = random_circuit(np.random.default_rng(1), 4, 2)
    bad.in0[0] = bad.n_inputs  # reads a wire of its own layer
    with pytest.raises(C.CufheError, match="same or a later layer"):
        srv.submit(bad, np.zeros((4, W1), np.uint32))
    srv.destroy()


def test_config5_sizes_poisson_alone_equals_mixed(served):
    """Circuits at the serving workload's sizes (width U[64, 1024], depth
    U[4, 32], 10% MUX, binary gates uniform) arriving as a Poisson stream:
    each result is bit-identical to the same circuit run alone, and every
    output decrypts to the plaintext evaluation. The pipelined worker keeps
    up to two rounds in flight, so consecutive layers of one circuit are
    issued back to back."""
    keys, _, _, _ = served
    k1 = keys["k1"]
    rng = np.random.default_rng(505)
    circuits, inputs, bits = [], [], []
    for i in range(8):
        c = random_circuit(rng, int(rng.integers(64, 1025)), int(rng.integers(4, 33)))
        b = rng.integers(0, 2, c.n_inputs).astype(np.uint8)
        circuits.append(c)
        bits.append(b)
        inputs.append(C.encrypt_bits(1, 700 + i, k1, b))
    print("circuits (width, depth):", [(c.n_inputs, c.depth) for c in circuits], flush=True)
    srv = C.Server(pool_bytes=1 << 30)
    alone = []
    for c, x in zip(circuits, inputs):
        alone.append(srv.result(srv.submit(c, x)))
    srv.destroy()
    for c, b, out in zip(circuits, bits, alone):
        assert np.array_equal(C.decrypt_bits(1, k1, out), c.evaluate(b))

    arrivals = np.cumsum(rng.exponential(1.0 / 40.0, len(circuits)))  # 40 circuits/s
    srv = C.Server(pool_bytes=1 << 30)
    ids = [None] * len(circuits)
    t0 = time.perf_counter()

    def client():
        for i, (c, x) in enumerate(zip(circuits, inputs)):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit(c, x)

    t = threading.Thread(target=client)
    t.start()
    t.join()
    mixed = [srv.result(r) for r in ids]
    st = srv.stats()
    srv.destroy()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m)
    assert st["rounds"] < sum(c.depth for c in circuits)
    assert 0 < st["busy_s"]
