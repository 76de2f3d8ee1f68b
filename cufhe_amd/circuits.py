"""Layered Boolean circuits for the serving loop: random generation from a
seed and plaintext evaluation (the expected decryptions).

Wires [0, n_inputs) are the inputs; gate g writes wire n_inputs + g. Layer l
holds `width` gates whose operands are wires of layer l - 1 (the inputs for
l = 0); the outputs are the last layer's wires.
"""
import ctypes

import numpy as np

from . import GATES, TRUTH

BINARY = ["NAND", "OR", "ORYN", "ORNY", "AND", "ANDYN", "ANDNY", "NOR", "XOR", "XNOR"]
_NAME = {v: k for k, v in GATES.items()}


class Circuit:
    def __init__(self, n_inputs, ops, in0, in1, inc, layer_end, outputs):
        self.n_inputs = int(n_inputs)
        self.ops = np.asarray(ops, np.uint32)
        self.in0 = np.asarray(in0, np.uint32)
        self.in1 = np.asarray(in1, np.uint32)
        self.inc = np.asarray(inc, np.uint32)
        self.layer_end = np.asarray(layer_end, np.int32)
        self.outputs = np.asarray(outputs, np.uint32)

    @property
    def n_gates(self):
        return len(self.ops)

    @property
    def depth(self):
        return len(self.layer_end)

    @property
    def bootstraps(self):
        mux = np.isin(self.ops, [GATES["MUX"], GATES["NMUX"]])
        lin = np.isin(self.ops, [GATES["NOT"], GATES["COPY"]])
        return int(2 * mux.sum() + (~mux & ~lin).sum())

    def evaluate(self, bits):
        """Plaintext outputs for input bits (the expected decryptions)."""
        w = np.zeros(self.n_inputs + self.n_gates, np.uint8)
        w[: self.n_inputs] = np.asarray(bits, np.uint8)
        for g in range(self.n_gates):
            name = _NAME[int(self.ops[g])]
            x, y, s = int(w[self.in0[g]]), int(w[self.in1[g]]), int(w[self.inc[g]])
            if name == "MUX":
                v = y if s else x
            elif name == "NMUX":
                v = 1 - (y if s else x)
            elif name == "NOT":
                v = 1 - x
            elif name == "COPY":
                v = x
            else:
                v = TRUTH[name](x, y)
            w[self.n_inputs + g] = v
        return w[self.outputs]


def random_circuit(rng, width, depth, mux_frac=0.1):
    """width gates per layer, depth layers; gate type MUX with probability
    mux_frac, otherwise uniform over the ten binary gates."""
    n = width * depth
    is_mux = rng.random(n) < mux_frac
    binops = np.array([GATES[b] for b in BINARY], np.uint32)[rng.integers(0, len(BINARY), n)]
    ops = np.where(is_mux, GATES["MUX"], binops).astype(np.uint32)
    in0 = np.zeros(n, np.uint32)
    in1 = np.zeros(n, np.uint32)
    inc = np.zeros(n, np.uint32)
    for l in range(depth):
        lo = 0 if l == 0 else width + (l - 1) * width  # first wire of the previous layer
        sl = slice(l * width, (l + 1) * width)
        in0[sl] = lo + rng.integers(0, width, width)
        in1[sl] = lo + rng.integers(0, width, width)
        inc[sl] = np.where(is_mux[sl], lo + rng.integers(0, width, width), 0)
    layer_end = np.arange(1, depth + 1, dtype=np.int32) * width
    outputs = np.arange(width + (depth - 1) * width, width + depth * width, dtype=np.uint32)
    return Circuit(width, ops, in0, in1, inc, layer_end, outputs)


class _Sizes(ctypes.Structure):
    _fields_ = [("n_inputs", ctypes.c_int), ("n_gates", ctypes.c_int),
                ("n_layers", ctypes.c_int), ("n_outputs", ctypes.c_int)]


def _native(fn, *args):
    """A circuit built by the library (include/cufhe_circuits.h): sizes
    first, then the arrays."""
    from . import check
    sz = _Sizes()
    check(fn(*args, ctypes.byref(sz), None, None, None), fn.__name__)
    gates = np.zeros((sz.n_gates, 5), np.uint32)
    layer_end = np.zeros(sz.n_layers, np.int32)
    outputs = np.zeros(sz.n_outputs, np.uint32)
    check(fn(*args, ctypes.byref(sz), gates.ctypes.data, layer_end.ctypes.data,
             outputs.ctypes.data), fn.__name__)
    return Circuit(sz.n_inputs, gates[:, 0].view(np.int32).astype(np.uint32), gates[:, 2],
                   gates[:, 3], gates[:, 4], layer_end, outputs)


def native_aes128_decrypt():
    """AES-128 decryption as built by the library's C++ builder (the circuit
    behind the C++ AESDec), in aes.py's wire layout."""
    from . import lib
    return _native(lib().cufhe_circuit_aes128_decrypt)


def native_ascon_xof(in_bytes, out_bytes):
    """ASCON-XOF as built by the library's C++ builder (behind the C++
    ASCONXOF), in ascon.py's wire layout."""
    from . import lib
    return _native(lib().cufhe_circuit_ascon_xof, int(in_bytes), int(out_bytes))


def native_ascon_absorb(in_bytes):
    """The XOF's absorb over an encrypted state (cufhe_circuit_ascon_absorb):
    inputs the 320 state bits, the message, 0 and 1; outputs the state."""
    from . import lib
    return _native(lib().cufhe_circuit_ascon_absorb, int(in_bytes))


def native_ascon_squeeze(out_bytes):
    """The XOF's squeeze (cufhe_circuit_ascon_squeeze): inputs the 320 state
    bits, 0 and 1; outputs the output bits, then the state."""
    from . import lib
    return _native(lib().cufhe_circuit_ascon_squeeze, int(out_bytes))
