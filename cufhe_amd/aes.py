"""AES-128 decryption as a layered Boolean circuit for the serving loop
(reference src/aes_gpu.cu:904-1118 AESDec: 128 encrypted ciphertext bits
and the encrypted expanded key in, 128 encrypted plaintext bits out).

The reference evaluates InvSubBytes with circuit bootstrapping and CMUX
lookups into ROM polynomials whose tables come from TFHEpp. Here the whole
decryption is a circuit of the engine's own gates, run by the continuous-
batching server (cufhe_amd.Server) or as layered plans:

  state = C ^ K10; for r = 9 .. 0: InvShiftRows (wire renaming),
  InvSubBytes, AddRoundKey(K_r), InvMixColumns (r > 0)       (FIPS-197 5.3)

* InvSubBytes: each of the 8 output bits of the inverse S-box is expanded
  over the input bits x0 .. x7 (Shannon expansion, x7 outermost), and the
  sub-functions are shared by truth table across the 8 outputs of an S-box.
  A node MUX(x_k, hi, lo) becomes one binary gate when a branch is constant
  (AND, ANDYN, ORNY, OR) or the branches are complements (XOR); functions of
  x0, x1 alone are one binary gate; the rest are MUX gates (two bootstraps).
* AddRoundKey: one XOR per bit; InvMixColumns: the GF(2)-linear map of each
  column (32 bits) as balanced XOR trees.
* The S-box is computed from its definition (inverse in GF(2^8) modulo
  x^8 + x^4 + x^3 + x + 1, then the affine map with 0x63; FIPS-197 5.1.1).

Wire layout: inputs [0, 128) the ciphertext, [128 + 128 r, 256 + 128 r) round
key r (r = 0 .. 10, FIPS-197 5.2), then one wire holding an encryption of 0
and one of 1 (trivial_bits builds them without the key). Bit b of byte i is
wire 8 i + b (least significant bit first); the 128 outputs follow the
same order.
"""
import numpy as np

from . import GATES, TLWE1_WORDS
from .circuits import Circuit

N_KEY_BITS = 11 * 128
N_INPUTS = 128 + N_KEY_BITS + 2
CONST0 = 128 + N_KEY_BITS
CONST1 = CONST0 + 1


def _gf_mul(a, b):
    r = 0
    for _ in range(8):
        if b & 1:
            r ^= a
        a = ((a << 1) ^ (0x1B if a & 0x80 else 0)) & 0xFF
        b >>= 1
    return r


def _inv_sbox():
    inv = [0] * 256
    for x in range(256):
        g = 0
        if x:
            g = next(y for y in range(1, 256) if _gf_mul(x, y) == 1)
        s = 0x63
        for i in range(8):
            bit = (g >> i) ^ (g >> ((i + 4) % 8)) ^ (g >> ((i + 5) % 8)) ^ \
                (g >> ((i + 6) % 8)) ^ (g >> ((i + 7) % 8))
            s ^= (bit & 1) << i
        inv[s] = x
    return inv


INV_SBOX = _inv_sbox()


class _Builder:
    """Gates with hash-consing; a gate's layer is one more than the latest
    layer among its operands (inputs are layer 0)."""

    def __init__(self, n_inputs):
        self.n_inputs = n_inputs
        self.gates = []   # (op, in0, in1, inc)
        self.level = [0] * n_inputs
        self.memo = {}

    def gate(self, op, a, b=0, c=0):
        key = (op, a, b, c)
        if op in ("AND", "OR", "XOR", "XNOR", "NAND", "NOR") and a > b:
            key = (op, b, a, c)
        w = self.memo.get(key)
        if w is None:
            w = self.n_inputs + len(self.gates)
            self.gates.append((GATES[key[0]], key[1], key[2], key[3]))
            ops = (key[1], key[2], key[3]) if op in ("MUX", "NMUX") else (key[1], key[2])
            self.level.append(1 + max(self.level[o] for o in ops))
            self.memo[key] = w
        return w

    def xor_tree(self, wires):
        wires = sorted(wires)
        while len(wires) > 1:
            nxt = [self.gate("XOR", wires[i], wires[i + 1]) for i in range(0, len(wires) - 1, 2)]
            if len(wires) % 2:
                nxt.append(wires[-1])
            wires = nxt
        return wires[0]

    def circuit(self, outputs):
        # gates in layer order; gate g then writes wire n_inputs + g
        order = sorted(range(len(self.gates)), key=lambda g: (self.level[self.n_inputs + g], g))
        remap = list(range(self.n_inputs)) + [0] * len(self.gates)
        for new, g in enumerate(order):
            remap[self.n_inputs + g] = self.n_inputs + new
        ops, in0, in1, inc, layer_end = [], [], [], [], []
        last = None
        for new, g in enumerate(order):
            op, a, b, c = self.gates[g]
            lv = self.level[self.n_inputs + g]
            if last is not None and lv != last:
                layer_end.append(new)
            last = lv
            ops.append(op)
            in0.append(remap[a])
            in1.append(remap[b])
            inc.append(remap[c])
        layer_end.append(len(order))
        return Circuit(self.n_inputs, ops, in0, in1, inc, layer_end, [remap[w] for w in outputs])


def _gate2(b, tt, x0, x1):
    """A function of (x0, x1) given by its 4-bit truth table (bit 2*x1 + x0)."""
    table = {
        0b1000: ("AND", x0, x1), 0b0111: ("NAND", x0, x1), 0b1110: ("OR", x0, x1),
        0b0001: ("NOR", x0, x1), 0b0110: ("XOR", x0, x1), 0b1001: ("XNOR", x0, x1),
        0b0010: ("ANDYN", x0, x1),  # x0 & !x1
        0b0100: ("ANDNY", x0, x1),  # !x0 & x1
        0b1011: ("ORYN", x0, x1),   # x0 | !x1
        0b1101: ("ORNY", x0, x1),   # !x0 | x1
    }
    if tt == 0b1100:
        return x1
    if tt == 0b0011:
        return b.gate("NOT", x1)
    op, a, c = table[tt]
    return b.gate(op, a, c)


def _sbox(b, xs):
    """Inverse S-box on 8 input wires (x0 least significant); 8 output wires."""
    memo = {}

    def build(tt, k):
        size = 1 << (k + 1)
        full = (1 << size) - 1
        if tt == 0:
            return CONST0
        if tt == full:
            return CONST1
        half = size // 2
        lo, hi = tt & ((1 << half) - 1), tt >> half
        if k > 0 and lo == hi:
            return build(lo, k - 1)
        key = (tt, k)
        if key in memo:
            return memo[key]
        if k == 0:
            w = xs[0] if tt == 0b10 else b.gate("NOT", xs[0])
        elif k == 1:
            w = _gate2(b, tt, xs[0], xs[1])
        else:
            s = xs[k]
            mask = (1 << half) - 1
            if lo == 0 and hi == mask:
                w = s
            elif lo == mask and hi == 0:
                w = b.gate("NOT", s)
            elif lo == 0:
                w = b.gate("AND", s, build(hi, k - 1))
            elif hi == 0:
                w = b.gate("ANDYN", build(lo, k - 1), s)
            elif lo == mask:
                w = b.gate("ORNY", s, build(hi, k - 1))
            elif hi == mask:
                w = b.gate("OR", s, build(lo, k - 1))
            elif hi == lo ^ mask:
                w = b.gate("XOR", s, build(lo, k - 1))
            else:
                w = b.gate("MUX", build(lo, k - 1), build(hi, k - 1), s)
        memo[key] = w
        return w

    outs = []
    for j in range(8):
        tt = 0
        for x in range(256):
            if (INV_SBOX[x] >> j) & 1:
                tt |= 1 << x
        outs.append(build(tt, 7))
    return outs


def _inv_mix_matrix():
    """rows[o] = the 32 input bits XORed into output bit o of a column."""
    coef = [0x0E, 0x0B, 0x0D, 0x09]
    rows = [[] for _ in range(32)]
    for ib in range(32):
        byte, bit = divmod(ib, 8)
        col = [0, 0, 0, 0]
        col[byte] = 1 << bit
        for r in range(4):
            v = 0
            for k in range(4):
                v ^= _gf_mul(coef[(k - r) % 4], col[k])
            for ob in range(8):
                if (v >> ob) & 1:
                    rows[8 * r + ob].append(ib)
    return rows


_CIRCUIT = None


def decrypt_circuit():
    """The AES-128 decryption circuit (built once per process)."""
    global _CIRCUIT
    if _CIRCUIT is not None:
        return _CIRCUIT
    b = _Builder(N_INPUTS)
    key = lambda r, i: 128 + 128 * r + i  # noqa: E731
    state = [b.gate("XOR", i, key(10, i)) for i in range(128)]
    mix = _inv_mix_matrix()
    for rnd in range(9, -1, -1):
        # InvShiftRows: byte (row r, column c) <- (row r, column c - r)
        shifted = [0] * 128
        for c in range(4):
            for r in range(4):
                src = 4 * ((c - r) % 4) + r
                for bit in range(8):
                    shifted[8 * (4 * c + r) + bit] = state[8 * src + bit]
        sub = []
        for byte in range(16):
            sub += _sbox(b, shifted[8 * byte:8 * byte + 8])
        state = [b.gate("XOR", sub[i], key(rnd, i)) for i in range(128)]
        if rnd > 0:
            nxt = []
            for c in range(4):
                col = state[32 * c:32 * c + 32]
                nxt += [b.xor_tree([col[i] for i in mix[o]]) for o in range(32)]
            state = nxt
    _CIRCUIT = b.circuit(state)
    return _CIRCUIT


def round_keys(key):
    """FIPS-197 5.2 key expansion (on the key owner's side, before the round
    keys are encrypted): 16-byte key -> 11 round keys of 16 bytes."""
    sbox = [0] * 256
    for x, y in enumerate(INV_SBOX):
        sbox[y] = x
    rcon = [0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1B, 0x36]
    w = [list(key[4 * i:4 * i + 4]) for i in range(4)]
    for i in range(4, 44):
        t = list(w[i - 1])
        if i % 4 == 0:
            t = [sbox[v] for v in t[1:] + t[:1]]
            t[0] ^= rcon[i // 4 - 1]
        w.append([w[i - 4][j] ^ t[j] for j in range(4)])
    return [bytes(sum(w[4 * r:4 * r + 4], [])) for r in range(11)]


def bits(data):
    """Bytes -> bits, least significant bit of each byte first."""
    a = np.frombuffer(bytes(data), np.uint8)
    return ((a[:, None] >> np.arange(8)) & 1).astype(np.uint8).reshape(-1)


def from_bits(bitvec):
    b = np.asarray(bitvec, np.uint8).reshape(-1, 8)
    return bytes((b << np.arange(8)).sum(axis=1).astype(np.uint8).tolist())


def input_bits(ciphertext, round_keys):
    """Plaintext input vector: ciphertext block, the 11 round keys, 0, 1."""
    assert len(ciphertext) == 16 and len(round_keys) == 11
    return np.concatenate([bits(ciphertext)] + [bits(k) for k in round_keys] +
                          [np.array([0, 1], np.uint8)])


def trivial_bits(values):
    """Noiseless lvl1 TLWEs of the given bits (a = 0, b = +-mu): constants
    that need no key."""
    v = np.asarray(values, np.uint8)
    out = np.zeros((len(v), TLWE1_WORDS), np.uint32)
    mu = np.uint32(1 << 29)
    out[:, -1] = np.where(v == 1, mu, np.uint32(0) - mu)
    return out
