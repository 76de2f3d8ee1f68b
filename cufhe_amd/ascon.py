"""ASCON-XOF as a layered Boolean circuit for the serving loop (reference
src/ascon_gpu.cu:1040-1104 ASCONXOFInitialize / Absorb / Squeeze and
:1239-1282 ASCONXOF: encrypted message bits in, encrypted output bits out).

The reference evaluates the permutation's S-boxes with circuit bootstrapping
and ROM lookups (its state layout and constants come from TFHEpp). Here the
whole XOF is a circuit of the engine's own gates:

* Initialize: the state after p^12 of (IV, 0, 0, 0, 0) is public, so it
  enters as constants (the reference permutes the trivially encrypted
  state, which yields the same values).
* Absorb (rate 64 bits): message byte j of a block is XORed into byte j of
  x0 (most significant byte first); full blocks are followed by p^12; the
  last, partial or empty, block gets the padding byte 0x80 at its length
  (a NOT of one bit) and p^12.
* Squeeze: 8 output bytes from x0, then p^12 while more than 8 remain.
* p^12: round constant into x2 (NOTs), the bitsliced S-box (x0 ^= x4,
  x4 ^= x3, x2 ^= x1, x_i ^= ~x_(i+1) & x_(i+2), x1 ^= x0, x0 ^= x4,
  x3 ^= x2, x2 = ~x2) and the linear layer x_i ^= (x_i >>> a_i) ^ (x_i >>> b_i)
  with (a, b) = (19, 28), (61, 39), (1, 6), (10, 17), (7, 41) (Ascon v1.2).

Negations are carried as flags on wires and folded into the consuming gate
(XOR of negated operands, ANDNY / ANDYN / NOR for the S-box products), and
constants propagate, so the circuit has no NOT gates except at outputs.

Wire layout: inputs [0, 8 L) the message (bit b of byte i is wire 8 i + b,
least significant bit first, the reference's byte-major order), then one
encryption of 0 and one of 1 (aes.trivial_bits builds them); outputs in the
same bit order.
"""
from .aes import _Builder, _Lits
from .circuits import Circuit  # noqa: F401  (the return type)

M64 = (1 << 64) - 1
_RC = [0xF0, 0xE1, 0xD2, 0xC3, 0xB4, 0xA5, 0x96, 0x87, 0x78, 0x69, 0x5A, 0x4B]
_ROT = [(19, 28), (61, 39), (1, 6), (10, 17), (7, 41)]
_XOF_IV = 0x00400C0000000000


def _rotr(x, n):
    return ((x >> n) | (x << (64 - n))) & M64


def _p12_plain(x):
    x = list(x)
    for r in range(12):
        x[2] ^= _RC[r]
        x[0] ^= x[4]
        x[4] ^= x[3]
        x[2] ^= x[1]
        t = [(~x[i] & M64) & x[(i + 1) % 5] for i in range(5)]
        for i in range(5):
            x[i] ^= t[(i + 1) % 5]
        x[1] ^= x[0]
        x[0] ^= x[4]
        x[3] ^= x[2]
        x[2] = ~x[2] & M64
        for i, (a, b) in enumerate(_ROT):
            x[i] ^= _rotr(x[i], a) ^ _rotr(x[i], b)
    return x


def _permutation(L, x):
    """p^12 on five words of 64 literals (bit k = 2^k)."""
    for r in range(12):
        x[2] = [L.neg(v) if (_RC[r] >> k) & 1 else v for k, v in enumerate(x[2])]
        for k in range(64):
            s = [x[i][k] for i in range(5)]
            s[0] = L.xor(s[0], s[4])
            s[4] = L.xor(s[4], s[3])
            s[2] = L.xor(s[2], s[1])
            t = [L.and_(L.neg(s[i]), s[(i + 1) % 5]) for i in range(5)]
            for i in range(5):
                s[i] = L.xor(s[i], t[(i + 1) % 5])
            s[1] = L.xor(s[1], s[0])
            s[0] = L.xor(s[0], s[4])
            s[3] = L.xor(s[3], s[2])
            s[2] = L.neg(s[2])
            for i in range(5):
                x[i][k] = s[i]
        for i, (a, b) in enumerate(_ROT):
            w = x[i]
            x[i] = [L.xor(L.xor(w[k], w[(k + a) % 64]), w[(k + b) % 64]) for k in range(64)]
    return x


def _pos(j, bit):
    """Word bit of rate byte j (most significant byte first), bit `bit`."""
    return 8 * (7 - j) + bit


def xof_circuit(in_bytes, out_bytes):
    """ASCON-XOF of an in_bytes message into out_bytes output bytes."""
    if in_bytes < 0 or out_bytes <= 0:
        raise ValueError("message length >= 0 and output length > 0 (bytes)")
    n_inputs = 8 * in_bytes + 2
    c0, c1 = 8 * in_bytes, 8 * in_bytes + 1
    b = _Builder(n_inputs)
    L = _Lits(b, c0, c1)
    s0 = _p12_plain([_XOF_IV, 0, 0, 0, 0])
    x = [[L.const((s0[i] >> k) & 1) for k in range(64)] for i in range(5)]
    msg = lambda i: (i, False)  # noqa: E731  (message bit i as a literal)
    left, off = in_bytes, 0
    while left >= 8:
        for j in range(8):
            for bit in range(8):
                x[0][_pos(j, bit)] = L.xor(x[0][_pos(j, bit)], msg(8 * (off + j) + bit))
        x = _permutation(L, x)
        off += 8
        left -= 8
    for j in range(left):
        for bit in range(8):
            x[0][_pos(j, bit)] = L.xor(x[0][_pos(j, bit)], msg(8 * (off + j) + bit))
    x[0][_pos(left, 7)] = L.neg(x[0][_pos(left, 7)])  # padding byte 0x80
    x = _permutation(L, x)
    outs, left = [], out_bytes
    while True:
        n = min(left, 8)
        outs += [L.wire(x[0][_pos(j, bit)]) for j in range(n) for bit in range(8)]
        left -= n
        if left == 0:
            break
        x = _permutation(L, x)
    if not b.gates:
        # an empty message: the output is public (constants); linear COPY
        # gates give the server a circuit with one layer
        copy = {w: b.gate("COPY", w) for w in sorted(set(outs))}
        outs = [copy[w] for w in outs]
    return b.circuit(outs)
