"""Ascon v1.2 permutation, Ascon-Hash and Ascon-Xof on the CPU -- TEST
INFRASTRUCTURE: the expected outputs of the ASCON-XOF circuit workload in
tests/ and the bench. Never imported by the product package (cufhe_amd/).

Restated from the Ascon v1.2 specification (Dobraunig, Eichlseder, Mendel,
Schlaeffer; NIST lightweight-cryptography finalist): 320-bit state of five
64-bit words, p^12 with round constants 0xf0, 0xe1, ..., 0x4b added to x2,
the bitsliced 5-bit S-box, the linear layer with rotations (19, 28),
(61, 39), (1, 6), (10, 17), (7, 41); hashing with rate 64 bits, padding
0x80 then zeros, big-endian words. Pinned by the published empty-message
digests of Ascon-Hash and Ascon-Xof (tests/test_ascon_circuit.py).
The reference's ASCON-XOF path (src/ascon_gpu.cu:1040-1104) takes the state
layout and constants from TFHEpp, which is absent from the reference checkout.
"""

M64 = (1 << 64) - 1
ROUND_CONSTANTS = [0xF0, 0xE1, 0xD2, 0xC3, 0xB4, 0xA5, 0x96, 0x87, 0x78, 0x69, 0x5A, 0x4B]
ROTATIONS = [(19, 28), (61, 39), (1, 6), (10, 17), (7, 41)]
HASH_IV = 0x00400C0000000100
XOF_IV = 0x00400C0000000000


def _rotr(x, n):
    return ((x >> n) | (x << (64 - n))) & M64


def permutation(state, rounds=12):
    x = list(state)
    for r in range(12 - rounds, 12):
        x[2] ^= ROUND_CONSTANTS[r]
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
        for i, (a, b) in enumerate(ROTATIONS):
            x[i] ^= _rotr(x[i], a) ^ _rotr(x[i], b)
    return x


def _sponge(iv, msg, outlen):
    s = permutation([iv, 0, 0, 0, 0])
    m = bytes(msg) + b"\x80" + b"\x00" * ((8 - (len(msg) + 1) % 8) % 8)
    for i in range(0, len(m), 8):
        s[0] ^= int.from_bytes(m[i:i + 8], "big")
        if i + 8 < len(m):
            s = permutation(s)
    s = permutation(s)
    out = b""
    while len(out) < outlen:
        out += s[0].to_bytes(8, "big")
        if len(out) < outlen:
            s = permutation(s)
    return out[:outlen]


def ascon_hash(msg):
    return _sponge(HASH_IV, msg, 32)


def ascon_xof(msg, outlen):
    return _sponge(XOF_IV, msg, outlen)
