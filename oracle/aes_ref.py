"""AES-128 on the CPU (FIPS-197) -- TEST INFRASTRUCTURE.

The plaintext reference for the AES-decryption circuit workload: the
expected decryptions in tests/ and the bench. It is never imported by the
product package (cufhe_amd/), which builds its circuit from its own tables.

Restated from FIPS-197 (Advanced Encryption Standard, NIST 2001): the S-box
is the multiplicative inverse in GF(2^8) = GF(2)[x]/(x^8 + x^4 + x^3 + x + 1)
followed by the affine map b_i ^ b_(i+4) ^ b_(i+5) ^ b_(i+6) ^ b_(i+7) ^ c_i
with c = 0x63 (FIPS-197 5.1.1); the key expansion (5.2), Cipher (5.1) and
InvCipher (5.3). Pinned by the FIPS-197 Appendix C.1 known-answer vector and
by S-box entries listed in FIPS-197 Figure 7 (tests/test_aes_circuit.py).
The reference's own AES path (src/aes_gpu.cu) takes the S-box from TFHEpp,
which is absent from the reference checkout.
"""


def gf_mul(a, b):
    """Product in GF(2^8) modulo x^8 + x^4 + x^3 + x + 1."""
    r = 0
    for _ in range(8):
        if b & 1:
            r ^= a
        hi = a & 0x80
        a = (a << 1) & 0xFF
        if hi:
            a ^= 0x1B
        b >>= 1
    return r


def gf_inv(a):
    """Multiplicative inverse (0 -> 0): a^254."""
    if a == 0:
        return 0
    r, e, base = 1, 254, a
    while e:
        if e & 1:
            r = gf_mul(r, base)
        base = gf_mul(base, base)
        e >>= 1
    return r


def _affine(b):
    r = 0
    for i in range(8):
        bit = ((b >> i) ^ (b >> ((i + 4) % 8)) ^ (b >> ((i + 5) % 8)) ^ (b >> ((i + 6) % 8)) ^
               (b >> ((i + 7) % 8)) ^ (0x63 >> i)) & 1
        r |= bit << i
    return r


SBOX = [_affine(gf_inv(x)) for x in range(256)]
INV_SBOX = [0] * 256
for _x, _y in enumerate(SBOX):
    INV_SBOX[_y] = _x

RCON = [0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80, 0x1B, 0x36]


def expand_key(key):
    """FIPS-197 5.2: 16-byte key -> 11 round keys of 16 bytes."""
    w = [list(key[4 * i:4 * i + 4]) for i in range(4)]
    for i in range(4, 44):
        t = list(w[i - 1])
        if i % 4 == 0:
            t = t[1:] + t[:1]
            t = [SBOX[b] for b in t]
            t[0] ^= RCON[i // 4 - 1]
        w.append([w[i - 4][j] ^ t[j] for j in range(4)])
    return [bytes(sum(w[4 * r:4 * r + 4], [])) for r in range(11)]


# state as 16 bytes in FIPS-197 input order: byte 4c + r is row r, column c
def _shift_rows(s, inv):
    out = [0] * 16
    for c in range(4):
        for r in range(4):
            src = (c + r) % 4 if not inv else (c - r) % 4
            out[4 * c + r] = s[4 * src + r]
    return out


def _mix_columns(s, m):
    out = [0] * 16
    for c in range(4):
        col = s[4 * c:4 * c + 4]
        for r in range(4):
            v = 0
            for k in range(4):
                v ^= gf_mul(m[(k - r) % 4], col[k])
            out[4 * c + r] = v
    return out


MIX = [0x02, 0x03, 0x01, 0x01]
INV_MIX = [0x0E, 0x0B, 0x0D, 0x09]


def encrypt_block(key, block):
    rk = expand_key(key)
    s = [b ^ k for b, k in zip(block, rk[0])]
    for rnd in range(1, 11):
        s = [SBOX[b] for b in s]
        s = _shift_rows(s, False)
        if rnd < 10:
            s = _mix_columns(s, MIX)
        s = [b ^ k for b, k in zip(s, rk[rnd])]
    return bytes(s)


def decrypt_block(key, block):
    rk = expand_key(key)
    s = [b ^ k for b, k in zip(block, rk[10])]
    for rnd in range(9, -1, -1):
        s = _shift_rows(s, True)
        s = [INV_SBOX[b] for b in s]
        s = [b ^ k for b, k in zip(s, rk[rnd])]
        if rnd > 0:
            s = _mix_columns(s, INV_MIX)
    return bytes(s)
