"""AES-128 decryption as a layered Boolean circuit for the serving loop
(reference src/aes_gpu.cu:904-1118 AESDec: 128 encrypted ciphertext bits
and the encrypted expanded key in, 128 encrypted plaintext bits out).

The reference evaluates InvSubBytes with circuit bootstrapping and CMUX
lookups into ROM polynomials whose tables come from TFHEpp. Here the whole
decryption is a circuit of the engine's own gates, run by the continuous-
batching server (cufhe_amd.Server) or as layered plans:

  state = C ^ K10; for r = 9 .. 0: InvShiftRows (wire renaming),
  InvSubBytes, AddRoundKey(K_r), InvMixColumns (r > 0)       (FIPS-197 5.3)

* InvSubBytes is the inverse affine map followed by inversion in GF(2^8),
  computed in the tower field GF((2^4)^2) = GF(2^4)[Y]/(Y^2 + Y + lam) over
  GF(2^4) = GF(2)[z]/(z^4 + z + 1) (the composite-field method of Satoh et
  al., ASIACRYPT 2001, and Canright, CHES 2005): a Y + b has the inverse
  (a D) Y + (a + b) D with D = (lam a^2 + a b + b^2)^-1. An isomorphism
  from the AES field is fixed by a root of x^8 + x^4 + x^3 + x + 1 in the
  tower field. Per S-box:
    - one linear map from the input byte to every GF(2)-linear form the
      products need (inverse affine map, isomorphism, the Karatsuba operand
      sums of a and b, and lam a^2 + b^2);
    - 9 ANDs for a b (two-level Karatsuba), XORs to D's argument;
    - the 4-bit inverse by Shannon expansion;
    - 18 ANDs for a D and b D (Karatsuba again);
    - one linear map back (a D and a D + b D through the inverse
      isomorphism) that also adds the round key.
* InvMixColumns: the GF(2)-linear 32 x 32 map of each column.
* Linear maps are XOR networks from a greedy common-subexpression search
  (Paar, ISIT 1997) held to each output's minimum depth. Complements (the
  affine constants, negated inputs) ride on wires as flags and are folded
  into the consuming gate, so no NOT gate sits on the data path.
* The S-box is computed from its definition (inverse in GF(2^8) modulo
  x^8 + x^4 + x^3 + x + 1, then the affine map with 0x63; FIPS-197 5.1.1).

Wire layout: inputs [0, 128) the ciphertext, [128 + 128 r, 256 + 128 r) round
key r (r = 0 .. 10, FIPS-197 5.2), then one wire holding an encryption of 0
and one of 1 (trivial_bits builds them without the key). Bit b of byte i is
wire 8 i + b (least significant bit first); the 128 outputs follow the
same order.
"""
import bisect
import itertools

import numpy as np

from . import GATES, TLWE1_WORDS
from .circuits import Circuit

N_KEY_BITS = 11 * 128
N_INPUTS = 128 + N_KEY_BITS + 2
CONST0 = 128 + N_KEY_BITS
CONST1 = CONST0 + 1

_NAME = {v: k for k, v in GATES.items()}


def _gf_mul(a, b):
    """Product in GF(2^8) modulo x^8 + x^4 + x^3 + x + 1."""
    r = 0
    for _ in range(8):
        if b & 1:
            r ^= a
        a = ((a << 1) ^ (0x1B if a & 0x80 else 0)) & 0xFF
        b >>= 1
    return r


def _affine_linear(g):
    """The linear part of the S-box's affine map (FIPS-197 5.1.1, without
    the constant 0x63)."""
    s = 0
    for i in range(8):
        bit = (g >> i) ^ (g >> ((i + 4) % 8)) ^ (g >> ((i + 5) % 8)) ^ \
            (g >> ((i + 6) % 8)) ^ (g >> ((i + 7) % 8))
        s |= (bit & 1) << i
    return s


def _inv_sbox():
    inv = [0] * 256
    for x in range(256):
        g = next(y for y in range(1, 256) if _gf_mul(x, y) == 1) if x else 0
        inv[_affine_linear(g) ^ 0x63] = x
    return inv


INV_SBOX = _inv_sbox()


# ------------------------------------------------------------ tower field
def _mul4(a, b):
    """Product in GF(2^4) = GF(2)[z]/(z^4 + z + 1)."""
    r = 0
    for i in range(4):
        if (b >> i) & 1:
            r ^= a << i
    for i in (6, 5, 4):
        if (r >> i) & 1:
            r ^= 0b10011 << (i - 4)
    return r


_INV4 = [0] + [next(y for y in range(1, 16) if _mul4(x, y) == 1) for x in range(1, 16)]
# the first lam for which Y^2 + Y + lam has no root in GF(2^4)
_LAM = next(v for v in range(1, 16) if all(_mul4(t, t) ^ t ^ v for t in range(16)))


def _mul8(p, q):
    """Product in GF((2^4)^2); p = (a << 4) | b stands for a Y + b."""
    a, b, c, d = p >> 4, p & 15, q >> 4, q & 15
    ac = _mul4(a, c)
    return ((ac ^ _mul4(a, d) ^ _mul4(b, c)) << 4) | (_mul4(ac, _LAM) ^ _mul4(b, d))


def _tower_basis():
    """Images of x^0 .. x^7 under an isomorphism GF(2^8) -> GF((2^4)^2): the
    powers of the first root of x^8 + x^4 + x^3 + x + 1 in the tower field."""
    for beta in range(2, 256):
        pw = [1]
        for _ in range(8):
            pw.append(_mul8(pw[-1], beta))
        if pw[8] ^ pw[4] ^ pw[3] ^ pw[1] ^ pw[0] == 0:
            return pw[:8]
    raise AssertionError("x^8 + x^4 + x^3 + x + 1 has a root in every field of 256 elements")


_BASIS = _tower_basis()


def _to_tower(x):
    r = 0
    for i in range(8):
        if (x >> i) & 1:
            r ^= _BASIS[i]
    return r


_FROM_TOWER = {_to_tower(x): x for x in range(256)}
_AFFINE_INV = {_affine_linear(g): g for g in range(256)}


def _rows(f, n_in, n_out):
    """Rows of the GF(2)-linear map f: bit i of rows[k] is set when input
    bit i feeds output bit k."""
    return [sum(((f(1 << i) >> k) & 1) << i for i in range(n_in)) for k in range(n_out)]


def _parity(x):
    return bin(x).count("1") & 1


# two-level Karatsuba in GF(2^4): the products (sum_{i in m} a_i)(sum_{i in m} b_i)
# for these index sets m span a b over GF(2) with 9 ANDs instead of 16
_KARA = [0b0001, 0b0010, 0b0011, 0b0100, 0b1000, 0b1100, 0b0101, 0b1010, 0b1111]


def _kara_subsets():
    """For each bit k of a GF(2^4) product: the Karatsuba products whose XOR
    is that bit (a search over the 512 subsets of bilinear forms)."""
    forms = [[(m >> i) & (m >> j) & 1 for i in range(4) for j in range(4)] for m in _KARA]
    out = []
    for k in range(4):
        target = [(_mul4(1 << i, 1 << j) >> k) & 1 for i in range(4) for j in range(4)]
        for sub in range(1 << len(_KARA)):
            acc = [0] * 16
            for t in range(len(_KARA)):
                if (sub >> t) & 1:
                    acc = [x ^ y for x, y in zip(acc, forms[t])]
            if acc == target:
                out.append([t for t in range(len(_KARA)) if (sub >> t) & 1])
                break
    return out


_KARA_SUB = _kara_subsets()


# ------------------------------------------------------------ XOR networks
def _depth(levels):
    """Depth of a balanced XOR of wires at these levels."""
    h = sorted(levels)
    while len(h) > 1:
        x, y = h.pop(0), h.pop(0)
        bisect.insort(h, max(x, y) + 1)
    return h[0] if h else 0


def _splitmix64(seed):
    """Uniform doubles in [0, 1) from the splitmix64 sequence (Steele,
    Lea and Flood 2014): the tie-breaking stream of the XOR-network search,
    the same as the library's C++ builder uses."""
    mask = (1 << 64) - 1
    s = seed & mask
    while True:
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        yield (z >> 11) * 2.0 ** -53


def _xor_network(rows, n_in, in_level=None, tries=8):
    """XOR network computing the GF(2) rows (bit masks over n_in signals).

    Greedy common-subexpression search (Paar 1997): repeatedly XOR the pair
    of signals that occurs in the most rows, here only among pairs that keep
    every row at its minimum depth; ties are broken at random (one draw per
    candidate pair, in pair order) and the best of `tries` seeds (fewest
    XORs, then least depth) is kept. Returns (pairs, rowsigs): pairs[t]
    makes signal n_in + t, rowsigs[r] lists the signals XORed into row r."""
    base = list(in_level or [0] * n_in)
    best = None
    for seed in range(tries):
        rng = _splitmix64(seed + 1)
        lv = list(base)
        sets = [{i for i in range(n_in) if (m >> i) & 1} for m in rows]
        bound = [_depth([lv[i] for i in s]) for s in sets]
        pairs = []
        while True:
            count = {}
            for s in sets:
                for p in itertools.combinations(sorted(s), 2):
                    count[p] = count.get(p, 0) + 1
            cands = [(p, count[p], max(lv[p[0]], lv[p[1]]), next(rng))
                     for p in sorted(count) if count[p] >= 2]
            cands.sort(key=lambda c: (-c[1], c[2], c[3]))
            chosen = None
            for p, _, _, _ in cands:
                new = max(lv[p[0]], lv[p[1]]) + 1
                if all(_depth([lv[t] for t in s if t not in p] + [new]) <= bd
                       for s, bd in zip(sets, bound) if p[0] in s and p[1] in s):
                    chosen = p
                    break
            if chosen is None:
                break
            sig = n_in + len(pairs)
            pairs.append(chosen)
            lv.append(max(lv[chosen[0]], lv[chosen[1]]) + 1)
            for s in sets:
                if chosen[0] in s and chosen[1] in s:
                    s -= set(chosen)
                    s.add(sig)
        xors = len(pairs) + sum(max(0, len(s) - 1) for s in sets)
        depth = max(_depth([lv[t] for t in s]) for s in sets)
        if best is None or (xors, depth) < best[0]:
            best = ((xors, depth), (pairs, [sorted(s) for s in sets]))
    return best[1]


class _Lits:
    """Literals (wire, complemented) over a builder; (c0, False) is 0 and
    (c0, True) is 1. Complements ride on the literal and are folded into the
    gate that consumes it."""

    _COMPLEMENT = {"AND": "NAND", "NAND": "AND", "OR": "NOR", "NOR": "OR", "XOR": "XNOR",
                   "XNOR": "XOR", "ANDNY": "ORYN", "ORYN": "ANDNY", "ANDYN": "ORNY",
                   "ORNY": "ANDYN", "MUX": "NMUX", "NMUX": "MUX"}

    def __init__(self, b, c0, c1):
        self.b, self.c0, self.c1 = b, c0, c1

    def const(self, v):
        return (self.c0, bool(v))

    def neg(self, p):
        return (p[0], not p[1])

    def level(self, p):
        return self.b.level[p[0]]

    def xor(self, p, q):
        (a, na), (b, nb) = p, q
        if a == self.c0:
            return (b, nb ^ na)
        if b == self.c0:
            return (a, na ^ nb)
        if a == b:
            return (self.c0, na ^ nb)
        return (self.b.gate("XOR", a, b), na ^ nb)

    def xor_all(self, lits):
        """XOR of the literals, the two earliest first."""
        lits = sorted(lits, key=self.level)
        if not lits:
            return self.const(0)
        while len(lits) > 1:
            x, y = lits.pop(0), lits.pop(0)
            v = self.xor(x, y)
            keys = [self.level(t) for t in lits]
            lits.insert(bisect.bisect_right(keys, self.level(v)), v)
        return lits[0]

    def and_(self, p, q):
        (a, na), (b, nb) = p, q
        if a == self.c0:
            return q if na else self.const(0)
        if b == self.c0:
            return p if nb else self.const(0)
        if a == b:
            return p if na == nb else self.const(0)
        op = {(False, False): "AND", (True, False): "ANDNY", (False, True): "ANDYN",
              (True, True): "NOR"}[(na, nb)]
        return (self.b.gate(op, a, b), False)

    def mux(self, s, hi, lo):
        """s ? hi : lo for a select wire s."""
        if hi == lo:
            return hi
        if lo == self.const(0):
            return self.and_((s, False), hi)
        if hi == self.const(0):
            return self.and_((s, True), lo)
        if lo == self.const(1):
            return self.neg(self.and_((s, False), self.neg(hi)))
        if hi == self.const(1):
            return self.neg(self.and_((s, True), self.neg(lo)))
        if hi == self.neg(lo):
            return self.xor((s, False), lo)
        if hi[1] == lo[1]:
            return (self.b.gate("MUX", lo[0], hi[0], s), hi[1])
        return (self.b.gate("MUX", self.wire(lo, True), self.wire(hi, True), s), False)

    def wire(self, p, complement_gate=False):
        """A wire carrying the literal. A complemented one becomes a NOT gate,
        or, with complement_gate, the complement of the gate that made the
        wire (one more bootstrap, no extra layer) where there is one."""
        a, na = p
        if a == self.c0:
            return self.c1 if na else self.c0
        if not na:
            return a
        g = a - self.b.n_inputs
        if complement_gate and g >= 0:
            op, x, y, c = self.b.gates[g]
            comp = self._COMPLEMENT.get(_NAME[op])
            if comp is not None:
                return self.b.gate(comp, x, y, c)
        return self.b.gate("NOT", a)


def _run(L, net, ins, flips=None):
    """Literals of an XOR network's rows over the input literals; flips[r]
    complements row r (an affine constant)."""
    pairs, rowsigs = net
    sig = list(ins)
    for i, j in pairs:
        sig.append(L.xor(sig[i], sig[j]))
    out = []
    for r, s in enumerate(rowsigs):
        v = L.xor_all([sig[t] for t in s])
        out.append(L.neg(v) if flips and flips[r] else v)
    return out


def _table4(L, tt, xs, order, memo):
    """Literal of the 4-input function with truth table tt (bit x of tt is
    f(x), input k is bit k of x) by Shannon expansion over xs[order[3]]
    (outermost) .. xs[order[0]]; sub-functions are shared through memo."""
    def build(t, k):
        if k < 0:
            return L.const(t & 1)
        half = 1 << k
        lo, hi = t & ((1 << half) - 1), t >> half
        if lo == hi:
            return build(lo, k - 1)
        if (t, k) not in memo:
            memo[(t, k)] = L.mux(xs[order[k]], build(hi, k - 1), build(lo, k - 1))
        return memo[(t, k)]

    # reindex the table so that bit k of the index is input order[k]
    t = 0
    for x in range(16):
        src = sum(((x >> k) & 1) << order[k] for k in range(4))
        t |= ((tt >> src) & 1) << x
    return build(t, 3)


# ------------------------------------------------------------ the S-box plan
_PLAN = None


def _plan():
    """The XOR networks and tables shared by every S-box of the circuit."""
    global _PLAN
    if _PLAN is not None:
        return _PLAN
    n = len(_KARA)
    # z = tower(AffineInv(y)) ^ tower(AffineInv(0x63)) of the S-box input
    # y, and from z = (a << 4) | b the linear forms the products need: the
    # a- and b-sums of the Karatsuba products and lam a^2 + b^2
    m_in = _rows(lambda y: _to_tower(_AFFINE_INV[y]), 8, 8)
    c_in = _to_tower(_AFFINE_INV[0x63])
    flips = [(c_in >> k) & 1 for k in range(8)]
    to_z = _xor_network(m_in, 8)
    ell = _rows(lambda z: _mul4(_mul4(z >> 4, z >> 4), _LAM) ^ _mul4(z & 15, z & 15), 8, 4)
    forms = _xor_network([m << 4 for m in _KARA] + list(_KARA) + ell, 8)
    # InvMixColumns followed by the map to z of each byte, as one 32 x 32
    # network per column (cheaper than the two apart, one layer shallower)
    mix = [sum(1 << i for i in row) for row in _inv_mix_matrix()]
    mix_z = []
    for byte in range(4):
        for k in range(8):
            r = 0
            for i in range(8):
                if (m_in[k] >> i) & 1:
                    r ^= mix[8 * byte + i]
            mix_z.append(r)
    mix_z = _xor_network(mix_z, 32)
    # D's argument: ell_k ^ (a b)_k over [ell 0..3, P 0..8]
    delta = _xor_network([(1 << k) | sum(1 << (4 + t) for t in _KARA_SUB[k]) for k in range(4)],
                         4 + n, in_level=[0] * 4 + [1] * n)
    # output byte: FromTower((a D) << 4 | (a D ^ b D)) ^ K over [Q 0..8, S 0..8, K 0..7]
    back = _rows(lambda v: _FROM_TOWER[v], 8, 8)

    def out_row(j):
        m = 0
        for k in range(8):
            if (back[j] >> k) & 1:
                for t in _KARA_SUB[k % 4]:
                    m ^= 1 << t                 # a D
                    if k < 4:
                        m ^= 1 << (n + t)       # + b D in the low half
        return m | (1 << (2 * n + j))

    # products with a sum of D bits come one XOR later than the others
    late = [1 if bin(m).count("1") == 2 else 0 for m in _KARA]
    out = _xor_network([out_row(j) for j in range(8)], 2 * n + 8,
                       in_level=[1 + x for x in late] * 2 + [0] * 8)
    # Shannon order for the 4-bit inverse: fewest bootstraps, then depth
    best = None
    for order in itertools.permutations(range(4)):
        b = _Builder(5)               # wire 4 stands for the constants
        L = _Lits(b, 4, 4)
        memo = {}
        for m in (1, 2, 4, 8, 15):
            tt = sum(1 << x for x in range(16) if _parity(_INV4[x] & m))
            _table4(L, tt, [0, 1, 2, 3], order, memo)
        c = b.circuit([])
        key = (c.bootstraps, max(b.level))
        if best is None or key < best[0]:
            best = (key, order)
    _PLAN = {"to_z": to_z, "flips": flips, "forms": forms, "mix_z": mix_z, "delta": delta,
             "out": out, "order": best[1]}
    return _PLAN


def _inv_sbox_bytes(L, plan, forms):
    """Products a D and b D of one S-box from its linear forms."""
    n = len(_KARA)
    fa, fb, ell = forms[:n], forms[n:2 * n], forms[2 * n:]
    prods = [L.and_(fa[t], fb[t]) for t in range(n)]
    delta = _run(L, plan["delta"], ell + prods)
    # complemented arguments are absorbed into the tables
    neg = sum(int(delta[k][1]) << k for k in range(4))
    dw = [delta[k][0] for k in range(4)]
    memo = {}

    def table(mask):
        tt = sum(1 << x for x in range(16) if _parity(_INV4[x ^ neg] & mask))
        return _table4(L, tt, dw, plan["order"], memo)

    d = {m: table(m) for m in (1, 2, 4, 8, 15)}
    for m in _KARA:
        if m not in d:
            lo, hi = m & -m, m & (m - 1)
            d[m] = L.xor(d[lo], d[hi])
    return [L.and_(fa[t], d[m]) for t, m in enumerate(_KARA)] + \
           [L.and_(fb[t], d[m]) for t, m in enumerate(_KARA)]


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
            if op in ("MUX", "NMUX"):
                ops = (key[1], key[2], key[3])
            elif op in ("NOT", "COPY"):
                ops = (key[1],)
            else:
                ops = (key[1], key[2])
            self.level.append(1 + max(self.level[o] for o in ops))
            self.memo[key] = w
        return w

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


def _inv_shift_rows(state_bytes):
    """Byte (row r, column c) <- (row r, column c - r); byte 4 c + r."""
    out = [None] * 16
    for c in range(4):
        for r in range(4):
            out[4 * c + r] = state_bytes[4 * ((c - r) % 4) + r]
    return out


def inv_sbox_circuit():
    """InvSubBytes of one byte on its own: inputs [0, 8) the byte (least
    significant bit first), 8 and 9 the constants 0 and 1; outputs the 8
    bits of the inverse S-box. The same construction as in decrypt_circuit,
    small enough to evaluate over all 256 inputs."""
    plan = _plan()
    b = _Builder(10)
    L = _Lits(b, 8, 9)
    z = _run(L, plan["to_z"], [(i, False) for i in range(8)], plan["flips"])
    out = _run(L, plan["out"], _inv_sbox_bytes(L, plan, _run(L, plan["forms"], z)) +
               [L.const(0)] * 8)
    return b.circuit([L.wire(v) for v in out])


_CIRCUIT = None


def decrypt_circuit():
    """The AES-128 decryption circuit (built once per process)."""
    global _CIRCUIT
    if _CIRCUIT is not None:
        return _CIRCUIT
    plan = _plan()
    b = _Builder(N_INPUTS)
    L = _Lits(b, CONST0, CONST1)
    key = lambda r, i: (128 + 128 * r + i, False)  # noqa: E731
    state = [L.xor((i, False), key(10, i)) for i in range(128)]
    byte_forms = lambda z: _run(L, plan["forms"], z)  # noqa: E731
    shifted = _inv_shift_rows([state[8 * i:8 * i + 8] for i in range(16)])
    forms = [byte_forms(_run(L, plan["to_z"], x, plan["flips"])) for x in shifted]
    for rnd in range(9, -1, -1):
        out = [_run(L, plan["out"], _inv_sbox_bytes(L, plan, forms[i]) +
                    [key(rnd, 8 * i + j) for j in range(8)]) for i in range(16)]
        if rnd == 0:
            state = sum(out, [])
            break
        zs = []
        for c in range(4):
            col = _run(L, plan["mix_z"], sum(out[4 * c:4 * c + 4], []), plan["flips"] * 4)
            zs += [col[8 * r:8 * r + 8] for r in range(4)]
        forms = [byte_forms(z) for z in _inv_shift_rows(zs)]
    _CIRCUIT = b.circuit([L.wire(v) for v in state])
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
    out[:, -1] = np.where(v == 1, np.uint32(1 << 29), np.uint32((1 << 32) - (1 << 29)))
    return out
