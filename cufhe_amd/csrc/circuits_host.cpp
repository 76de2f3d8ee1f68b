// Host-side builders of the AES-128 decryption and ASCON-XOF circuits of
// lvl1 gates (include/cufhe_circuits.h), for the C++ drop-ins of the
// reference's AESDec (src/aes_gpu.cu:904-1118) and ASCONXOF
// (src/ascon_gpu.cu:1206-1282), which evaluate the same functions with
// circuit bootstrapping instead.
//
// AES-128 decryption (FIPS-197 5.3): InvSubBytes is the inverse affine map
// and the GF(2^8) inverse, taken in the tower field GF((2^4)^2) =
// GF(2^4)[Y]/(Y^2 + Y + lam) over GF(2)[z]/(z^4 + z + 1) (Satoh et al.,
// ASIACRYPT 2001; Canright, CHES 2005): aY + b has the inverse
// (aD)Y + (a + b)D, D = (lam a^2 + ab + b^2)^-1, with two-level Karatsuba
// products, the 4-bit inverse by Shannon expansion, and XOR networks from a
// greedy common-subexpression search (Paar, ISIT 1997) held to each
// output's minimum depth. InvMixColumns and the next S-box's input map are
// one network per column. Complements ride on literals and fold into the
// consuming gate.
//
// ASCON-XOF (Ascon v1.2, rate 64): the state after p^12 of the IV is public
// and enters as constants; absorb XORs message bytes into x0 (most
// significant byte first), pads 0x80, permutes; squeeze outputs x0 with p^12
// between 8-byte blocks. p^12: round constants as complements, the bitsliced
// S-box, the linear layer x ^= (x >>> a) ^ (x >>> b).
//
// Wire layouts and bit order (bit b of byte i is wire 8 i + b) are those of
// include/cufhe_circuits.h.
#include "../../include/cufhe_circuits.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <stdexcept>
#include <utility>
#include <vector>

namespace {

// ------------------------------------------------------------ gate builder
struct Built {
    int n_inputs = 0;
    std::vector<cufhe_gate_desc> gates;
    std::vector<int> layer_end;
    std::vector<uint32_t> outputs;
};

bool symmetric(int op)
{
    return op == CUFHE_AND || op == CUFHE_OR || op == CUFHE_XOR || op == CUFHE_XNOR ||
           op == CUFHE_NAND || op == CUFHE_NOR;
}

// Gates with hash-consing; a gate's level is one more than the latest level
// among its operands (inputs are level 0).
struct Builder {
    int n_inputs;
    std::vector<std::array<uint32_t, 4>> gates;  // op, in0, in1, inc
    std::vector<int> level;
    std::map<std::array<uint32_t, 4>, uint32_t> memo;

    explicit Builder(int n) : n_inputs(n), level(n, 0) {}

    uint32_t gate(int op, uint32_t a, uint32_t b = 0, uint32_t c = 0)
    {
        if (symmetric(op) && a > b) std::swap(a, b);
        const std::array<uint32_t, 4> key{static_cast<uint32_t>(op), a, b, c};
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
        const uint32_t w = static_cast<uint32_t>(n_inputs + gates.size());
        gates.push_back(key);
        int lv = level[a];
        if (op == CUFHE_MUX || op == CUFHE_NMUX)
            lv = std::max({level[a], level[b], level[c]});
        else if (op != CUFHE_NOT && op != CUFHE_COPY)
            lv = std::max(level[a], level[b]);
        level.push_back(lv + 1);
        memo.emplace(key, w);
        return w;
    }

    // gates in level order; gate g then writes wire n_inputs + g
    Built finish(const std::vector<uint32_t>& outs) const
    {
        const int ng = static_cast<int>(gates.size());
        std::vector<int> order(ng);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return level[n_inputs + x] < level[n_inputs + y];
        });
        std::vector<uint32_t> remap(n_inputs + ng);
        std::iota(remap.begin(), remap.begin() + n_inputs, 0u);
        for (int i = 0; i < ng; i++) remap[n_inputs + order[i]] = static_cast<uint32_t>(n_inputs + i);
        Built b;
        b.n_inputs = n_inputs;
        int last = -1;
        for (int i = 0; i < ng; i++) {
            const auto& g = gates[order[i]];
            const int lv = level[n_inputs + order[i]];
            if (last >= 0 && lv != last) b.layer_end.push_back(i);
            last = lv;
            b.gates.push_back(cufhe_gate_desc{static_cast<int32_t>(g[0]),
                                              static_cast<uint32_t>(n_inputs + i), remap[g[1]],
                                              remap[g[2]], remap[g[3]]});
        }
        b.layer_end.push_back(ng);
        for (uint32_t w : outs) b.outputs.push_back(remap[w]);
        return b;
    }
};

// Literals (wire, complemented); (c0, false) is 0 and (c0, true) is 1.
struct Lit {
    uint32_t w;
    bool neg;
    bool operator==(const Lit& o) const { return w == o.w && neg == o.neg; }
};

struct Lits {
    Builder& b;
    uint32_t c0, c1;

    Lit konst(bool v) const { return Lit{c0, v}; }
    static Lit neg(Lit p) { return Lit{p.w, !p.neg}; }
    int lvl(Lit p) const { return b.level[p.w]; }

    Lit x(Lit p, Lit q)
    {
        if (p.w == c0) return Lit{q.w, q.neg != p.neg};
        if (q.w == c0) return Lit{p.w, p.neg != q.neg};
        if (p.w == q.w) return Lit{c0, p.neg != q.neg};
        return Lit{b.gate(CUFHE_XOR, p.w, q.w), p.neg != q.neg};
    }
    // XOR of the literals, the two earliest first
    Lit x_all(std::vector<Lit> v)
    {
        if (v.empty()) return konst(false);
        auto by_level = [&](const Lit& p, const Lit& q) { return lvl(p) < lvl(q); };
        std::stable_sort(v.begin(), v.end(), by_level);
        while (v.size() > 1) {
            const Lit r = x(v[0], v[1]);
            v.erase(v.begin(), v.begin() + 2);
            v.insert(std::upper_bound(v.begin(), v.end(), r, by_level), r);
        }
        return v[0];
    }
    Lit a(Lit p, Lit q)
    {
        if (p.w == c0) return p.neg ? q : konst(false);
        if (q.w == c0) return q.neg ? p : konst(false);
        if (p.w == q.w) return p.neg == q.neg ? p : konst(false);
        const int op = !p.neg && !q.neg  ? CUFHE_AND
                       : p.neg && !q.neg ? CUFHE_ANDNY
                       : !p.neg && q.neg ? CUFHE_ANDYN
                                         : CUFHE_NOR;
        return Lit{b.gate(op, p.w, q.w), false};
    }
    // a wire carrying the literal: a complemented gate output becomes the
    // complement gate (same level) when complement_gate, a NOT otherwise
    uint32_t wire(Lit p, bool complement_gate = false)
    {
        if (p.w == c0) return p.neg ? c1 : c0;
        if (!p.neg) return p.w;
        const int g = static_cast<int>(p.w) - b.n_inputs;
        if (complement_gate && g >= 0) {
            const auto& d = b.gates[g];
            int comp = -1;
            switch (static_cast<int>(d[0])) {
                case CUFHE_AND: comp = CUFHE_NAND; break;
                case CUFHE_NAND: comp = CUFHE_AND; break;
                case CUFHE_OR: comp = CUFHE_NOR; break;
                case CUFHE_NOR: comp = CUFHE_OR; break;
                case CUFHE_XOR: comp = CUFHE_XNOR; break;
                case CUFHE_XNOR: comp = CUFHE_XOR; break;
                case CUFHE_ANDNY: comp = CUFHE_ORYN; break;
                case CUFHE_ORYN: comp = CUFHE_ANDNY; break;
                case CUFHE_ANDYN: comp = CUFHE_ORNY; break;
                case CUFHE_ORNY: comp = CUFHE_ANDYN; break;
                case CUFHE_MUX: comp = CUFHE_NMUX; break;
                case CUFHE_NMUX: comp = CUFHE_MUX; break;
                default: break;
            }
            if (comp >= 0) return b.gate(comp, d[1], d[2], d[3]);
        }
        return b.gate(CUFHE_NOT, p.w);
    }
    // s ? hi : lo for a select wire s
    Lit mux(uint32_t s, Lit hi, Lit lo)
    {
        if (hi == lo) return hi;
        if (lo == konst(false)) return a(Lit{s, false}, hi);
        if (hi == konst(false)) return a(Lit{s, true}, lo);
        if (lo == konst(true)) return neg(a(Lit{s, false}, neg(hi)));
        if (hi == konst(true)) return neg(a(Lit{s, true}, neg(lo)));
        if (hi == neg(lo)) return x(Lit{s, false}, lo);
        if (hi.neg == lo.neg) return Lit{b.gate(CUFHE_MUX, lo.w, hi.w, s), hi.neg};
        const uint32_t lw = wire(lo, true);  // lo, then hi: a fixed gate order
        const uint32_t hw = wire(hi, true);
        return Lit{b.gate(CUFHE_MUX, lw, hw, s), false};
    }
};

// ------------------------------------------------------------ XOR networks
int xor_depth(std::vector<int> lv)
{
    if (lv.empty()) return 0;
    std::sort(lv.begin(), lv.end());
    while (lv.size() > 1) {
        const int v = std::max(lv[0], lv[1]) + 1;
        lv.erase(lv.begin(), lv.begin() + 2);
        lv.insert(std::upper_bound(lv.begin(), lv.end(), v), v);
    }
    return lv[0];
}

struct XorNet {
    std::vector<std::pair<int, int>> pairs;  // pairs[t] makes signal n_in + t
    std::vector<std::vector<int>> rows;      // signals XORed into each row
};

// splitmix64: the tie-breaking random stream of the search
struct Mix64 {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return static_cast<double>(z >> 11) * 0x1.0p-53;
    }
};

// Greedy common-subexpression search over rows (bit masks over n_in
// signals): XOR the pair of signals that occurs in the most rows, among
// pairs that keep every row at its minimum depth; ties at random, best of
// `tries` seeds (fewest XORs, then least depth).
XorNet xor_network(const std::vector<uint64_t>& masks, int n_in, std::vector<int> in_level = {},
                   int tries = 8)
{
    if (in_level.empty()) in_level.assign(n_in, 0);
    XorNet best;
    std::pair<int, int> best_cost{1 << 30, 1 << 30};
    for (int seed = 0; seed < tries; seed++) {
        Mix64 rng{static_cast<uint64_t>(seed) + 1};
        std::vector<int> lv = in_level;
        std::vector<std::vector<int>> sets;
        std::vector<int> bound;
        for (uint64_t m : masks) {
            std::vector<int> s;
            for (int i = 0; i < n_in; i++)
                if ((m >> i) & 1) s.push_back(i);
            std::vector<int> l;
            for (int i : s) l.push_back(lv[i]);
            bound.push_back(xor_depth(l));
            sets.push_back(s);
        }
        XorNet net;
        while (true) {
            std::map<std::pair<int, int>, int> count;
            for (const auto& s : sets)
                for (size_t x = 0; x < s.size(); x++)
                    for (size_t y = x + 1; y < s.size(); y++) count[{s[x], s[y]}]++;
            struct Cand {
                std::pair<int, int> p;
                int c, l;
                double r;
            };
            std::vector<Cand> cands;
            for (const auto& [p, c] : count)
                if (c >= 2) cands.push_back({p, c, std::max(lv[p.first], lv[p.second]), rng.next()});
            std::sort(cands.begin(), cands.end(), [](const Cand& u, const Cand& v) {
                if (u.c != v.c) return u.c > v.c;
                if (u.l != v.l) return u.l < v.l;
                return u.r < v.r;
            });
            int chosen = -1;
            for (size_t k = 0; k < cands.size() && chosen < 0; k++) {
                const auto [i, j] = cands[k].p;
                const int nl = std::max(lv[i], lv[j]) + 1;
                bool ok = true;
                for (size_t r = 0; r < sets.size() && ok; r++) {
                    const auto& s = sets[r];
                    if (!std::binary_search(s.begin(), s.end(), i) ||
                        !std::binary_search(s.begin(), s.end(), j))
                        continue;
                    std::vector<int> l{nl};
                    for (int t : s)
                        if (t != i && t != j) l.push_back(lv[t]);
                    ok = xor_depth(l) <= bound[r];
                }
                if (ok) chosen = static_cast<int>(k);
            }
            if (chosen < 0) break;
            const auto [i, j] = cands[chosen].p;
            const int sig = n_in + static_cast<int>(net.pairs.size());
            net.pairs.push_back({i, j});
            lv.push_back(std::max(lv[i], lv[j]) + 1);
            for (auto& s : sets) {
                if (std::binary_search(s.begin(), s.end(), i) &&
                    std::binary_search(s.begin(), s.end(), j)) {
                    s.erase(std::find(s.begin(), s.end(), i));
                    s.erase(std::find(s.begin(), s.end(), j));
                    s.insert(std::upper_bound(s.begin(), s.end(), sig), sig);
                }
            }
        }
        int xors = static_cast<int>(net.pairs.size()), depth = 0;
        for (const auto& s : sets) {
            xors += std::max<int>(0, static_cast<int>(s.size()) - 1);
            std::vector<int> l;
            for (int t : s) l.push_back(lv[t]);
            depth = std::max(depth, xor_depth(l));
        }
        net.rows = sets;
        if (std::make_pair(xors, depth) < best_cost) {
            best_cost = {xors, depth};
            best = net;
        }
    }
    return best;
}

std::vector<Lit> run(Lits& L, const XorNet& net, const std::vector<Lit>& ins,
                     const std::vector<int>& flips = {})
{
    std::vector<Lit> sig = ins;
    for (const auto& [i, j] : net.pairs) sig.push_back(L.x(sig[i], sig[j]));
    std::vector<Lit> out;
    for (size_t r = 0; r < net.rows.size(); r++) {
        std::vector<Lit> v;
        for (int t : net.rows[r]) v.push_back(sig[t]);
        Lit o = L.x_all(v);
        if (!flips.empty() && flips[r]) o = Lits::neg(o);
        out.push_back(o);
    }
    return out;
}

// ------------------------------------------------------------ AES fields
int gf_mul8(int a, int b)  // GF(2^8) modulo x^8 + x^4 + x^3 + x + 1
{
    int r = 0;
    for (int i = 0; i < 8; i++) {
        if (b & 1) r ^= a;
        a = ((a << 1) ^ ((a & 0x80) ? 0x1B : 0)) & 0xFF;
        b >>= 1;
    }
    return r;
}
int affine_linear(int g)  // FIPS-197 5.1.1 without the constant 0x63
{
    int s = 0;
    for (int i = 0; i < 8; i++) {
        const int bit = (g >> i) ^ (g >> ((i + 4) % 8)) ^ (g >> ((i + 5) % 8)) ^
                        (g >> ((i + 6) % 8)) ^ (g >> ((i + 7) % 8));
        s |= (bit & 1) << i;
    }
    return s;
}
int mul4(int a, int b)  // GF(2^4) = GF(2)[z]/(z^4 + z + 1)
{
    int r = 0;
    for (int i = 0; i < 4; i++)
        if ((b >> i) & 1) r ^= a << i;
    for (int i = 6; i >= 4; i--)
        if ((r >> i) & 1) r ^= 0b10011 << (i - 4);
    return r;
}
int parity(uint64_t x) { return __builtin_popcountll(x) & 1; }

struct Tower {
    int lam = 0;
    std::array<int, 16> inv4{};
    std::array<int, 8> basis{};        // images of x^0 .. x^7
    std::array<int, 256> from_tower{}, affine_inv{};
    Tower()
    {
        for (int x = 1; x < 16; x++)
            for (int y = 1; y < 16; y++)
                if (mul4(x, y) == 1) inv4[x] = y;
        for (int v = 1; v < 16 && !lam; v++) {
            bool root = false;
            for (int t = 0; t < 16; t++) root |= (mul4(t, t) ^ t ^ v) == 0;
            if (!root) lam = v;
        }
        for (int beta = 2; beta < 256; beta++) {
            std::array<int, 9> pw{1};
            for (int i = 1; i <= 8; i++) pw[i] = mul8(pw[i - 1], beta);
            if ((pw[8] ^ pw[4] ^ pw[3] ^ pw[1] ^ pw[0]) == 0) {
                std::copy(pw.begin(), pw.begin() + 8, basis.begin());
                break;
            }
        }
        for (int x = 0; x < 256; x++) {
            from_tower[to(x)] = x;
            affine_inv[affine_linear(x)] = x;
        }
    }
    int mul8(int p, int q) const  // GF((2^4)^2), p = (a << 4) | b is aY + b
    {
        const int a = p >> 4, b = p & 15, c = q >> 4, d = q & 15, ac = mul4(a, c);
        return ((ac ^ mul4(a, d) ^ mul4(b, c)) << 4) | (mul4(ac, lam) ^ mul4(b, d));
    }
    int to(int x) const
    {
        int r = 0;
        for (int i = 0; i < 8; i++)
            if ((x >> i) & 1) r ^= basis[i];
        return r;
    }
};

template <class F>
std::vector<uint64_t> lin_rows(F f, int n_in, int n_out)
{
    std::vector<uint64_t> rows(n_out, 0);
    for (int i = 0; i < n_in; i++) {
        const int v = f(1 << i);
        for (int k = 0; k < n_out; k++)
            if ((v >> k) & 1) rows[k] |= 1ULL << i;
    }
    return rows;
}

// two-level Karatsuba in GF(2^4): products (sum_{i in m} a_i)(sum_{i in m} b_i)
constexpr std::array<int, 9> kKara{0b0001, 0b0010, 0b0011, 0b0100, 0b1000,
                                   0b1100, 0b0101, 0b1010, 0b1111};

// Table from 4-bit inverse: Shannon expansion over xs[order[3]] .. xs[order[0]]
Lit table4(Lits& L, int tt, const std::array<uint32_t, 4>& xs, const std::array<int, 4>& order,
           std::map<std::pair<int, int>, Lit>& memo)
{
    int t = 0;
    for (int x = 0; x < 16; x++) {
        int src = 0;
        for (int k = 0; k < 4; k++) src |= ((x >> k) & 1) << order[k];
        t |= ((tt >> src) & 1) << x;
    }
    auto build = [&](auto&& self, int tab, int k) -> Lit {
        if (k < 0) return L.konst(tab & 1);
        const int half = 1 << k;
        const int lo = tab & ((1 << half) - 1), hi = tab >> half;
        if (lo == hi) return self(self, lo, k - 1);
        auto it = memo.find({tab, k});
        if (it != memo.end()) return it->second;
        const Lit h = self(self, hi, k - 1);  // hi, then lo: a fixed gate order
        const Lit l = self(self, lo, k - 1);
        const Lit v = L.mux(xs[order[k]], h, l);
        memo.emplace(std::make_pair(tab, k), v);
        return v;
    };
    return build(build, t, 3);
}

struct AesPlan {
    Tower tf;
    std::array<std::vector<int>, 4> sub;  // Karatsuba products of bit k of a product
    XorNet to_z, forms, mix_z, delta, out;
    std::vector<int> flips;  // tower(AffineInv(0x63)) bits
    std::array<int, 4> order{};

    AesPlan()
    {
        const int n = static_cast<int>(kKara.size());
        for (int k = 0; k < 4; k++) {
            for (int s = 0; s < (1 << n); s++) {
                bool ok = true;
                for (int i = 0; i < 4 && ok; i++)
                    for (int j = 0; j < 4 && ok; j++) {
                        int acc = 0;
                        for (int t = 0; t < n; t++)
                            if ((s >> t) & 1) acc ^= (kKara[t] >> i) & (kKara[t] >> j) & 1;
                        ok = acc == ((mul4(1 << i, 1 << j) >> k) & 1);
                    }
                if (ok) {
                    for (int t = 0; t < n; t++)
                        if ((s >> t) & 1) sub[k].push_back(t);
                    break;
                }
            }
        }
        const auto m_in = lin_rows([&](int y) { return tf.to(tf.affine_inv[y]); }, 8, 8);
        const int c_in = tf.to(tf.affine_inv[0x63]);
        for (int k = 0; k < 8; k++) flips.push_back((c_in >> k) & 1);
        to_z = xor_network(m_in, 8);
        const auto ell = lin_rows(
            [&](int z) { return mul4(mul4(z >> 4, z >> 4), tf.lam) ^ mul4(z & 15, z & 15); }, 8, 4);
        std::vector<uint64_t> fz;
        for (int m : kKara) fz.push_back(static_cast<uint64_t>(m) << 4);
        for (int m : kKara) fz.push_back(static_cast<uint64_t>(m));
        for (uint64_t m : ell) fz.push_back(m);
        forms = xor_network(fz, 8);
        // InvMixColumns (coefficients 0e 0b 0d 09) then the map to z of each byte
        std::vector<uint64_t> mix(32, 0);
        const int coef[4] = {0x0E, 0x0B, 0x0D, 0x09};
        for (int ib = 0; ib < 32; ib++) {
            const int byte = ib / 8, bit = ib % 8;
            for (int r = 0; r < 4; r++) {
                const int v = gf_mul8(coef[((byte - r) % 4 + 4) % 4], 1 << bit);
                for (int ob = 0; ob < 8; ob++)
                    if ((v >> ob) & 1) mix[8 * r + ob] |= 1ULL << ib;
            }
        }
        std::vector<uint64_t> mz;
        for (int byte = 0; byte < 4; byte++)
            for (int k = 0; k < 8; k++) {
                uint64_t r = 0;
                for (int i = 0; i < 8; i++)
                    if ((m_in[k] >> i) & 1) r ^= mix[8 * byte + i];
                mz.push_back(r);
            }
        mix_z = xor_network(mz, 32);
        std::vector<uint64_t> drows;
        for (int k = 0; k < 4; k++) {
            uint64_t m = 1ULL << k;
            for (int t : sub[k]) m |= 1ULL << (4 + t);
            drows.push_back(m);
        }
        std::vector<int> dl(4, 0);
        dl.resize(4 + n, 1);
        delta = xor_network(drows, 4 + n, dl);
        const auto back = lin_rows([&](int v) { return tf.from_tower[v]; }, 8, 8);
        std::vector<uint64_t> orows;
        for (int j = 0; j < 8; j++) {
            uint64_t m = 0;
            for (int k = 0; k < 8; k++) {
                if (!((back[j] >> k) & 1)) continue;
                for (int t : sub[k % 4]) {
                    m ^= 1ULL << t;                  // a D
                    if (k < 4) m ^= 1ULL << (n + t);  // + b D in the low half
                }
            }
            orows.push_back(m | (1ULL << (2 * n + j)));
        }
        std::vector<int> ol;
        for (int rep = 0; rep < 2; rep++)
            for (int m : kKara) ol.push_back(1 + (__builtin_popcount(m) == 2 ? 1 : 0));
        ol.resize(2 * n + 8, 0);
        out = xor_network(orows, 2 * n + 8, ol);
        // Shannon order of the 4-bit inverse: fewest bootstraps, then depth
        std::pair<int, int> best{1 << 30, 1 << 30};
        std::array<int, 4> o{0, 1, 2, 3};
        do {
            Builder b(5);
            Lits L{b, 4, 4};
            std::map<std::pair<int, int>, Lit> memo;
            for (int m : {1, 2, 4, 8, 15}) {
                int tt = 0;
                for (int x = 0; x < 16; x++)
                    if (parity(tf.inv4[x] & m)) tt |= 1 << x;
                table4(L, tt, {0, 1, 2, 3}, o, memo);
            }
            int boots = 0;
            for (const auto& g : b.gates)
                boots += g[0] == CUFHE_MUX || g[0] == CUFHE_NMUX ? 2
                         : g[0] == CUFHE_NOT || g[0] == CUFHE_COPY ? 0
                                                                    : 1;
            const std::pair<int, int> key{boots, *std::max_element(b.level.begin(), b.level.end())};
            if (key < best) {
                best = key;
                order = o;
            }
        } while (std::next_permutation(o.begin(), o.end()));
    }

    // products a D and b D of one S-box from its linear forms
    std::vector<Lit> sbox_products(Lits& L, const std::vector<Lit>& f) const
    {
        const int n = static_cast<int>(kKara.size());
        std::vector<Lit> v(f.begin() + 2 * n, f.end());
        for (int t = 0; t < n; t++) v.push_back(L.a(f[t], f[n + t]));
        const std::vector<Lit> d = run(L, delta, v);
        int neg = 0;
        std::array<uint32_t, 4> dw{};
        for (int k = 0; k < 4; k++) {
            neg |= (d[k].neg ? 1 : 0) << k;
            dw[k] = d[k].w;
        }
        std::map<std::pair<int, int>, Lit> memo;
        std::map<int, Lit> dm;
        for (int m : {1, 2, 4, 8, 15}) {
            int tt = 0;
            for (int x = 0; x < 16; x++)
                if (parity(tf.inv4[x ^ neg] & m)) tt |= 1 << x;
            dm.emplace(m, table4(L, tt, dw, order, memo));
        }
        for (int m : kKara)
            if (!dm.count(m)) dm.emplace(m, L.x(dm.at(m & -m), dm.at(m & (m - 1))));
        std::vector<Lit> prods;
        for (int t = 0; t < n; t++) prods.push_back(L.a(f[t], dm.at(kKara[t])));
        for (int t = 0; t < n; t++) prods.push_back(L.a(f[n + t], dm.at(kKara[t])));
        return prods;
    }
};

const AesPlan& aes_plan()
{
    static const AesPlan plan;
    return plan;
}

// byte (row r, column c) <- (row r, column c - r); byte index 4 c + r
template <class T>
std::vector<T> inv_shift_rows(const std::vector<T>& s)
{
    std::vector<T> out(16);
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) out[4 * c + r] = s[4 * ((c - r + 4) % 4) + r];
    return out;
}

Built build_aes_decrypt()
{
    const AesPlan& P = aes_plan();
    constexpr int kKeyBase = 128, kC0 = 128 + 11 * 128;
    Builder b(CUFHE_AES128_INPUTS);
    Lits L{b, kC0, kC0 + 1};
    auto key = [](int r, int i) { return Lit{static_cast<uint32_t>(kKeyBase + 128 * r + i), false}; };
    std::vector<std::vector<Lit>> bytes(16);
    for (int i = 0; i < 128; i++)
        bytes[i / 8].push_back(L.x(Lit{static_cast<uint32_t>(i), false}, key(10, i)));
    std::vector<std::vector<Lit>> forms;
    for (const auto& x : inv_shift_rows(bytes)) forms.push_back(run(L, P.forms, run(L, P.to_z, x, P.flips)));
    std::vector<Lit> state;
    for (int rnd = 9; rnd >= 0; rnd--) {
        std::vector<std::vector<Lit>> out(16);
        for (int i = 0; i < 16; i++) {
            std::vector<Lit> in = P.sbox_products(L, forms[i]);
            for (int j = 0; j < 8; j++) in.push_back(key(rnd, 8 * i + j));
            out[i] = run(L, P.out, in);
        }
        if (rnd == 0) {
            for (const auto& o : out) state.insert(state.end(), o.begin(), o.end());
            break;
        }
        std::vector<int> fl;
        for (int r = 0; r < 4; r++) fl.insert(fl.end(), P.flips.begin(), P.flips.end());
        std::vector<std::vector<Lit>> zs;
        for (int c = 0; c < 4; c++) {
            std::vector<Lit> col;
            for (int r = 0; r < 4; r++) col.insert(col.end(), out[4 * c + r].begin(), out[4 * c + r].end());
            const std::vector<Lit> z = run(L, P.mix_z, col, fl);
            for (int r = 0; r < 4; r++) zs.emplace_back(z.begin() + 8 * r, z.begin() + 8 * r + 8);
        }
        forms.clear();
        for (const auto& z : inv_shift_rows(zs)) forms.push_back(run(L, P.forms, z));
    }
    std::vector<uint32_t> outs;
    for (const Lit& v : state) outs.push_back(L.wire(v));
    return b.finish(outs);
}

// ------------------------------------------------------------ ASCON
constexpr uint64_t kXofIV = 0x00400C0000000000ULL;
constexpr int kRC[12] = {0xF0, 0xE1, 0xD2, 0xC3, 0xB4, 0xA5, 0x96, 0x87, 0x78, 0x69, 0x5A, 0x4B};
constexpr int kRot[5][2] = {{19, 28}, {61, 39}, {1, 6}, {10, 17}, {7, 41}};

uint64_t rotr(uint64_t x, int n) { return (x >> n) | (x << (64 - n)); }

std::array<uint64_t, 5> p12_plain(std::array<uint64_t, 5> x)
{
    for (int r = 0; r < 12; r++) {
        x[2] ^= kRC[r];
        x[0] ^= x[4];
        x[4] ^= x[3];
        x[2] ^= x[1];
        uint64_t t[5];
        for (int i = 0; i < 5; i++) t[i] = ~x[i] & x[(i + 1) % 5];
        for (int i = 0; i < 5; i++) x[i] ^= t[(i + 1) % 5];
        x[1] ^= x[0];
        x[0] ^= x[4];
        x[3] ^= x[2];
        x[2] = ~x[2];
        for (int i = 0; i < 5; i++) x[i] ^= rotr(x[i], kRot[i][0]) ^ rotr(x[i], kRot[i][1]);
    }
    return x;
}

using Word = std::array<Lit, 64>;  // bit k = 2^k

void permutation(Lits& L, std::array<Word, 5>& x)
{
    for (int r = 0; r < 12; r++) {
        for (int k = 0; k < 8; k++)  // the round constants have 8 bits
            if ((kRC[r] >> k) & 1) x[2][k] = Lits::neg(x[2][k]);
        for (int k = 0; k < 64; k++) {
            Lit s[5];
            for (int i = 0; i < 5; i++) s[i] = x[i][k];
            s[0] = L.x(s[0], s[4]);
            s[4] = L.x(s[4], s[3]);
            s[2] = L.x(s[2], s[1]);
            Lit t[5];
            for (int i = 0; i < 5; i++) t[i] = L.a(Lits::neg(s[i]), s[(i + 1) % 5]);
            for (int i = 0; i < 5; i++) s[i] = L.x(s[i], t[(i + 1) % 5]);
            s[1] = L.x(s[1], s[0]);
            s[0] = L.x(s[0], s[4]);
            s[3] = L.x(s[3], s[2]);
            s[2] = Lits::neg(s[2]);
            for (int i = 0; i < 5; i++) x[i][k] = s[i];
        }
        for (int i = 0; i < 5; i++) {
            const Word w = x[i];
            for (int k = 0; k < 64; k++)
                x[i][k] = L.x(L.x(w[k], w[(k + kRot[i][0]) % 64]), w[(k + kRot[i][1]) % 64]);
        }
    }
}

int rate_pos(int j, int bit) { return 8 * (7 - j) + bit; }  // rate byte j, MSB first

// Absorb of the reference's XOF (src/ascon_gpu.cu:1051-1077): whole 8-byte
// blocks XORed into x0 and permuted, the rest XORed, the padding bit of
// byte `left` flipped, p^12. Message bit i is wire msg0 + i.
void absorb(Lits& L, std::array<Word, 5>& x, uint32_t msg0, int in_bytes)
{
    auto msg = [&](int i) { return Lit{msg0 + static_cast<uint32_t>(i), false}; };
    int left = in_bytes, off = 0;
    while (left >= 8) {
        for (int j = 0; j < 8; j++)
            for (int bit = 0; bit < 8; bit++)
                x[0][rate_pos(j, bit)] = L.x(x[0][rate_pos(j, bit)], msg(8 * (off + j) + bit));
        permutation(L, x);
        off += 8;
        left -= 8;
    }
    for (int j = 0; j < left; j++)
        for (int bit = 0; bit < 8; bit++)
            x[0][rate_pos(j, bit)] = L.x(x[0][rate_pos(j, bit)], msg(8 * (off + j) + bit));
    x[0][rate_pos(left, 7)] = Lits::neg(x[0][rate_pos(left, 7)]);  // padding byte 0x80
    permutation(L, x);
}

// Squeeze (src/ascon_gpu.cu:1080-1104): x0's bytes, p^12 between blocks and
// none after the last.
std::vector<uint32_t> squeeze(Lits& L, std::array<Word, 5>& x, int out_bytes)
{
    std::vector<uint32_t> outs;
    for (int left = out_bytes;;) {
        const int n = std::min(left, 8);
        for (int j = 0; j < n; j++)
            for (int bit = 0; bit < 8; bit++) outs.push_back(L.wire(x[0][rate_pos(j, bit)]));
        left -= n;
        if (left == 0) break;
        permutation(L, x);
    }
    return outs;
}

std::array<Word, 5> state_inputs()
{
    std::array<Word, 5> x;
    for (int i = 0; i < 5; i++)
        for (int k = 0; k < 64; k++) x[i][k] = Lit{static_cast<uint32_t>(64 * i + k), false};
    return x;
}

void append_state(Lits& L, const std::array<Word, 5>& x, std::vector<uint32_t>& outs)
{
    for (int i = 0; i < 5; i++)
        for (int k = 0; k < 64; k++) outs.push_back(L.wire(x[i][k]));
}

Built build_ascon_xof(int in_bytes, int out_bytes)
{
    const uint32_t c0 = 8 * in_bytes, c1 = c0 + 1;
    Builder b(8 * in_bytes + 2);
    Lits L{b, c0, c1};
    const auto s0 = p12_plain({kXofIV, 0, 0, 0, 0});
    std::array<Word, 5> x;
    for (int i = 0; i < 5; i++)
        for (int k = 0; k < 64; k++) x[i][k] = L.konst((s0[i] >> k) & 1);
    absorb(L, x, 0, in_bytes);
    std::vector<uint32_t> outs = squeeze(L, x, out_bytes);
    if (b.gates.empty()) {
        // an empty message: the output is public; one COPY gate per distinct
        // constant wire gives the server a circuit with one layer
        std::map<uint32_t, uint32_t> copy;
        for (uint32_t w : outs) copy.emplace(w, 0);
        for (auto& [w, c] : copy) c = b.gate(CUFHE_COPY, w);  // in wire order
        for (uint32_t& w : outs) w = copy.at(w);
    }
    return b.finish(outs);
}

// state (320) and message in, state out
Built build_ascon_absorb(int in_bytes)
{
    const uint32_t c0 = 320 + 8 * in_bytes;
    Builder b(static_cast<int>(c0) + 2);
    Lits L{b, c0, c0 + 1};
    std::array<Word, 5> x = state_inputs();
    absorb(L, x, 320, in_bytes);
    std::vector<uint32_t> outs;
    append_state(L, x, outs);
    return b.finish(outs);
}

// state in, output bytes and the state after the squeeze out
Built build_ascon_squeeze(int out_bytes)
{
    const uint32_t c0 = 320;
    Builder b(322);
    Lits L{b, c0, c0 + 1};
    std::array<Word, 5> x = state_inputs();
    std::vector<uint32_t> outs = squeeze(L, x, out_bytes);
    append_state(L, x, outs);
    return b.finish(outs);
}

int deliver(const Built& c, cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates, int* layer_end,
            uint32_t* outputs)
{
    if (!sizes) return CUFHE_ERR_ARG;
    const bool fill = gates || layer_end || outputs;
    if (fill) {
        // a buffer may be NULL only where its array is empty (no gates)
        if ((!gates && !c.gates.empty()) || (!layer_end && !c.layer_end.empty()) ||
            (!outputs && !c.outputs.empty()))
            return CUFHE_ERR_ARG;
        if (sizes->n_gates < static_cast<int>(c.gates.size()) ||
            sizes->n_layers < static_cast<int>(c.layer_end.size()) ||
            sizes->n_outputs < static_cast<int>(c.outputs.size()))
            return CUFHE_ERR_SIZE;
        if (!c.gates.empty())
            std::memcpy(gates, c.gates.data(), c.gates.size() * sizeof(cufhe_gate_desc));
        if (!c.layer_end.empty())
            std::memcpy(layer_end, c.layer_end.data(), c.layer_end.size() * sizeof(int));
        if (!c.outputs.empty())
            std::memcpy(outputs, c.outputs.data(), c.outputs.size() * sizeof(uint32_t));
    }
    sizes->n_inputs = c.n_inputs;
    sizes->n_gates = static_cast<int>(c.gates.size());
    sizes->n_layers = static_cast<int>(c.layer_end.size());
    sizes->n_outputs = static_cast<int>(c.outputs.size());
    return CUFHE_OK;
}

}  // namespace

extern "C" {

int cufhe_circuit_aes128_decrypt(cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                                 int* layer_end, uint32_t* outputs)
{
    static const Built c = build_aes_decrypt();
    return deliver(c, sizes, gates, layer_end, outputs);
}

int cufhe_circuit_ascon_xof(int in_bytes, int out_bytes, cufhe_circuit_sizes* sizes,
                            cufhe_gate_desc* gates, int* layer_end, uint32_t* outputs)
{
    if (in_bytes < 0 || out_bytes <= 0 || in_bytes > CUFHE_ASCON_MAX_BYTES ||
        out_bytes > CUFHE_ASCON_MAX_BYTES)
        return CUFHE_ERR_ARG;
    try {
        const Built c = build_ascon_xof(in_bytes, out_bytes);
        return deliver(c, sizes, gates, layer_end, outputs);
    } catch (const std::bad_alloc&) {
        return CUFHE_ERR_NOMEM;
    }
}

int cufhe_ascon_xof_initial_state(uint64_t* words)
{
    if (!words) return CUFHE_ERR_ARG;
    const auto s0 = p12_plain({kXofIV, 0, 0, 0, 0});
    std::copy(s0.begin(), s0.end(), words);
    return CUFHE_OK;
}

int cufhe_circuit_ascon_absorb(int in_bytes, cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                               int* layer_end, uint32_t* outputs)
{
    if (in_bytes < 0 || in_bytes > CUFHE_ASCON_MAX_BYTES) return CUFHE_ERR_ARG;
    try {
        return deliver(build_ascon_absorb(in_bytes), sizes, gates, layer_end, outputs);
    } catch (const std::bad_alloc&) {
        return CUFHE_ERR_NOMEM;
    }
}

int cufhe_circuit_ascon_squeeze(int out_bytes, cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                                int* layer_end, uint32_t* outputs)
{
    if (out_bytes <= 0 || out_bytes > CUFHE_ASCON_MAX_BYTES) return CUFHE_ERR_ARG;
    try {
        return deliver(build_ascon_squeeze(out_bytes), sizes, gates, layer_end, outputs);
    } catch (const std::bad_alloc&) {
        return CUFHE_ERR_NOMEM;
    }
}

}  // extern "C"
