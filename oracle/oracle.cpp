// CPU ORACLE for the TFHE gate-bootstrapping path -- TEST INFRASTRUCTURE.
//
// A plain restatement of the reference (cuFHEpp) arithmetic, used only by
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
// checker. It is never linked into, loaded by, or called from the product
// library (cufhe_amd/), which has no CPU fallback.
//
// Three arithmetic modes for the external product:
//   ORACLE_EXACT : the ideal integer external product (what the FFT path
//                  approximates); exact negacyclic convolutions are computed
//                  through a Goldilocks-prime NTT (bounds checked below).
//   ORACLE_FFT   : FP64 mirror of the reference FFT backend: fold, twist,
//                  H-point FFT, complex MAC, inverse, untwist, llrint
//                  (include/gatebootstrapping_gpu.cuh:69-221, :895-1033).
//   ORACLE_NTT   : the reference small-modulus NTT backend semantics
//                  (include/ntt_small_modulus.cuh:42-74, :284-361;
//                  include/gatebootstrapping_gpu.cuh:522-659, :1187-1332):
//                  exact mod-P negacyclic convolution, P = 2147473409 for
//                  N = 1024 and the Goldilocks prime for N = 2048, with the
//                  reference's torus <-> Z_P conversions.
// Integer parts follow the reference line by line: gadget digits
// (gatebootstrapping_gpu.cuh:123-129), mod switch and test vector (:10-53),
// blind-rotation drivers (:801-884, :1335-1429), sample extraction
// (src/bootstrap_gpu.cu:1201-1221), identity key switching with pre-add
// (include/keyswitch_gpu.cuh:13-256), gate compositions
// (src/bootstrap_gpu.cu:1223-1331, :1885-1946, :2069-2197).
//
// Parity status per function is stated in DESIGN.md.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

using u32 = uint32_t;
using u64 = uint64_t;
using i32 = int32_t;
using i64 = int64_t;
using u128 = unsigned __int128;
using cd = std::complex<double>;

enum Mode { ORACLE_EXACT = 0, ORACLE_FFT = 1, ORACLE_NTT = 2 };

// Planted mistakes, for the tests that show the pinning checks catch them
// (oracle_set_mutation). 0 = the faithful restatement.
enum Mutation {
    MUT_NONE = 0,
    MUT_NTT_TORUS_OFFSET = 1,  // ntt_mod_to_torus32 rounds with 2^31 instead of 2^30
    MUT_BAR_OFF_BY_ONE = 2,    // bar = 2N - ms(b) + 1
    MUT_B_ROUNDED = 3,         // a rounding offset added to b before the mod switch
    MUT_BB_SIGN = 4,           // block-binary body + correction/2 instead of -
    MUT_BB_NOT_HALVED = 5,     // block-binary body - correction
    MUT_KS_NO_ROUND = 6,       // key-switch rounding offset dropped
    MUT_KB_ORDER = 7,          // key-bundle members 3i+1 and 3i+2 swapped
    MUT_A_TRUNCATED = 8,       // mask coefficients mod-switched without rounding
    MUT_AUTO_SIGN = 9,         // automorphism without the negacyclic sign
    MUT_AH_NO_HALVE = 10,      // annihilation rounds without the halving
    MUT_AH_KEY_ROW = 11,       // annihilation digits paired with the wrong key row
};
int g_mut = MUT_NONE;

// ---------------------------------------------------------------- parameters
// Same values as include/tfhe_params.hpp (checked against the library by
// tests/test_oracle.py).
constexpr u32 N0 = 636;
constexpr u32 KS_T = 7, KS_BASEBIT = 2;
constexpr u32 BB_ELL = 4;

struct L1 {
    using T = u32;
    static constexpr u32 N = 1024, nbit = 10, l = 2, Bgbit = 8;
    static constexpr T mu = 1u << 29;
};
struct L2 {
    using T = u64;
    static constexpr u32 N = 2048, nbit = 11, l = 4, Bgbit = 10;
    static constexpr T mu = 1ULL << 61;
};

// ---------------------------------------------------------------- modular NTT
struct Field {
    u64 p;
    u64 mul(u64 a, u64 b) const
    {
        // a, b < P31 < 2^32: the product fits, and a constant divisor lets
        // the compiler reduce with a multiplication
        if (p == 2147473409ULL) return a * b % 2147473409ULL;
        return static_cast<u64>((u128)a * b % p);
    }
    u64 add(u64 a, u64 b) const
    {
        if (p < (1ULL << 63)) {
            const u64 s = a + b;
            return s >= p ? s - p : s;
        }
        u128 s = (u128)a + b;
        return static_cast<u64>(s >= p ? s - p : s);
    }
    u64 sub(u64 a, u64 b) const { return a >= b ? a - b : static_cast<u64>((u128)a + p - b); }
    u64 pow(u64 b, u64 e) const
    {
        u64 r = 1 % p;
        b %= p;
        while (e) {
            if (e & 1) r = mul(r, b);
            b = mul(b, b);
            e >>= 1;
        }
        return r;
    }
};

constexpr u64 kP31 = 2147473409ULL;                // reference small_ntt31::P
constexpr u64 kGold = 0xFFFFFFFF00000001ULL;       // Goldilocks

// Negacyclic convolution mod p via psi-twisted cyclic NTT (any primitive
// 2N-th root psi gives the same exact result).
struct NegaNTT {
    Field f;
    u32 N;
    std::vector<u64> psi, ipsi, w, iw;
    u64 ninv;
    NegaNTT(u64 p, u32 n) : f{p}, N(n)
    {
        // psi = g^((p-1)/2N) for a quadratic non-residue g (psi^N = -1)
        u64 g = 3;
        while (f.pow(g, (p - 1) / 2) != p - 1) g++;
        const u64 ps = f.pow(g, (p - 1) / (2 * N));
        psi.resize(N);
        ipsi.resize(N);
        const u64 ips = f.pow(ps, 2 * N - 1);
        psi[0] = ipsi[0] = 1;
        for (u32 i = 1; i < N; i++) {
            psi[i] = f.mul(psi[i - 1], ps);
            ipsi[i] = f.mul(ipsi[i - 1], ips);
        }
        const u64 om = f.mul(ps, ps);  // primitive N-th root
        w.resize(N / 2);
        iw.resize(N / 2);
        const u64 iom = f.pow(om, N - 1);
        w[0] = iw[0] = 1;
        for (u32 i = 1; i < N / 2; i++) {
            w[i] = f.mul(w[i - 1], om);
            iw[i] = f.mul(iw[i - 1], iom);
        }
        ninv = f.pow(N, p - 2);
    }
    void cyclic(std::vector<u64>& a, bool inv) const
    {
        // iterative radix-2 with bit reversal
        for (u32 i = 1, j = 0; i < N; i++) {
            u32 bit = N >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (u32 len = 2; len <= N; len <<= 1) {
            const u32 step = N / len;
            for (u32 s = 0; s < N; s += len)
                for (u32 k = 0; k < len / 2; k++) {
                    const u64 tw = inv ? iw[k * step] : w[k * step];
                    const u64 u = a[s + k], v = f.mul(a[s + k + len / 2], tw);
                    a[s + k] = f.add(u, v);
                    a[s + k + len / 2] = f.sub(u, v);
                }
        }
    }
    void forward(std::vector<u64>& a) const
    {
        for (u32 i = 0; i < N; i++) a[i] = f.mul(a[i], psi[i]);
        cyclic(a, false);
    }
    void inverse(std::vector<u64>& a) const
    {
        cyclic(a, true);
        for (u32 i = 0; i < N; i++) a[i] = f.mul(f.mul(a[i], ipsi[i]), ninv);
    }
};

const NegaNTT& nega(u64 p, u32 N)
{
    static const NegaNTT g1(kGold, 1024), g2(kGold, 2048), s1(kP31, 1024);
    if (p == kP31) return s1;
    return N == 1024 ? g1 : g2;
}

u64 to_mod_signed(i64 v, u64 p) { return v >= 0 ? static_cast<u64>(v) % p : p - (static_cast<u64>(-v) % p); }
i64 center(u64 v, u64 p) { return v > p / 2 ? -static_cast<i64>(p - v) : static_cast<i64>(v); }

// ---------------------------------------------------------------- torus <-> Z_P
// include/ntt_small_modulus.cuh:284-295
template <class T>
u64 torus_to_ntt_mod(T a, u64 P)
{
    constexpr int bits = std::numeric_limits<T>::digits;
    u128 prod = (u128)a * P;
    prod += (u128)1 << (bits - 1);
    return static_cast<u64>(prod >> bits);
}
// :297-306 (also correct for the Goldilocks sub, which yields p - |v|)
u64 signed_int_to_ntt_mod(i32 v, u64 P) { return v < 0 ? P - static_cast<u64>(-static_cast<i64>(v)) : static_cast<u64>(v); }
// :308-315
u64 ntt_abs_to_torus64_goldilocks(u64 val)
{
    const u128 mul = val;
    return static_cast<u64>(((mul << 64) + (mul << 32) - mul + ((u128)1 << 63)) >> 64);
}
// :317-361
u32 ntt_mod_to_torus32(u64 val, u64 P)
{
    if (P == kP31) {
        const u64 half = kP31 / 2;
        const bool neg = val > half;
        const u32 mag = static_cast<u32>(neg ? (kP31 - val) : val);
        const u64 inv_mul = (1ULL << 63) / kP31;
        const u64 temp = static_cast<u64>(mag) * inv_mul;
        const u64 half_unit = g_mut == MUT_NTT_TORUS_OFFSET ? (1ULL << 31) : (1ULL << 30);
        const u32 t32 = static_cast<u32>((temp + half_unit) >> 31);
        return neg ? static_cast<u32>(0u - t32) : t32;
    }
    const bool neg = val > kGold / 2;
    const u64 t64 = ntt_abs_to_torus64_goldilocks(neg ? (kGold - val) : val);
    u32 t32 = static_cast<u32>(t64 >> 32);
    t32 += static_cast<u32>((t64 & 0xFFFFFFFFULL) >= 0x80000000ULL);
    return neg ? static_cast<u32>(0u - t32) : t32;
}
u64 ntt_mod_to_torus64(u64 val, u64 P)
{
    if (P == kP31) {
        const u64 half = kP31 / 2;
        const bool neg = val > half;
        const u64 mag = neg ? (kP31 - val) : val;
        const u128 scaled = ((u128)mag << 64) + half;
        const u64 t64 = static_cast<u64>(scaled / kP31);
        return neg ? static_cast<u64>(0 - t64) : t64;
    }
    if (val > kGold / 2) return static_cast<u64>(0 - ntt_abs_to_torus64_goldilocks(kGold - val));
    return ntt_abs_to_torus64_goldilocks(val);
}

// ---------------------------------------------------------------- FP64 FFNT
struct FFNT {
    u32 N, H;
    std::vector<cd> twist, root;
    explicit FFNT(u32 n) : N(n), H(n / 2)
    {
        twist.resize(H);
        root.resize(H);
        for (u32 i = 0; i < H; i++) {
            twist[i] = std::polar(1.0, M_PI * i / N);
            root[i] = std::polar(1.0, -2.0 * M_PI * i / H);
        }
    }
    void fft(std::vector<cd>& a, bool inv) const
    {
        for (u32 i = 1, j = 0; i < H; i++) {
            u32 bit = H >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (u32 len = 2; len <= H; len <<= 1) {
            const u32 step = H / len;
            for (u32 s = 0; s < H; s += len)
                for (u32 k = 0; k < len / 2; k++) {
                    const cd tw = inv ? std::conj(root[k * step]) : root[k * step];
                    const cd u = a[s + k], v = a[s + k + len / 2] * tw;
                    a[s + k] = u + v;
                    a[s + k + len / 2] = u - v;
                }
        }
    }
    // real coefficients -> spectrum (fold, twist, FFT)
    std::vector<cd> forward(const double* p) const
    {
        std::vector<cd> z(H);
        for (u32 i = 0; i < H; i++) z[i] = cd(p[i], p[i + H]) * twist[i];
        fft(z, false);
        return z;
    }
    // spectrum -> real coefficients (IFFT, untwist incl. 1/H)
    void inverse(std::vector<cd> z, double* out) const
    {
        fft(z, true);
        for (u32 i = 0; i < H; i++) {
            const cd v = z[i] * std::conj(twist[i]) / static_cast<double>(H);
            out[i] = v.real();
            out[i + H] = v.imag();
        }
    }
};

const FFNT& ffnt(u32 N)
{
    static const FFNT f1(1024), f2(2048);
    return N == 1024 ? f1 : f2;
}

// double (scaled to torus units * 2^w) -> torus word, round-to-nearest,
// wrapping mod 2^w (include/utils_gpu.cuh:48-65 for Torus64).
template <class T>
T double_to_torus(double v)
{
    const double r = std::nearbyint(v);
    if constexpr (sizeof(T) == 4) {
        return static_cast<T>(static_cast<i64>(r));
    }
    else {
        if (std::fabs(r) < 9223372036854775808.0) return static_cast<T>(static_cast<i64>(r));
        int ex;
        const double m = std::frexp(std::fabs(r), &ex);  // r = m * 2^ex, m in [0.5,1)
        const u64 mant = static_cast<u64>(std::ldexp(m, 53));
        const int sh = ex - 53;
        const u64 mag = sh >= 64 ? 0 : (mant << sh);
        return r < 0 ? static_cast<T>(0 - mag) : static_cast<T>(mag);
    }
}

// ---------------------------------------------------------------- gadget
template <class P>
i32 gadget_digit(typename P::T c, u32 q)
{
    using T = typename P::T;
    constexpr int W = std::numeric_limits<T>::digits;
    T off = 0;
    for (u32 i = 1; i <= P::l; i++) off += static_cast<T>(1u << (P::Bgbit - 1)) << (W - i * P::Bgbit);
    off += static_cast<T>(1) << (W - P::l * P::Bgbit - 1);
    const T t = c + off;
    return static_cast<i32>((t >> (W - (q + 1) * P::Bgbit)) & ((1u << P::Bgbit) - 1)) -
           static_cast<i32>(1u << (P::Bgbit - 1));
}
template <class P>
typename P::T gadget_h(u32 q)
{
    using T = typename P::T;
    return static_cast<T>(1) << (std::numeric_limits<T>::digits - (q + 1) * P::Bgbit);
}

// X^a * p (negacyclic), a in [0, 2N)
template <class T>
void rotate(const T* p, u32 N, u32 a, T* out)
{
    for (u32 i = 0; i < N; i++) {
        const u32 s = (i - a) & (N - 1);
        const bool neg = (i < (a & (N - 1))) ^ ((a / N) != 0);
        out[i] = neg ? static_cast<T>(0 - p[s]) : p[s];
    }
}

// ---------------------------------------------------------------- external products
// C: TRGSW in the polynomial domain, rows*(k+1) polys of N. d: TRLWE whose
// digits are decomposed. Returns ExtProd(d, C) (k = 1).
template <class P>
void extprod_exact(const typename P::T* d, const typename P::T* C, typename P::T* out)
{
    using T = typename P::T;
    constexpr u32 N = P::N, rows = 2 * P::l;
    const NegaNTT& nt = nega(kGold, N);
    const Field& f = nt.f;
    // digit spectra
    std::vector<std::vector<u64>> D(rows, std::vector<u64>(N));
    for (u32 r = 0; r < rows; r++) {
        const u32 j = r / P::l, q = r % P::l;
        for (u32 i = 0; i < N; i++) D[r][i] = to_mod_signed(gadget_digit<P>(d[j * N + i], q), kGold);
        nt.forward(D[r]);
    }
    for (u32 o = 0; o < 2; o++) {
        // lvl1: key words as signed 32-bit; lvl2: lo/hi 32-bit halves
        const int halves = sizeof(T) == 4 ? 1 : 2;
        std::vector<i64> res[2];
        for (int hf = 0; hf < halves; hf++) {
            std::vector<u64> acc(N, 0);
            for (u32 r = 0; r < rows; r++) {
                std::vector<u64> c(N);
                const T* poly = C + (r * 2 + o) * N;
                for (u32 i = 0; i < N; i++) {
                    i64 v;
                    if constexpr (sizeof(T) == 4)
                        v = static_cast<i32>(poly[i]);
                    else
                        v = hf == 0 ? static_cast<i64>(poly[i] & 0xFFFFFFFFULL)
                                    : static_cast<i64>(poly[i] >> 32);
                    c[i] = to_mod_signed(v, kGold);
                }
                nt.forward(c);
                for (u32 i = 0; i < N; i++) acc[i] = f.add(acc[i], f.mul(c[i], D[r][i]));
            }
            nt.inverse(acc);
            res[hf].resize(N);
            for (u32 i = 0; i < N; i++) res[hf][i] = center(acc[i], kGold);
        }
        for (u32 i = 0; i < N; i++) {
            if constexpr (sizeof(T) == 4)
                out[o * N + i] = static_cast<T>(res[0][i]);
            else
                out[o * N + i] = static_cast<T>(res[0][i]) + (static_cast<T>(res[1][i]) << 32);
        }
    }
}

// FFT mirror: Cspec holds the FP64 spectra of the (combined) key rows.
template <class P>
void extprod_fft(const typename P::T* d, const std::vector<std::vector<cd>>& Cspec,
                 typename P::T* out)
{
    using T = typename P::T;
    using ST = std::make_signed_t<T>;
    constexpr u32 N = P::N, H = N / 2, rows = 2 * P::l;
    const FFNT& F = ffnt(N);
    std::vector<cd> acc[2] = {std::vector<cd>(H, 0.0), std::vector<cd>(H, 0.0)};
    std::vector<double> tmp(N);
    for (u32 r = 0; r < rows; r++) {
        const u32 j = r / P::l, q = r % P::l;
        for (u32 i = 0; i < N; i++) tmp[i] = static_cast<double>(gadget_digit<P>(d[j * N + i], q));
        const std::vector<cd> z = F.forward(tmp.data());
        for (u32 o = 0; o < 2; o++)
            for (u32 b = 0; b < H; b++) acc[o][b] += z[b] * Cspec[r * 2 + o][b];
    }
    constexpr double scale = sizeof(T) == 4 ? 4294967296.0 : 18446744073709551616.0;
    for (u32 o = 0; o < 2; o++) {
        F.inverse(acc[o], tmp.data());
        for (u32 i = 0; i < N; i++) out[o * N + i] = double_to_torus<T>(tmp[i] * scale);
    }
    (void)sizeof(ST);
}

template <class P>
std::vector<std::vector<cd>> key_spectra(const typename P::T* C)
{
    using T = typename P::T;
    using ST = std::make_signed_t<T>;
    constexpr u32 N = P::N, rows = 2 * P::l;
    constexpr double s = sizeof(T) == 4 ? 1.0 / 4294967296.0 : 1.0 / 18446744073709551616.0;
    std::vector<std::vector<cd>> out(rows * 2);
    std::vector<double> tmp(N);
    for (u32 p = 0; p < rows * 2; p++) {
        for (u32 i = 0; i < N; i++) tmp[i] = static_cast<double>(static_cast<ST>(C[p * N + i])) * s;
        out[p] = ffnt(N).forward(tmp.data());
    }
    return out;
}

std::vector<cd> xai_spectrum(u32 N, u32 a)
{
    std::vector<double> p(N, 0.0);
    p[0] = -1.0;
    if (a < N)
        p[a] += 1.0;
    else
        p[a - N] -= 1.0;
    return ffnt(N).forward(p.data());
}

// NTT backend: exact mod-P with the reference's conversions; Cmod holds the
// key rows already mapped to Z_P (torus_to_ntt_mod) in the coefficient domain.
template <class P>
void extprod_ntt(const typename P::T* d, const std::vector<std::vector<u64>>& Cmod, u64 Pm,
                 typename P::T* out, bool add)
{
    using T = typename P::T;
    constexpr u32 N = P::N, rows = 2 * P::l;
    const NegaNTT& nt = nega(Pm, N);
    const Field& f = nt.f;
    std::vector<u64> acc[2] = {std::vector<u64>(N, 0), std::vector<u64>(N, 0)};
    for (u32 r = 0; r < rows; r++) {
        const u32 j = r / P::l, q = r % P::l;
        std::vector<u64> D(N);
        for (u32 i = 0; i < N; i++) D[i] = signed_int_to_ntt_mod(gadget_digit<P>(d[j * N + i], q), Pm);
        nt.forward(D);
        for (u32 o = 0; o < 2; o++) {
            std::vector<u64> c = Cmod[r * 2 + o];
            for (u32 i = 0; i < N; i++) c[i] %= Pm;
            nt.forward(c);
            for (u32 i = 0; i < N; i++) acc[o][i] = f.add(acc[o][i], f.mul(c[i], D[i]));
        }
    }
    for (u32 o = 0; o < 2; o++) {
        nt.inverse(acc[o]);
        for (u32 i = 0; i < N; i++) {
            T v;
            if constexpr (sizeof(T) == 4)
                v = ntt_mod_to_torus32(acc[o][i], Pm);
            else
                v = ntt_mod_to_torus64(acc[o][i], Pm);
            out[o * N + i] = add ? static_cast<T>(out[o * N + i] + v) : v;
        }
    }
}

template <class P>
u64 ntt_prime() { return P::N == 1024 ? kP31 : kGold; }

template <class P>
std::vector<std::vector<u64>> key_mod(const typename P::T* C)
{
    constexpr u32 N = P::N, rows = 2 * P::l;
    const u64 Pm = ntt_prime<P>();
    std::vector<std::vector<u64>> out(rows * 2, std::vector<u64>(N));
    for (u32 p = 0; p < rows * 2; p++)
        for (u32 i = 0; i < N; i++) out[p][i] = torus_to_ntt_mod<typename P::T>(C[p * N + i], Pm);
    return out;
}

// Spectral form of the NTT backend for whole rotations: the key polynomials
// are mapped to Z_P (torus_to_ntt_mod) and transformed once per call, and the
// X^a - 1 factors are multiplied in the transform domain. Arithmetic mod P is
// exact and NTT(x) . NTT(y) = NTT(x * y) in Z_P[X]/(X^N + 1), so the results
// are the same residues as the coefficient-domain combination of the
// reference (include/gatebootstrapping_gpu.cuh:747-761, :1187-1332).
template <class P>
std::vector<u64> key_spectra_ntt(const typename P::T* bk, size_t ntrgsw)
{
    using T = typename P::T;
    constexpr u32 N = P::N, polys = 4 * P::l;
    const u64 Pm = ntt_prime<P>();
    const NegaNTT& nt = nega(Pm, N);
    std::vector<u64> out(ntrgsw * polys * N);
    std::vector<u64> c(N);
    for (size_t t = 0; t < ntrgsw * polys; t++) {
        const T* src = bk + t * N;
        for (u32 i = 0; i < N; i++) c[i] = torus_to_ntt_mod<T>(src[i], Pm) % Pm;
        nt.forward(c);
        std::copy(c.begin(), c.end(), out.begin() + t * N);
    }
    return out;
}

// Spectrum of X^a - 1 (a in [0, 2N)) for the NTT prime of ring size N.
const u64* xai_ntt(u32 N, u32 a)
{
    struct Tab {
        std::once_flag once;
        std::vector<u64> v;
    };
    static Tab t1, t2;
    Tab& t = N == 1024 ? t1 : t2;
    const u64 Pm = N == 1024 ? kP31 : kGold;
    std::call_once(t.once, [&]() {
        const NegaNTT& nt = nega(Pm, N);
        t.v.resize(static_cast<size_t>(2 * N) * N);
        std::vector<u64> c(N);
        for (u32 e = 0; e < 2 * N; e++) {
            std::fill(c.begin(), c.end(), 0);
            // X^e - 1 (X^e = -X^(e-N) for e >= N); field ops, as c[0] = P - 1
            // plus P - 1 overflows 64 bits for the Goldilocks prime (e = N)
            c[0] = Pm - 1;
            if (e < N)
                c[e] = nt.f.add(c[e], 1);
            else
                c[e - N] = nt.f.sub(c[e - N], 1);
            nt.forward(c);
            std::copy(c.begin(), c.end(), t.v.begin() + static_cast<size_t>(e) * N);
        }
    });
    return t.v.data() + static_cast<size_t>(a) * N;
}

// out (+)= ExtProd(d, K) with K given as spectra (rows * 2 polys of N bins).
template <class P>
void extprod_ntt_spec(const typename P::T* d, const u64* K, typename P::T* out, bool add)
{
    using T = typename P::T;
    constexpr u32 N = P::N, rows = 2 * P::l;
    const u64 Pm = ntt_prime<P>();
    const NegaNTT& nt = nega(Pm, N);
    const Field& f = nt.f;
    std::vector<u64> acc[2] = {std::vector<u64>(N, 0), std::vector<u64>(N, 0)};
    std::vector<u64> D(N);
    for (u32 r = 0; r < rows; r++) {
        const u32 j = r / P::l, q = r % P::l;
        for (u32 i = 0; i < N; i++) D[i] = signed_int_to_ntt_mod(gadget_digit<P>(d[j * N + i], q), Pm);
        nt.forward(D);
        for (u32 o = 0; o < 2; o++) {
            const u64* k = K + static_cast<size_t>(r * 2 + o) * N;
            for (u32 i = 0; i < N; i++) acc[o][i] = f.add(acc[o][i], f.mul(k[i], D[i]));
        }
    }
    for (u32 o = 0; o < 2; o++) {
        nt.inverse(acc[o]);
        for (u32 i = 0; i < N; i++) {
            T v;
            if constexpr (sizeof(T) == 4)
                v = ntt_mod_to_torus32(acc[o][i], Pm);
            else
                v = ntt_mod_to_torus64(acc[o][i], Pm);
            out[o * N + i] = add ? static_cast<T>(out[o * N + i] + v) : v;
        }
    }
}

// ---------------------------------------------------------------- blind rotation
u32 mod_switch(u32 phase, u32 nbit) { return phase >> (32 - 1 - nbit); }

// Blind-rotation variants: plain CMUX per key bit, 2-bit key bundle
// (gatebootstrapping_gpu.cuh:1187-1332), block-binary key in blocks of
// BB_ELL bits (gatebootstrapping_gpu.cuh:661-797, :801-884).
enum Variant { BR_PLAIN = 0, BR_KB = 1, BR_BB = 2 };

// Block-binary body: b - correction/2 rounded, correction = sum_i (a_i -
// round(a_i)) as a wrapped signed 32-bit value (gatebootstrapping_gpu.cuh:814-838).
template <class P>
u32 bb_corrected_body(const u32* tl0)
{
    constexpr u32 shift = 32 - 1 - P::nbit;
    constexpr u32 rnd = 1u << (shift - 1);
    u32 sum = 0;
    for (u32 i = 0; i < N0; i++) sum += tl0[i] - ((mod_switch(tl0[i] + rnd, P::nbit)) << shift);
    const i32 correction = static_cast<i32>(sum);
    if (g_mut == MUT_BB_SIGN) return tl0[N0] + static_cast<u32>(correction / 2) + rnd;
    if (g_mut == MUT_BB_NOT_HALVED) return tl0[N0] - static_cast<u32>(correction) + rnd;
    return tl0[N0] - static_cast<u32>(correction / 2) + rnd;
}

// Number of TRGSWs of a lvl0 -> P bootstrapping key of the given variant.
inline size_t bk_trgsws(int variant) { return variant == BR_KB ? 3 * (N0 / 2) : N0; }

// kspec: NTT mode only, key_spectra_ntt of bk (computed here when null).
template <class P>
void blind_rotate(int mode, int variant, const u32* tl0, const typename P::T* bk, typename P::T mu,
                  typename P::T* acc, const std::vector<u64>* kspec = nullptr)
{
    using T = typename P::T;
    constexpr u32 N = P::N, rows = 2 * P::l, polys = 2 * rows;
    constexpr size_t TW = static_cast<size_t>(rows) * 2 * N;
    const u32 rnd = g_mut == MUT_A_TRUNCATED ? 0u : 1u << (32 - 2 - P::nbit);
    const bool kb = variant == BR_KB;
    std::vector<u64> kspec_local;
    if (mode == ORACLE_NTT && !kspec) {
        kspec_local = key_spectra_ntt<P>(bk, bk_trgsws(variant));
        kspec = &kspec_local;
    }
    auto kpoly = [&](size_t t) { return kspec->data() + t * polys * N; };
    // test vector (gatebootstrapping_gpu.cuh:31-53)
    u32 body = variant == BR_BB ? bb_corrected_body<P>(tl0) : tl0[N0];
    if (g_mut == MUT_B_ROUNDED && variant != BR_BB) body += 1u << (32 - 2 - P::nbit);
    u32 bar = 2 * N - mod_switch(body, P::nbit);
    if (g_mut == MUT_BAR_OFF_BY_ONE) bar = (bar + 1) % (2 * N);
    for (u32 i = 0; i < N; i++) {
        acc[i] = 0;
        if (bar == 2 * N)
            acc[N + i] = mu;
        else
            acc[N + i] = ((i < (bar & (N - 1))) ^ ((bar >> P::nbit) != 0)) ? static_cast<T>(0 - mu) : mu;
    }
    std::vector<T> tmp(2 * N), rot(2 * N), comb(TW);
    std::vector<u64> cm(kspec ? static_cast<size_t>(polys) * N : 0);
    if (kb) {
        for (u32 it = 0; it < N0 / 2; it++) {
            const u32 a0 = mod_switch(tl0[2 * it] + rnd, P::nbit);
            const u32 a1 = mod_switch(tl0[2 * it + 1] + rnd, P::nbit);
            const u32 a01 = (a0 + a1) & (2 * N - 1);
            size_t m0 = 3 * it + 0, m1 = 3 * it + 1, m2 = 3 * it + 2;
            if (g_mut == MUT_KB_ORDER) std::swap(m1, m2);
            const T* b0 = bk + m0 * TW;
            const T* b1 = bk + m1 * TW;
            const T* b2 = bk + m2 * TW;
            if (mode == ORACLE_FFT) {
                // combined = one + b2*xai[a1] + b1*xai[a0] + b0*xai[a01] in FP64
                const auto s0 = key_spectra<P>(b0), s1 = key_spectra<P>(b1), s2 = key_spectra<P>(b2);
                const auto x0 = xai_spectrum(N, a0), x1 = xai_spectrum(N, a1), x01 = xai_spectrum(N, a01);
                std::vector<std::vector<cd>> cs(rows * 2, std::vector<cd>(N / 2));
                for (u32 r = 0; r < rows; r++)
                    for (u32 o = 0; o < 2; o++) {
                        const double h = (o == r / P::l)
                                             ? std::ldexp(1.0, -static_cast<int>((r % P::l + 1) * P::Bgbit))
                                             : 0.0;
                        for (u32 b = 0; b < N / 2; b++) {
                            cd v(h, 0.0);
                            v += s2[r * 2 + o][b] * x1[b];
                            v += s1[r * 2 + o][b] * x0[b];
                            v += s0[r * 2 + o][b] * x01[b];
                            cs[r * 2 + o][b] = v;
                        }
                    }
                extprod_fft<P>(acc, cs, tmp.data());
                std::memcpy(acc, tmp.data(), 2 * N * sizeof(T));
                continue;
            }
            if (mode == ORACLE_NTT) {
                // one + b2 (X^a1-1) + b1 (X^a0-1) + b0 (X^a01-1) in Z_P
                const u64 Pm = ntt_prime<P>();
                const Field f{Pm};
                const u64 *k0 = kpoly(m0), *k1 = kpoly(m1), *k2 = kpoly(m2);
                const u64 *x0 = xai_ntt(N, a0), *x1 = xai_ntt(N, a1), *x01 = xai_ntt(N, a01);
                for (u32 p = 0; p < polys; p++) {
                    const u32 r = p / 2, o = p % 2;
                    // the identity term is a constant polynomial: the same value in every bin
                    const u64 one =
                        o == r / P::l ? torus_to_ntt_mod<T>(gadget_h<P>(r % P::l), Pm) % Pm : 0;
                    const size_t off = static_cast<size_t>(p) * N;
                    for (u32 i = 0; i < N; i++) {
                        u64 v = one;
                        v = f.add(v, f.mul(k2[off + i], x1[i]));
                        v = f.add(v, f.mul(k1[off + i], x0[i]));
                        v = f.add(v, f.mul(k0[off + i], x01[i]));
                        cm[off + i] = v;
                    }
                }
                extprod_ntt_spec<P>(acc, cm.data(), acc, false);
                continue;
            }
            // exact: combined key in the polynomial domain mod 2^w
            for (size_t i = 0; i < TW; i++) comb[i] = 0;
            for (u32 r = 0; r < rows; r++) comb[(r * 2 + r / P::l) * N] = gadget_h<P>(r % P::l);
            auto acc_rot = [&](const T* src, u32 a) {
                for (u32 p = 0; p < rows * 2; p++) {
                    rotate<T>(src + p * N, N, a, rot.data());
                    for (u32 i = 0; i < N; i++) comb[p * N + i] += rot[i] - src[p * N + i];
                }
            };
            acc_rot(b2, a1);
            acc_rot(b1, a0);
            acc_rot(b0, a01);
            extprod_exact<P>(acc, comb.data(), tmp.data());
            std::memcpy(acc, tmp.data(), 2 * N * sizeof(T));
        }
        return;
    }
    if (variant == BR_BB) {
        // acc += ExtProd(acc, sum_e BK[ell b + e] (X^a_e - 1)), no identity term
        for (u32 blk = 0; blk < N0 / BB_ELL; blk++) {
            u32 a[BB_ELL];
            for (u32 e = 0; e < BB_ELL; e++) a[e] = mod_switch(tl0[blk * BB_ELL + e] + rnd, P::nbit);
            if (mode == ORACLE_NTT) {
                const Field f{ntt_prime<P>()};
                std::fill(cm.begin(), cm.end(), 0);
                for (u32 e = 0; e < BB_ELL; e++) {
                    const u64* k = kpoly(blk * BB_ELL + e);
                    const u64* x = xai_ntt(N, a[e]);
                    for (u32 p = 0; p < polys; p++) {
                        const size_t off = static_cast<size_t>(p) * N;
                        for (u32 i = 0; i < N; i++) cm[off + i] = f.add(cm[off + i], f.mul(k[off + i], x[i]));
                    }
                }
                for (u32 i = 0; i < 2 * N; i++) tmp[i] = 0;
                extprod_ntt_spec<P>(acc, cm.data(), tmp.data(), true);
            }
            else {
                for (size_t i = 0; i < TW; i++) comb[i] = 0;
                for (u32 e = 0; e < BB_ELL; e++) {
                    const T* src = bk + (blk * BB_ELL + e) * TW;
                    for (u32 p = 0; p < rows * 2; p++) {
                        rotate<T>(src + p * N, N, a[e], rot.data());
                        for (u32 i = 0; i < N; i++) comb[p * N + i] += rot[i] - src[p * N + i];
                    }
                }
                extprod_exact<P>(acc, comb.data(), tmp.data());
            }
            for (u32 i = 0; i < 2 * N; i++) acc[i] += tmp[i];
        }
        return;
    }
    for (u32 it = 0; it < N0; it++) {
        const u32 a = mod_switch(tl0[it] + rnd, P::nbit);
        const T* C = bk + it * TW;
        // d = X^a acc - acc
        for (u32 j = 0; j < 2; j++) {
            rotate<T>(acc + j * N, N, a, rot.data() + j * N);
            for (u32 i = 0; i < N; i++) rot[j * N + i] -= acc[j * N + i];
        }
        if (mode == ORACLE_FFT)
            extprod_fft<P>(rot.data(), key_spectra<P>(C), tmp.data());
        else if (mode == ORACLE_NTT) {
            for (u32 i = 0; i < 2 * N; i++) tmp[i] = 0;
            extprod_ntt_spec<P>(rot.data(), kpoly(it), tmp.data(), true);
        }
        else
            extprod_exact<P>(rot.data(), C, tmp.data());
        for (u32 i = 0; i < 2 * N; i++) acc[i] += tmp[i];
    }
}

template <class T>
void sample_extract(const T* trlwe, u32 N, T* tlwe)
{
    tlwe[0] = trlwe[0];
    for (u32 i = 1; i < N; i++) tlwe[i] = static_cast<T>(0 - trlwe[N - i]);
    tlwe[N] = trlwe[N];
}

// ---------------------------------------------------------------- key switching
template <class DT>
void identity_key_switch(u32 dn, bool subset, const DT* in0, const DT* in1, i32 ca, i32 cb,
                         DT offset, const u32* ksk, u32* out)
{
    constexpr int W = std::numeric_limits<DT>::digits;
    constexpr u32 numbase = 1u << (KS_BASEBIT - 1);
    constexpr u32 n0 = N0;
    DT decomp = 0;
    for (u32 i = 1; i <= KS_T; i++) decomp += static_cast<DT>(numbase) << (W - i * KS_BASEBIT);
    const DT round = g_mut == MUT_KS_NO_ROUND ? 0 : static_cast<DT>(1) << (W - (1 + KS_BASEBIT * KS_T));
    auto conv = [](DT v) -> u32 {
        if constexpr (W == 32)
            return static_cast<u32>(v);
        else
            return static_cast<u32>((v + (static_cast<DT>(1) << 31)) >> 32);
    };
    const DT ua = static_cast<DT>(static_cast<std::make_signed_t<DT>>(ca));
    const DT ub = static_cast<DT>(static_cast<std::make_signed_t<DT>>(cb));
    auto comb = [&](u32 j) -> DT {
        DT v = ua * in0[j];
        if (in1) v += ub * in1[j];
        return v;
    };
    const u32 first = subset ? n0 : 0;
    for (u32 i = 0; i <= n0; i++) {
        u32 res = 0;
        if (i == n0)
            res = conv(comb(dn) + offset);
        else if (subset)
            res = conv(comb(i));
        for (u32 j = first; j < dn; j++) {
            const DT tmp = comb(j) + decomp + round;
            for (u32 k = 0; k < KS_T; k++) {
                const i32 val = static_cast<i32>((tmp >> (W - (k + 1) * KS_BASEBIT)) &
                                                 ((static_cast<DT>(1) << KS_BASEBIT) - 1)) -
                                static_cast<i32>(numbase);
                if (val == 0) continue;
                const u32 x = ksk[((static_cast<size_t>(j - first) * KS_T + k) * numbase +
                                   (std::abs(val) - 1)) *
                                      (n0 + 1) +
                                  i];
                res = val > 0 ? res - x : res + x;
            }
        }
        out[i] = res;
    }
}

// ---------------------------------------------------------------- annihilation
// p(X) -> p(X^d) in Z_{2^w}[X]/(X^N + 1) for odd d: coefficient i moves to
// i*d mod 2N, negated when that lands in [N, 2N) (reference
// src/annihilate_gpu.cu:204-219 __AutomorphismPolynomial__).
template <class T>
void automorphism(const T* in, u32 N, u32 d, T* out)
{
    for (u32 i = 0; i < N; i++) {
        const u32 idx = i * d;
        const bool neg = (idx & N) != 0 && g_mut != MUT_AUTO_SIGN;
        out[idx & (N - 1)] = neg ? static_cast<T>(0 - in[i]) : in[i];
    }
}

// Annihilation key switching of one TRLWE (reference src/annihilate_gpu.cu
// :676-703 and the kernels it launches: :180-202 halving, :436-462 EvalAuto,
// :196-202 the sum). nbit rounds, round `bit` with d = 2^(bit+1) + 1:
//   c >>= 1 (every word, logical shift of the unsigned torus word);
//   c += EvalAuto_d(c) = (0, tau_d(b)) - ExtProd_half(tau_d(a), AHK[bit]),
// where the half external product decomposes tau_d(a) alone over the l rows
// (each 2 polynomials) of AHK[bit] (:245-434). That equals the full external
// product of the TRLWE (tau_d(a), 0) with a TRGSW whose rows l .. 2l-1 are
// zero (every gadget digit of the zero polynomial is 0): the exact and NTT
// modes compute it that way, the FFT mode skips the zero rows
// (extprod_half_fft). ahk: nbit x l x 2 polynomials.
// FP64 half external product (FFT mode): the rows r < l of extprod_fft in the
// same order; the skipped rows l .. 2l-1 multiply zero digit spectra, whose
// products would add exact zeros.
template <class P>
void extprod_half_fft(const typename P::T* t, const std::vector<std::vector<cd>>& Cspec,
                      typename P::T* out)
{
    using T = typename P::T;
    constexpr u32 N = P::N, H = N / 2;
    const FFNT& F = ffnt(N);
    std::vector<cd> acc[2] = {std::vector<cd>(H, 0.0), std::vector<cd>(H, 0.0)};
    std::vector<double> tmp(N);
    for (u32 q = 0; q < P::l; q++) {
        for (u32 i = 0; i < N; i++) tmp[i] = static_cast<double>(gadget_digit<P>(t[i], q));
        const std::vector<cd> z = F.forward(tmp.data());
        for (u32 o = 0; o < 2; o++)
            for (u32 b = 0; b < H; b++) acc[o][b] += z[b] * Cspec[q * 2 + o][b];
    }
    constexpr double scale = sizeof(T) == 4 ? 4294967296.0 : 18446744073709551616.0;
    for (u32 o = 0; o < 2; o++) {
        F.inverse(acc[o], tmp.data());
        for (u32 i = 0; i < N; i++) out[o * N + i] = double_to_torus<T>(tmp[i] * scale);
    }
}

// count TRLWEs, one after the other; the key is prepared once per call
// (FFT mode: its spectra; NTT mode: its rows mapped to Z_P).
template <class P>
void annihilate(int mode, const typename P::T* in, const typename P::T* ahk, typename P::T* out,
                size_t count)
{
    using T = typename P::T;
    constexpr u32 N = P::N, L = P::l;
    std::vector<std::vector<T>> Cfull(P::nbit, std::vector<T>(4 * L * N, 0));
    std::vector<std::vector<std::vector<cd>>> spec(P::nbit);
    std::vector<std::vector<std::vector<u64>>> kmod(P::nbit);
    for (u32 bit = 0; bit < P::nbit; bit++) {
        const T* kb = ahk + static_cast<size_t>(bit) * 2 * L * N;
        for (u32 q = 0; q < L; q++) {
            const u32 src = g_mut == MUT_AH_KEY_ROW ? (q + 1) % L : q;
            std::copy(kb + src * 2 * N, kb + (src + 1) * 2 * N, Cfull[bit].begin() + q * 2 * N);
        }
        if (mode == ORACLE_FFT) spec[bit] = key_spectra<P>(Cfull[bit].data());
        if (mode == ORACLE_NTT) kmod[bit] = key_mod<P>(Cfull[bit].data());
    }
    std::vector<T> c(2 * N), d(2 * N, 0), tb(N), prod(2 * N);
    for (size_t j = 0; j < count; j++) {
        std::copy(in + j * 2 * N, in + (j + 1) * 2 * N, c.begin());
        for (u32 bit = 0; bit < P::nbit; bit++) {
            if (g_mut != MUT_AH_NO_HALVE)
                for (auto& v : c) v >>= 1;
            const u32 dd = (2u << bit) + 1;
            automorphism<T>(c.data(), N, dd, d.data());
            automorphism<T>(c.data() + N, N, dd, tb.data());
            if (mode == ORACLE_FFT)
                extprod_half_fft<P>(d.data(), spec[bit], prod.data());
            else if (mode == ORACLE_NTT) {
                std::fill(prod.begin(), prod.end(), T(0));
                extprod_ntt<P>(d.data(), kmod[bit], ntt_prime<P>(), prod.data(), true);
            }
            else
                extprod_exact<P>(d.data(), Cfull[bit].data(), prod.data());
            for (u32 i = 0; i < N; i++) {
                c[i] += static_cast<T>(0 - prod[i]);
                c[N + i] += static_cast<T>(tb[i] - prod[N + i]);
            }
        }
        std::copy(c.begin(), c.end(), out + j * 2 * N);
    }
}

struct GateLin {
    i32 ca, cb, off;
};
// (ca, cb, offset in units of mu): src/bootstrap_gpu.cu:1335-1881, :1905-1946
bool gate_lin(int op, GateLin& g)
{
    static const GateLin t[10] = {{-1, -1, 1}, {1, 1, 1},   {1, -1, 1},  {-1, 1, 1},
                                  {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1},
                                  {2, 2, 2},   {-2, -2, -2}};
    if (op < 0 || op > 9) return false;
    g = t[op];
    return true;
}

}  // namespace

extern "C" {

// ---- primitives (for tests) ------------------------------------------------
int oracle_version(void) { return 1; }

// out = a * b mod (X^N + 1, p) for p in {P31 (N=1024), Goldilocks}; a, b
// canonical residues. Brute force when brute != 0.
int oracle_negacyclic_mod(u64 p, u32 N, const u64* a, const u64* b, u64* out, int brute)
{
    if (!((p == kP31 && N == 1024) || (p == kGold && (N == 1024 || N == 2048)))) return -1;
    Field f{p};
    if (brute) {
        for (u32 i = 0; i < N; i++) out[i] = 0;
        for (u32 i = 0; i < N; i++)
            for (u32 j = 0; j < N; j++) {
                const u64 v = f.mul(a[i], b[j]);
                const u32 k = i + j;
                if (k < N)
                    out[k] = f.add(out[k], v);
                else
                    out[k - N] = f.sub(out[k - N], v);
            }
        return 0;
    }
    const NegaNTT& nt = nega(p, N);
    std::vector<u64> A(a, a + N), B(b, b + N);
    nt.forward(A);
    nt.forward(B);
    for (u32 i = 0; i < N; i++) A[i] = f.mul(A[i], B[i]);
    nt.inverse(A);
    for (u32 i = 0; i < N; i++) out[i] = A[i];
    return 0;
}

int oracle_torus_to_ntt_mod32(const u32* a, size_t n, u64 p, u64* out)
{
    for (size_t i = 0; i < n; i++) out[i] = torus_to_ntt_mod<u32>(a[i], p);
    return 0;
}
int oracle_torus_to_ntt_mod64(const u64* a, size_t n, u64 p, u64* out)
{
    for (size_t i = 0; i < n; i++) out[i] = torus_to_ntt_mod<u64>(a[i], p);
    return 0;
}
int oracle_ntt_mod_to_torus32(const u64* v, size_t n, u64 p, u32* out)
{
    for (size_t i = 0; i < n; i++) out[i] = ntt_mod_to_torus32(v[i], p);
    return 0;
}
int oracle_ntt_mod_to_torus64(const u64* v, size_t n, u64 p, u64* out)
{
    for (size_t i = 0; i < n; i++) out[i] = ntt_mod_to_torus64(v[i], p);
    return 0;
}

// FP64 FFNT: spectrum (natural bin order k, Z[k] = p(exp(i*pi*(1-4k)/N)))
// as interleaved doubles; and its inverse.
int oracle_ffnt_forward(u32 N, const double* poly, double* spec)
{
    if (N != 1024 && N != 2048) return -1;
    const auto z = ffnt(N).forward(poly);
    for (u32 k = 0; k < N / 2; k++) {
        spec[2 * k] = z[k].real();
        spec[2 * k + 1] = z[k].imag();
    }
    return 0;
}
int oracle_ffnt_inverse(u32 N, const double* spec, double* poly)
{
    if (N != 1024 && N != 2048) return -1;
    std::vector<cd> z(N / 2);
    for (u32 k = 0; k < N / 2; k++) z[k] = cd(spec[2 * k], spec[2 * k + 1]);
    ffnt(N).inverse(z, poly);
    return 0;
}

int oracle_gadget_digits(int level, const void* coef, size_t n, i32* digits)
{
    // digits[i*l + q]
    if (level == 1) {
        const u32* c = static_cast<const u32*>(coef);
        for (size_t i = 0; i < n; i++)
            for (u32 q = 0; q < L1::l; q++) digits[i * L1::l + q] = gadget_digit<L1>(c[i], q);
        return 0;
    }
    if (level == 2) {
        const u64* c = static_cast<const u64*>(coef);
        for (size_t i = 0; i < n; i++)
            for (u32 q = 0; q < L2::l; q++) digits[i * L2::l + q] = gadget_digit<L2>(c[i], q);
        return 0;
    }
    return -1;
}

// acc += ExtProd((X^abar - 1) acc, C): one CMUX accumulate step.
int oracle_cmux_accumulate(int level, int mode, void* trlwe, const void* trgsw, u32 abar)
{
    auto run = [&](auto tag) {
        using P = decltype(tag);
        using T = typename P::T;
        constexpr u32 N = P::N;
        T* acc = static_cast<T*>(trlwe);
        const T* C = static_cast<const T*>(trgsw);
        std::vector<T> rot(2 * N), tmp(2 * N, 0);
        for (u32 j = 0; j < 2; j++) {
            rotate<T>(acc + j * N, N, abar, rot.data() + j * N);
            for (u32 i = 0; i < N; i++) rot[j * N + i] -= acc[j * N + i];
        }
        if (mode == ORACLE_FFT)
            extprod_fft<P>(rot.data(), key_spectra<P>(C), tmp.data());
        else if (mode == ORACLE_NTT)
            extprod_ntt<P>(rot.data(), key_mod<P>(C), ntt_prime<P>(), tmp.data(), true);
        else
            extprod_exact<P>(rot.data(), C, tmp.data());
        for (u32 i = 0; i < 2 * N; i++) acc[i] += tmp[i];
    };
    if (abar >= 2 * (level == 1 ? L1::N : L2::N)) return -1;
    if (level == 1)
        run(L1{});
    else if (level == 2)
        run(L2{});
    else
        return -1;
    return 0;
}

// ExtProd(d, C) (no rotation), the external product itself.
int oracle_external_product(int level, int mode, const void* d, const void* trgsw, void* out)
{
    auto run = [&](auto tag) {
        using P = decltype(tag);
        using T = typename P::T;
        const T* C = static_cast<const T*>(trgsw);
        T* o = static_cast<T*>(out);
        if (mode == ORACLE_FFT)
            extprod_fft<P>(static_cast<const T*>(d), key_spectra<P>(C), o);
        else if (mode == ORACLE_NTT) {
            for (u32 i = 0; i < 2 * P::N; i++) o[i] = 0;
            extprod_ntt<P>(static_cast<const T*>(d), key_mod<P>(C), ntt_prime<P>(), o, true);
        }
        else
            extprod_exact<P>(static_cast<const T*>(d), C, o);
    };
    if (level == 1)
        run(L1{});
    else if (level == 2)
        run(L2{});
    else
        return -1;
    return 0;
}

// Blind rotation of a lvl0 TLWE into a TRLWE of the target level; kb is the
// variant (0 plain, 1 key bundle, 2 block-binary: exact and NTT modes only).
int oracle_blind_rotate(int level, int mode, int kb, const u32* tl0, const void* bk, u64 mu,
                        void* trlwe_out)
{
    if (kb < 0 || kb > 2 || (kb == BR_BB && mode == ORACLE_FFT)) return -1;
    if (level == 1)
        blind_rotate<L1>(mode, kb, tl0, static_cast<const u32*>(bk), static_cast<u32>(mu),
                         static_cast<u32*>(trlwe_out));
    else if (level == 2)
        blind_rotate<L2>(mode, kb, tl0, static_cast<const u64*>(bk), mu,
                         static_cast<u64*>(trlwe_out));
    else
        return -1;
    return 0;
}

int oracle_sample_extract(int level, const void* trlwe, void* tlwe)
{
    if (level == 1)
        sample_extract<u32>(static_cast<const u32*>(trlwe), L1::N, static_cast<u32*>(tlwe));
    else if (level == 2)
        sample_extract<u64>(static_cast<const u64*>(trlwe), L2::N, static_cast<u64*>(tlwe));
    else
        return -1;
    return 0;
}

// Identity key switch with pre-add (domain level 1: uint32 in, 2: uint64 in).
int oracle_identity_key_switch(int level, int subset, const void* in0, const void* in1, i32 ca,
                               i32 cb, u64 offset, const u32* ksk, u32* out)
{
    if (level == 1)
        identity_key_switch<u32>(L1::N, subset != 0, static_cast<const u32*>(in0),
                                 static_cast<const u32*>(in1), ca, cb, static_cast<u32>(offset), ksk,
                                 out);
    else if (level == 2 && !subset)
        identity_key_switch<u64>(L2::N, false, static_cast<const u64*>(in0),
                                 static_cast<const u64*>(in1), ca, cb, offset, ksk, out);
    else
        return -1;
    return 0;
}

// Automorphism X -> X^d (d odd) of one polynomial of the level's ring.
int oracle_automorphism(int level, const void* in, u32 d, void* out)
{
    if ((d & 1u) == 0) return -1;
    if (level == 1)
        automorphism<u32>(static_cast<const u32*>(in), L1::N, d, static_cast<u32*>(out));
    else if (level == 2)
        automorphism<u64>(static_cast<const u64*>(in), L2::N, d, static_cast<u64*>(out));
    else
        return -1;
    return 0;
}

// Annihilation key switching of count TRLWEs (a then b each) with the key in
// the polynomial domain (nbit x l x 2 polynomials).
int oracle_annihilate(int level, int mode, const void* in, const void* ahk, void* out,
                      size_t count)
{
    if (mode < ORACLE_EXACT || mode > ORACLE_NTT) return -1;
    if (level == 1)
        annihilate<L1>(mode, static_cast<const u32*>(in), static_cast<const u32*>(ahk),
                       static_cast<u32*>(out), count);
    else if (level == 2)
        annihilate<L2>(mode, static_cast<const u64*>(in), static_cast<const u64*>(ahk),
                       static_cast<u64*>(out), count);
    else
        return -1;
    return 0;
}

// lvl1 gate API (IKS-BR order): lvl1 TLWEs in and out. kb: blind-rotation
// variant (0 plain, 1 key bundle, 2 block-binary) plus 4 for a subset KSK.
}  // extern "C"

namespace {
int gate_lvl1_impl(int mode, int kb, int op, const u32* in0, const u32* in1, const u32* inc,
                   const u32* bk, const u32* ksk, u32* out, const std::vector<u64>* kspec)
{
    if ((kb & 3) == BR_BB && mode == ORACLE_FFT) return -1;
    const bool subset = (kb & 4) != 0;
    constexpr u32 N = L1::N, mu = L1::mu;
    std::vector<u32> tl0(N0 + 1), acc(2 * N), r1(N + 1), r0(N + 1);
    if (op == 12 || op == 13) {  // NOT / COPY
        for (u32 i = 0; i <= N; i++) out[i] = op == 12 ? 0u - in0[i] : in0[i];
        return 0;
    }
    if (op == 10 || op == 11) {  // MUX / NMUX
        identity_key_switch<u32>(N, subset, inc, in1, 1, 1, 0u - mu, ksk, tl0.data());
        blind_rotate<L1>(mode, kb & 3, tl0.data(), bk, mu, acc.data(), kspec);
        sample_extract<u32>(acc.data(), N, r1.data());
        identity_key_switch<u32>(N, subset, inc, in0, -1, 1, 0u - mu, ksk, tl0.data());
        blind_rotate<L1>(mode, kb & 3, tl0.data(), bk, mu, acc.data(), kspec);
        sample_extract<u32>(acc.data(), N, r0.data());
        for (u32 i = 0; i <= N; i++) {
            u32 v = r1[i] + r0[i];
            if (i == N) v += mu;
            out[i] = op == 10 ? v : 0u - v;
        }
        return 0;
    }
    GateLin g;
    if (!gate_lin(op, g)) return -1;
    identity_key_switch<u32>(N, subset, in0, in1, g.ca, g.cb, static_cast<u32>(g.off) * mu, ksk,
                             tl0.data());
    blind_rotate<L1>(mode, kb & 3, tl0.data(), bk, mu, acc.data(), kspec);
    sample_extract<u32>(acc.data(), N, out);
    return 0;
}

}  // namespace

extern "C" {

int oracle_gate_lvl1(int mode, int kb, int op, const u32* in0, const u32* in1, const u32* inc,
                     const u32* bk, const u32* ksk, u32* out)
{
    return gate_lvl1_impl(mode, kb, op, in0, in1, inc, bk, ksk, out, nullptr);
}

// lvl02 gate API (BR-IKS order): lvl0 TLWEs in and out, BR at lvl2.
int oracle_gate_lvl02(int mode, int kb, int op, const u32* in0, const u32* in1, const u32* inc,
                      const u64* bk, const u32* ksk, u32* out)
{
    constexpr u32 N = L2::N;
    constexpr u64 mu = L2::mu;
    constexpr u32 mu0 = 1u << 29;
    std::vector<u32> tl0(N0 + 1);
    std::vector<u64> acc(2 * N), r1(N + 1), r0(N + 1);
    if (op == 12 || op == 13) {
        for (u32 i = 0; i <= N0; i++) out[i] = op == 12 ? 0u - in0[i] : in0[i];
        return 0;
    }
    auto pre = [&](const u32* a, const u32* b, i32 sa, i32 sb, u32 off) {
        for (u32 i = 0; i <= N0; i++) {
            u32 v = static_cast<u32>(sa) * a[i] + static_cast<u32>(sb) * b[i];
            if (i == N0) v += off;
            tl0[i] = v;
        }
    };
    if (op == 10 || op == 11) {
        pre(inc, in1, 1, 1, 0u - mu0);
        blind_rotate<L2>(mode, kb & 3, tl0.data(), bk, mu, acc.data());
        sample_extract<u64>(acc.data(), N, r1.data());
        pre(inc, in0, -1, 1, 0u - mu0);
        blind_rotate<L2>(mode, kb & 3, tl0.data(), bk, mu, acc.data());
        sample_extract<u64>(acc.data(), N, r0.data());
        const i32 s = op == 10 ? 1 : -1;
        identity_key_switch<u64>(N, false, r1.data(), r0.data(), s, s, s > 0 ? mu : 0 - mu, ksk,
                                 out);
        return 0;
    }
    GateLin g;
    if (!gate_lin(op, g)) return -1;
    pre(in0, in1, g.ca, g.cb, static_cast<u32>(g.off) * mu0);
    blind_rotate<L2>(mode, kb & 3, tl0.data(), bk, mu, acc.data());
    sample_extract<u64>(acc.data(), N, r1.data());
    identity_key_switch<u64>(N, false, r1.data(), nullptr, 1, 0, 0, ksk, out);
    return 0;
}

// Batch of lvl1 gates on nthreads host threads (one gate per thread):
// the CPU baseline of the benchmark and the checker of full-size batches.
// Ciphertexts are rows of 1025 words; inc may be null when no gate is a MUX.
// The NTT-mode key spectra are computed once for the whole batch.
int oracle_gate_batch_lvl1_ex(int mode, int kb, const int32_t* ops, const u32* in0, const u32* in1,
                              const u32* inc, const u32* bk, const u32* ksk, u32* out, int n,
                              int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if ((kb & 3) == BR_BB && mode == ORACLE_FFT) return -1;
    std::vector<u64> kspec;
    if (mode == ORACLE_NTT) kspec = key_spectra_ntt<L1>(bk, bk_trgsws(kb & 3));
    const std::vector<u64>* ks = mode == ORACLE_NTT ? &kspec : nullptr;
    std::vector<std::thread> pool;
    std::vector<int> rc(nthreads, 0);
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([&, t]() {
            for (int g = t; g < n; g += nthreads) {
                const size_t o = static_cast<size_t>(g) * (L1::N + 1);
                if (gate_lvl1_impl(mode, kb, ops[g], in0 + o, in1 + o, inc ? inc + o : nullptr, bk,
                                   ksk, out + o, ks))
                    rc[t] = -1;
            }
        });
    for (auto& th : pool) th.join();
    for (int r : rc)
        if (r) return r;
    return 0;
}

int oracle_gate_batch_lvl1(int mode, int kb, const int32_t* ops, const u32* in0, const u32* in1,
                           const u32* bk, const u32* ksk, u32* out, int n, int nthreads)
{
    return oracle_gate_batch_lvl1_ex(mode, kb, ops, in0, in1, nullptr, bk, ksk, out, n, nthreads);
}

// Test hook: plant one of the Mutation mistakes (0 restores the faithful
// oracle). Used only by the tests that show the pinning checks catch them.
int oracle_set_mutation(int m)
{
    if (m < 0 || m > MUT_AH_KEY_ROW) return -1;
    g_mut = m;
    return 0;
}

}  // extern "C"
