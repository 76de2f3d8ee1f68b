// CPU BASELINE ENGINE for lvl1 gates -- TEST INFRASTRUCTURE.
//
// The CPU column of the benchmark: lvl1 gate bootstrapping (identity key
// switch, blind rotation with the FP64 FFNT, sample extraction) the way a
// CPU library does it (TFHEpp keeps its bootstrapping key in the Fourier
// domain): the key is transformed ONCE per key load, each rotation step
// transforms only the accumulator's gadget digits (4 forward, 2 inverse
// transforms per step at lvl1) and forms the key-bundle combination
//   one + BK2 X(a1) + BK1 X(a0) + BK0 X(a0 + a1)
// bin by bin from the stored spectra and a table of X^a - 1 spectra
// (reference include/gatebootstrapping_gpu.cuh:895-1033; plain CMUX :69-221).
// It is NOT TFHEpp (TFHEpp is absent from the reference tree): it is this
// project's own CPU implementation of the same arithmetic, built with
// AVX2/FMA code generation and run one gate per host thread. It is timed by
// bench.py's cpu_baseline leg and checked against the oracle's FFT mode by
// tests/test_cpu_baseline.py; nothing in the product library calls it.
//
// The key switch is the oracle's formula (keyswitch_gpu.cuh:189-256) with
// the loops interchanged so each digit adds one whole KSK row to the 637
// output words (vectorised); the test vector, mod switch and gadget digits
// follow gatebootstrapping_gpu.cuh:10-53, :123-129.
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

using u32 = uint32_t;
using i32 = int32_t;

constexpr u32 N = 1024, H = 512, NBIT = 10, L = 2, BGBIT = 8, N0 = 636;
constexpr u32 ROWS = 2 * L;       // digit rows of a TRGSW
constexpr u32 POLYS = 2 * ROWS;   // key polynomials of a TRGSW
constexpr u32 MU = 1u << 29;

// Split-array H-point FFT: forward is decimation in frequency (natural
// order in, bit-reversed order out), inverse decimation in time (bit-reversed
// in, natural out). Pointwise products never look at the bin order, and the
// key, the X^a - 1 table and the digits all use this one forward transform.
struct Fft {
    std::vector<double> wr, wi;  // e^{-2 pi i k / H}, k < H/2
    std::vector<double> tr, ti;  // twist e^{i pi k / N}, k < H
    Fft()
    {
        wr.resize(H / 2);
        wi.resize(H / 2);
        for (u32 k = 0; k < H / 2; k++) {
            wr[k] = std::cos(-2.0 * M_PI * k / H);
            wi[k] = std::sin(-2.0 * M_PI * k / H);
        }
        tr.resize(H);
        ti.resize(H);
        for (u32 k = 0; k < H; k++) {
            tr[k] = std::cos(M_PI * k / N);
            ti[k] = std::sin(M_PI * k / N);
        }
    }
    void forward(double* __restrict re, double* __restrict im) const
    {
        for (u32 len = H; len >= 2; len >>= 1) {
            const u32 half = len / 2, stride = H / len;
            for (u32 s = 0; s < H; s += len) {
                double* ar = re + s;
                double* ai = im + s;
                for (u32 k = 0; k < half; k++) {
                    const double cr = wr[k * stride], ci = wi[k * stride];
                    const double ur = ar[k], ui = ai[k], vr = ar[k + half], vi = ai[k + half];
                    ar[k] = ur + vr;
                    ai[k] = ui + vi;
                    const double dr = ur - vr, di = ui - vi;
                    ar[k + half] = dr * cr - di * ci;
                    ai[k + half] = dr * ci + di * cr;
                }
            }
        }
    }
    void inverse(double* __restrict re, double* __restrict im) const
    {
        for (u32 len = 2; len <= H; len <<= 1) {
            const u32 half = len / 2, stride = H / len;
            for (u32 s = 0; s < H; s += len) {
                double* ar = re + s;
                double* ai = im + s;
                for (u32 k = 0; k < half; k++) {
                    const double cr = wr[k * stride], ci = -wi[k * stride];
                    const double xr = ar[k + half], xi = ai[k + half];
                    const double vr = xr * cr - xi * ci, vi = xr * ci + xi * cr;
                    const double ur = ar[k], ui = ai[k];
                    ar[k] = ur + vr;
                    ai[k] = ui + vi;
                    ar[k + half] = ur - vr;
                    ai[k + half] = ui - vi;
                }
            }
        }
    }
    // real polynomial p (N values) -> spectrum: fold (re = p[i], im = p[i+H]), twist, FFT
    void poly_forward(const double* p, double* re, double* im) const
    {
        for (u32 i = 0; i < H; i++) {
            const double a = p[i], b = p[i + H];
            re[i] = a * tr[i] - b * ti[i];
            im[i] = a * ti[i] + b * tr[i];
        }
        forward(re, im);
    }
    // spectrum -> real polynomial scaled by `scale` (IFFT, untwist, 1/H)
    void poly_inverse(double* re, double* im, double scale, double* p) const
    {
        inverse(re, im);
        const double s = scale / H;
        for (u32 i = 0; i < H; i++) {
            const double a = re[i], b = im[i];
            p[i] = (a * tr[i] + b * ti[i]) * s;
            p[i + H] = (b * tr[i] - a * ti[i]) * s;
        }
    }
};

const Fft& fft()
{
    static const Fft f;
    return f;
}

// Spectra of the key: [trgsw][poly][re H | im H]; and of X^a - 1 for a < 2N.
struct Spectra {
    int variant = 0;  // 0 plain, 1 key bundle
    size_t ntrgsw = 0;
    std::vector<double> key;
    std::vector<double> xai;  // [a][re H | im H]
    const double* poly(size_t t, u32 p) const { return key.data() + (t * POLYS + p) * 2 * H; }
    const double* x(u32 a) const { return xai.data() + static_cast<size_t>(a) * 2 * H; }
};

void build_spectra(Spectra& S, const u32* bk, int variant)
{
    const Fft& F = fft();
    S.variant = variant;
    S.ntrgsw = variant == 1 ? 3 * (N0 / 2) : N0;
    S.key.assign(S.ntrgsw * POLYS * 2 * H, 0.0);
    std::vector<double> tmp(N);
    const double s = 1.0 / 4294967296.0;
    for (size_t t = 0; t < S.ntrgsw; t++)
        for (u32 p = 0; p < POLYS; p++) {
            const u32* c = bk + (t * POLYS + p) * N;
            for (u32 i = 0; i < N; i++) tmp[i] = static_cast<double>(static_cast<i32>(c[i])) * s;
            double* out = S.key.data() + (t * POLYS + p) * 2 * H;
            F.poly_forward(tmp.data(), out, out + H);
        }
    S.xai.assign(static_cast<size_t>(2 * N) * 2 * H, 0.0);
    for (u32 a = 0; a < 2 * N; a++) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        tmp[0] = -1.0;
        if (a < N)
            tmp[a] += 1.0;
        else
            tmp[a - N] -= 1.0;
        double* out = S.xai.data() + static_cast<size_t>(a) * 2 * H;
        F.poly_forward(tmp.data(), out, out + H);
    }
}

inline u32 mod_switch(u32 v) { return v >> (32 - 1 - NBIT); }

inline i32 digit(u32 c, u32 q)
{
    constexpr u32 off = (1u << (BGBIT - 1)) * ((1u << (32 - BGBIT)) + (1u << (32 - 2 * BGBIT))) +
                        (1u << (32 - L * BGBIT - 1));
    const u32 t = c + off;
    return static_cast<i32>((t >> (32 - (q + 1) * BGBIT)) & ((1u << BGBIT) - 1)) -
           static_cast<i32>(1u << (BGBIT - 1));
}

// Per-thread scratch of one rotation.
struct Work {
    std::vector<double> d[ROWS][2];  // digit spectra, re/im
    std::vector<double> k[POLYS][2];  // combined key spectra (key bundle)
    std::vector<double> o[2][2];     // output spectra
    std::vector<double> poly;
    std::vector<u32> rot;
    Work()
    {
        for (auto& r : d)
            for (auto& v : r) v.resize(H);
        for (auto& r : k)
            for (auto& v : r) v.resize(H);
        for (auto& r : o)
            for (auto& v : r) v.resize(H);
        poly.resize(N);
        rot.resize(2 * N);
    }
};

// acc (+)= ExtProd(src, K) with K[p] = key spectra of poly p.
template <class KeyFn>
void ext_product(const u32* src, KeyFn K, Work& w, u32* acc, bool add)
{
    const Fft& F = fft();
    for (u32 r = 0; r < ROWS; r++) {
        const u32* c = src + (r / L) * N;
        for (u32 i = 0; i < N; i++) w.poly[i] = static_cast<double>(digit(c[i], r % L));
        F.poly_forward(w.poly.data(), w.d[r][0].data(), w.d[r][1].data());
    }
    for (u32 j = 0; j < 2; j++) {
        double* __restrict orr = w.o[j][0].data();
        double* __restrict oi = w.o[j][1].data();
        std::fill(orr, orr + H, 0.0);
        std::fill(oi, oi + H, 0.0);
        for (u32 r = 0; r < ROWS; r++) {
            const double* __restrict dr = w.d[r][0].data();
            const double* __restrict di = w.d[r][1].data();
            const double* kr;
            const double* ki;
            K(r * 2 + j, kr, ki);
            for (u32 b = 0; b < H; b++) {
                orr[b] += dr[b] * kr[b] - di[b] * ki[b];
                oi[b] += dr[b] * ki[b] + di[b] * kr[b];
            }
        }
        F.poly_inverse(w.o[j][0].data(), w.o[j][1].data(), 4294967296.0, w.poly.data());
        u32* a = acc + j * N;
        for (u32 i = 0; i < N; i++) {
            const u32 v = static_cast<u32>(static_cast<int64_t>(std::nearbyint(w.poly[i])));
            a[i] = add ? a[i] + v : v;
        }
    }
}

void blind_rotate(const Spectra& S, const u32* tl0, u32* acc, Work& w)
{
    const u32 bar = 2 * N - mod_switch(tl0[N0]);
    for (u32 i = 0; i < N; i++) {
        acc[i] = 0;
        if (bar == 2 * N)
            acc[N + i] = MU;
        else
            acc[N + i] = ((i < (bar & (N - 1))) ^ ((bar >> NBIT) != 0)) ? 0u - MU : MU;
    }
    constexpr u32 rnd = 1u << (32 - 2 - NBIT);
    if (S.variant == 1) {
        for (u32 it = 0; it < N0 / 2; it++) {
            const u32 a0 = mod_switch(tl0[2 * it] + rnd), a1 = mod_switch(tl0[2 * it + 1] + rnd);
            const u32 a01 = (a0 + a1) & (2 * N - 1);
            const double *x0 = S.x(a0), *x1 = S.x(a1), *x01 = S.x(a01);
            for (u32 p = 0; p < POLYS; p++) {
                const u32 r = p / 2, j = p % 2;
                const double one = j == r / L ? std::ldexp(1.0, -static_cast<int>((r % L + 1) * BGBIT)) : 0.0;
                const double* b0 = S.poly(3 * it + 0, p);
                const double* b1 = S.poly(3 * it + 1, p);
                const double* b2 = S.poly(3 * it + 2, p);
                double* __restrict kr = w.k[p][0].data();
                double* __restrict ki = w.k[p][1].data();
                for (u32 b = 0; b < H; b++) {
                    double vr = one, vi = 0.0;
                    vr += b2[b] * x1[b] - b2[b + H] * x1[b + H];
                    vi += b2[b] * x1[b + H] + b2[b + H] * x1[b];
                    vr += b1[b] * x0[b] - b1[b + H] * x0[b + H];
                    vi += b1[b] * x0[b + H] + b1[b + H] * x0[b];
                    vr += b0[b] * x01[b] - b0[b + H] * x01[b + H];
                    vi += b0[b] * x01[b + H] + b0[b + H] * x01[b];
                    kr[b] = vr;
                    ki[b] = vi;
                }
            }
            ext_product(
                acc,
                [&](u32 p, const double*& kr, const double*& ki) {
                    kr = w.k[p][0].data();
                    ki = w.k[p][1].data();
                },
                w, acc, false);
        }
        return;
    }
    for (u32 it = 0; it < N0; it++) {
        const u32 a = mod_switch(tl0[it] + rnd);
        // d = X^a acc - acc
        for (u32 j = 0; j < 2; j++)
            for (u32 i = 0; i < N; i++) {
                const u32 s = (i - a) & (N - 1);
                const bool neg = (i < (a & (N - 1))) ^ ((a / N) != 0);
                const u32 v = neg ? 0u - acc[j * N + s] : acc[j * N + s];
                w.rot[j * N + i] = v - acc[j * N + i];
            }
        ext_product(
            w.rot.data(),
            [&](u32 p, const double*& kr, const double*& ki) {
                kr = S.poly(it, p);
                ki = kr + H;
            },
            w, acc, true);
    }
}

// lvl1 -> lvl0 identity key switch of ca*in0 + cb*in1 (+offset on the
// body), t = 7 digits of 2 bits: row-wise over the TFHEpp KSK layout
// [j][k][|v| - 1][n0 + 1].
void key_switch(const u32* in0, const u32* in1, i32 ca, i32 cb, u32 offset, const u32* ksk, u32* out)
{
    constexpr u32 T = 7, BASEBIT = 2, NUMBASE = 2, W0 = N0 + 1;
    constexpr u32 decomp = (2u << 30) + (2u << 28) + (2u << 26) + (2u << 24) + (2u << 22) +
                           (2u << 20) + (2u << 18);
    constexpr u32 round = 1u << (32 - (1 + BASEBIT * T));
    const u32 ua = static_cast<u32>(ca), ub = static_cast<u32>(cb);
    for (u32 i = 0; i < N0; i++) out[i] = 0;
    out[N0] = ua * in0[N] + ub * in1[N] + offset;
    for (u32 j = 0; j < N; j++) {
        const u32 tmp = ua * in0[j] + ub * in1[j] + decomp + round;
        for (u32 k = 0; k < T; k++) {
            const i32 val = static_cast<i32>((tmp >> (32 - (k + 1) * BASEBIT)) & 3u) - static_cast<i32>(NUMBASE);
            if (val == 0) continue;
            const u32* __restrict row = ksk + ((static_cast<size_t>(j) * T + k) * NUMBASE + (std::abs(val) - 1)) * W0;
            u32* __restrict o = out;
            if (val > 0)
                for (u32 i = 0; i < W0; i++) o[i] -= row[i];
            else
                for (u32 i = 0; i < W0; i++) o[i] += row[i];
        }
    }
}

struct GateLin {
    i32 ca, cb, off;
};
// (ca, cb, offset in units of mu): reference src/bootstrap_gpu.cu:1335-1881, :1905-1946
bool gate_lin(int op, GateLin& g)
{
    static const GateLin t[10] = {{-1, -1, 1}, {1, 1, 1},   {1, -1, 1},  {-1, 1, 1},
                                  {1, 1, -1},  {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1},
                                  {2, 2, 2},   {-2, -2, -2}};
    if (op < 0 || op > 9) return false;
    g = t[op];
    return true;
}

int gate(const Spectra& S, int op, const u32* in0, const u32* in1, const u32* ksk, u32* out, Work& w)
{
    GateLin g;
    if (!gate_lin(op, g)) return -1;
    std::vector<u32> tl0(N0 + 1), acc(2 * N);
    key_switch(in0, in1, g.ca, g.cb, static_cast<u32>(g.off) * MU, ksk, tl0.data());
    blind_rotate(S, tl0.data(), acc.data(), w);
    out[0] = acc[0];
    for (u32 i = 1; i < N; i++) out[i] = 0u - acc[N - i];
    out[N] = acc[N];
    return 0;
}

}  // namespace

extern "C" {

// Key spectra for the baseline engine (variant 0 plain, 1 key bundle);
// returns an opaque handle or null.
void* oracle_cpu_baseline_prepare(const uint32_t* bk, int variant)
{
    if (variant != 0 && variant != 1) return nullptr;
    auto* S = new Spectra;
    build_spectra(*S, bk, variant);
    return S;
}

void oracle_cpu_baseline_release(void* h) { delete static_cast<Spectra*>(h); }

// n binary lvl1 gates (ops 0-9), rows of 1025 words, on nthreads threads
// (one gate per thread at a time).
int oracle_cpu_baseline_gates(void* h, const int32_t* ops, const uint32_t* in0, const uint32_t* in1,
                              const uint32_t* ksk, uint32_t* out, int n, int nthreads)
{
    if (!h || n < 0) return -1;
    if (nthreads < 1) nthreads = 1;
    const Spectra& S = *static_cast<const Spectra*>(h);
    std::vector<std::thread> pool;
    std::vector<int> rc(nthreads, 0);
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([&, t]() {
            Work w;
            for (int g = t; g < n; g += nthreads) {
                const size_t o = static_cast<size_t>(g) * (N + 1);
                if (gate(S, ops[g], in0 + o, in1 + o, ksk, out + o, w)) rc[t] = -1;
            }
        });
    for (auto& th : pool) th.join();
    for (int r : rc)
        if (r) return r;
    return 0;
}

}  // extern "C"
