// lvl2 (N = 2048, Torus64) NTT backend over the Goldilocks prime
// P = 2^64 - 2^32 + 1 (reference include/ntt_small_modulus.cuh:60-74
// constants, :154-193 arithmetic, :284-361 torus <-> Z_P conversions;
// src/ntt_small_modulus.cu:85-113 the lvl2 root, src/bootstrap_gpu.cu:213-340
// the key transform).
//
// Arithmetic. Residues are canonical in [0, P). A product is reduced with the
// identities 2^64 = 2^32 - 1 and 2^96 = -1 (mod P): for x = hi 2^64 + lo,
// hi = h1 2^32 + h0, x = lo - h1 + h0 (2^32 - 1) (mod P) -- the published
// Goldilocks reduction (Plonky2's reduce128) -- followed by one conditional
// subtraction.
//
// Transform (one wavefront per 2048-point negacyclic NTT, 32 values per
// lane). Forward: Cooley-Tukey with the merged psi twist (psi of order 2N,
// table psi^brv(i)), natural order in, output position p holding the
// evaluation at psi^(2 brv(p) + 1); inverse: Gentleman-Sande with psi^-brv(i)
// and N^-1, positions in, natural order out (Longa and Naehrig, "Speeding up
// the Number Theoretic Transform for Faster Ideal Lattice-Based
// Cryptography", 2016, Algorithms 1 and 2). Modular arithmetic is exact, so
// pointwise products of forward transforms followed by the inverse give
// exactly the negacyclic convolution mod P whatever the bin order or root.
//   * stages with t >= 64 run in registers with lane l holding elements
//     l + 64 r (the twiddle is wave-uniform);
//   * the t = 32 stage runs through the wave's LDS buffer;
//   * stages with t <= 16 run in registers with lane l holding elements
//     32 l + r (per-lane twiddles).
// LDS element e sits at e + e / 32 (one pad per 32), so both layouts' 64-bit
// accesses are bank-conflict-minimal.
#pragma once

#include "br_fft.hpp"

namespace cufhe_amd {

namespace gold {
constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 - P
constexpr uint64_t HALF_P = P / 2;
// Primitive 2^32-th root of unity (reference ntt_small_modulus.cuh:67).
constexpr uint64_t ROOT_2_32 = 12037493425763644479ull;

// Branch-free forms (selects, no divergent branches): canonical in, canonical out.
__host__ __device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    s += s < a ? EPS : 0;  // wrapped: a + b - 2^64 + EPS = a + b - P < P
    return s >= P ? s - P : s;
}
__host__ __device__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b)
{
    const uint64_t d = a - b;
    return d - (a < b ? EPS : 0);  // a - b + 2^64 - EPS = a - b + P
}
__host__ __device__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi)
{
    const uint64_t hh = hi >> 32, hl = hi & EPS;
    uint64_t t0 = lo - hh;
    t0 -= lo < hh ? EPS : 0;
    const uint64_t t1 = (hl << 32) - hl;  // hl * EPS
    uint64_t r = t0 + t1;
    r += r < t1 ? EPS : 0;  // carry: cannot overflow again
    return r >= P ? r - P : r;
}
__device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b)
{
    return reduce128(a * b, __umul64hi(a, b));
}
}  // namespace gold

// torus_to_ntt_mod for Torus64: floor((a P + 2^63) / 2^64) (reference
// ntt_small_modulus.cuh:284-295); < P for every a.
__host__ __device__ __forceinline__ uint64_t torus64_to_gold(uint64_t a)
{
    const unsigned __int128 prod = static_cast<unsigned __int128>(a) * gold::P +
                                   (static_cast<unsigned __int128>(1) << 63);
    return static_cast<uint64_t>(prod >> 64);
}
// ntt_mod_to_torus64, Goldilocks branch (reference :308-315, :320-335).
__host__ __device__ __forceinline__ uint64_t gold_abs_to_torus64(uint64_t v)
{
    const unsigned __int128 m = v;
    return static_cast<uint64_t>(((m << 64) + (m << 32) - m + (static_cast<unsigned __int128>(1) << 63)) >>
                                 64);
}
__host__ __device__ __forceinline__ uint64_t gold_to_torus64(uint64_t v)
{
    return v > gold::HALF_P ? 0 - gold_abs_to_torus64(gold::P - v) : gold_abs_to_torus64(v);
}
// signed_int_to_ntt_mod (reference :297-306)
__device__ __forceinline__ uint64_t signed_to_gold(int32_t v)
{
    return v < 0 ? gold::P - static_cast<uint64_t>(-static_cast<int64_t>(v)) : static_cast<uint64_t>(v);
}

// A global-memory table pointer whose provenance the compiler forgets at each
// use site, so table loads are not hoisted out of the rotation's loops (the
// twiddles are re-read from L1 rather than held in VGPRs); the address space
// stays global so the loads are not flat.
using gtab64 = const __attribute__((address_space(1))) uint64_t*;
__device__ __forceinline__ gtab64 opaque_table(const uint64_t* p)
{
    gtab64 g = (gtab64)p;
    asm volatile("" : "+s"(g));
    return g;
}

struct Ntt64Tables {
    const uint64_t* psi_rev;      // [N] psi^brv(i)
    const uint64_t* psi_inv_rev;  // [N] psi^-brv(i)
    const uint64_t* pw;           // [2N] psi^e
    uint64_t n_inv;
};

struct Ntt64Keys {
    // [trgsw][row][out][N]: forward transforms of torus64_to_gold(BK); word
    // 64 r + l of a polynomial holds transform position 32 l + r
    const uint64_t* bk;
    Ntt64Tables tb;
    uint64_t one[lvl2param::l];  // torus64_to_gold(h_q), the key bundle's identity term
};

constexpr int kNtt64N = lvl2param::n;
constexpr int kNtt64Buf = kNtt64N + kNtt64N / 32;  // padded LDS buffer, u64
constexpr size_t kTrgswNtt64 =
    static_cast<size_t>(TFHEpp::BootstrappingTRGSWRows<lvl2param>) * (lvl2param::k + 1) * kNtt64N;

__device__ __forceinline__ int ntt64_addr(int e) { return e + (e >> 5); }

// Forward transform of x (lane l holds elements l + 64 r on entry); on exit
// lane l holds output positions 32 l + r. buf: the wave's LDS scratch.
__device__ __forceinline__ void ntt64_forward(uint64_t (&x)[32], uint64_t* buf, const Ntt64Tables& tb,
                                              int lane)
{
    const gtab64 psi = opaque_table(tb.psi_rev);
    // t = 1024 .. 64 (m = 1 .. 16): pairs (r0, r0 + tr), tr = t / 64
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int m = 1 << s, tr = 16 >> s;
#pragma unroll
        for (int r0 = 0; r0 < 32; r0++) {
            if ((r0 / tr) % 2 == 0) {
                const uint64_t S = psi[m + r0 / (2 * tr)];
                const uint64_t U = x[r0], V = gold::mul(x[r0 + tr], S);
                x[r0] = gold::add(U, V);
                x[r0 + tr] = gold::sub(U, V);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 32; r++) buf[ntt64_addr(lane + 64 * r)] = x[r];
    wave_lds_sync();
    // t = 32 (m = 32): butterfly k = lane + 64 b -> block i = 2 b + h, offset l
    {
        const int h = lane >> 5, l = lane & 31;
#pragma unroll
        for (int b = 0; b < 16; b++) {
            const int p0 = 128 * b + 64 * h + l;
            const uint64_t S = psi[32 + 2 * b + h];
            const uint64_t U = buf[ntt64_addr(p0)], V = gold::mul(buf[ntt64_addr(p0 + 32)], S);
            buf[ntt64_addr(p0)] = gold::add(U, V);
            buf[ntt64_addr(p0 + 32)] = gold::sub(U, V);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = buf[ntt64_addr(32 * lane + r)];
    wave_lds_sync();
    // t = 16 .. 1 (m = 64 .. 1024): element 32 lane + r0, block i = lane (16 / t) + r0 / (2 t)
    // (one stage's twiddles in flight at a time)
#pragma unroll
    for (int s = 0; s < 5; s++) {
        __builtin_amdgcn_sched_barrier(0);
        const int t = 16 >> s, m = 64 << s;
#pragma unroll
        for (int r0 = 0; r0 < 32; r0++) {
            if ((r0 / t) % 2 == 0) {
                const uint64_t S = psi[m + lane * (16 / t) + r0 / (2 * t)];
                const uint64_t U = x[r0], V = gold::mul(x[r0 + t], S);
                x[r0] = gold::add(U, V);
                x[r0 + t] = gold::sub(U, V);
            }
        }
    }
}

// Inverse transform: lane l holds positions 32 l + r on entry, elements
// l + 64 r (natural order, scaled by N^-1) on exit.
__device__ __forceinline__ void ntt64_inverse(uint64_t (&x)[32], uint64_t* buf, const Ntt64Tables& tb,
                                              int lane)
{
    const gtab64 psi = opaque_table(tb.psi_inv_rev);
    // t = 1 .. 16 (h = 1024 .. 64): element 32 lane + r0, block i = lane (16 / t) + r0 / (2 t)
#pragma unroll
    for (int s = 0; s < 5; s++) {
        __builtin_amdgcn_sched_barrier(0);
        const int t = 1 << s, hh = 1024 >> s;
#pragma unroll
        for (int r0 = 0; r0 < 32; r0++) {
            if ((r0 / t) % 2 == 0) {
                const uint64_t S = psi[hh + lane * (16 / t) + r0 / (2 * t)];
                const uint64_t U = x[r0], V = x[r0 + t];
                x[r0] = gold::add(U, V);
                x[r0 + t] = gold::mul(gold::sub(U, V), S);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 32; r++) buf[ntt64_addr(32 * lane + r)] = x[r];
    wave_lds_sync();
    // t = 32 (h = 32)
    {
        const int h = lane >> 5, l = lane & 31;
#pragma unroll
        for (int b = 0; b < 16; b++) {
            const int p0 = 128 * b + 64 * h + l;
            const uint64_t S = psi[32 + 2 * b + h];
            const uint64_t U = buf[ntt64_addr(p0)], V = buf[ntt64_addr(p0 + 32)];
            buf[ntt64_addr(p0)] = gold::add(U, V);
            buf[ntt64_addr(p0 + 32)] = gold::mul(gold::sub(U, V), S);
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = buf[ntt64_addr(lane + 64 * r)];
    wave_lds_sync();
    // t = 64 .. 1024 (h = 16 .. 1): pairs (r0, r0 + tr), block i = r0 / (2 tr)
#pragma unroll
    for (int s = 0; s < 5; s++) {
        const int tr = 1 << s, hh = 16 >> s;
#pragma unroll
        for (int r0 = 0; r0 < 32; r0++) {
            if ((r0 / tr) % 2 == 0) {
                const uint64_t S = psi[hh + r0 / (2 * tr)];
                const uint64_t U = x[r0], V = x[r0 + tr];
                x[r0] = gold::add(U, V);
                x[r0 + tr] = gold::mul(gold::sub(U, V), S);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = gold::mul(x[r], tb.n_inv);
}

// One blind-rotation step at lvl2 on the NTT backend for the wave owning
// component j (two waves per job), mirroring br_step_fft / br_step_ntt:
//   KB   : acc  = ExtProd(acc, one + BK2 X(a1) + BK1 X(a0) + BK0 X(a0+a1))
//   plain: acc += ExtProd((X^a0 - 1) acc, BK)
// (reference include/gatebootstrapping_gpu.cuh:522-659, :1187-1332). Per
// gadget digit q: the wave transforms its digit polynomial, publishes the
// spectrum in its LDS buffer, the pair synchronises, each wave multiplies
// both components' spectra by the key column of its own output component,
// and the pair synchronises again before the buffers are reused. X(a) at
// position p is psi^(a (2 brv(p) + 1)) - 1.
// GHOST: the same loop nest and workgroup barriers with no loads, stores or
// arithmetic, for waves that keep step with a rotation they take no part in.
template <bool KB, bool GHOST = false>
__device__ __forceinline__ void br_step_ntt64(AccStore<lvl2param>& acc, int j, uint32_t a0, uint32_t a1,
                                              const uint64_t* trgsw0, uint64_t* my_buf,
                                              const uint64_t* partner_buf, const Ntt64Keys& K)
{
    using P = lvl2param;
    constexpr int N = P::n, L = P::l;
    const int lane = threadIdx.x & 63;
    uint64_t sp[32];
#pragma unroll
    for (int r = 0; r < 32; r++) sp[r] = 0;
    const uint32_t a01 = (a0 + a1) & (2 * N - 1);
    // position p = 32 lane + r is evaluation point psi^(2 brv11(p) + 1) and
    // brv11(p) = 64 brv5(r) + brv6(lane): X^a at p is pw[(a eb + 128 a brv5(r)) mod 2N]
    // with eb = 2 brv6(lane) + 1 per lane and the second term wave-uniform
    const uint32_t eb = 2u * (__brev(static_cast<uint32_t>(lane)) >> 26) + 1u;
    // this lane's 32 spectrum words in the padded LDS layout: 33 lane + r
    const uint64_t* const own = my_buf + 33 * lane;
    const uint64_t* const oth = partner_buf + 33 * lane;
#pragma unroll 1
    for (int q = 0; q < L; q++) {
        if constexpr (!GHOST) {
            uint64_t x[32];
            // groups of eight coefficients: bounds the LDS loads in flight
            static_for<4>([&](auto gc) {
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 8 * decltype(gc)::value; r < 8 * decltype(gc)::value + 8; r++) {
                    const int e = lane + 64 * r;
                    uint64_t src;
                    if constexpr (KB)
                        src = acc.lds[e];
                    else
                        src = rotated_coef<P>(acc.lds, e, a0) - acc.lds[e];
                    x[r] = signed_to_gold(Gadget<P>::digit(src, q));
                }
            });
            __builtin_amdgcn_sched_barrier(0);
            ntt64_forward(x, my_buf, K.tb, lane);
#pragma unroll
            for (int r = 0; r < 32; r++) my_buf[ntt64_addr(32 * lane + r)] = x[r];
        }
        // key column j of row (j, q) for this wave's own digit spectrum and
        // of row (1 - j, q) for the partner's; word 64 r + lane holds position p
        const uint64_t* const kown =
            GHOST ? nullptr : trgsw0 + static_cast<size_t>((j * L + q) * 2 + j) * N;
        const uint64_t* const koth =
            GHOST ? nullptr : trgsw0 + static_cast<size_t>(((1 - j) * L + q) * 2 + j) * N;
        // the first group's key words are loaded while the pair meets at the
        // barrier (their L2 latency overlaps the wait)
        constexpr int KM = KB ? 3 : 1;
        uint64_t kpre[4][2][KM];
        if constexpr (!GHOST) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int e = 0; e < KM; e++) {
                    kpre[r][0][e] = kown[e * kTrgswNtt64 + 64 * r + lane];
                    kpre[r][1][e] = koth[e * kTrgswNtt64 + 64 * r + lane];
                }
        }
        __syncthreads();  // both spectra of digit q published
        if constexpr (!GHOST) {
            uint32_t b0 = 0, b1 = 0, b01 = 0;
            uint64_t one = 0;
            if constexpr (KB) {
                // per-lane parts of the X^a exponents, recomputed per digit so
                // they are not held across the digit loop
                b0 = a0 * eb;
                b1 = a1 * eb;
                b01 = a01 * eb;
                asm volatile("" : "+v"(b0), "+v"(b1), "+v"(b01));
                one = K.one[q];
            }
            // groups of four bins: bounds the loads the scheduler keeps in flight
            static_for<8>([&](auto gc) {
                __builtin_amdgcn_sched_barrier(0);
                static_for<4>([&](auto rc) {
                    constexpr int r = 4 * decltype(gc)::value + decltype(rc)::value;
                    constexpr uint32_t br5 = __builtin_bitreverse32(static_cast<uint32_t>(r)) >> 27;
                    const uint64_t d_own = own[r], d_oth = oth[r];
                    // key word e of the own / partner row at this register
                    auto kw = [&](int side, int e) -> uint64_t {
                        if constexpr (r < 4)
                            return kpre[r][side][e];
                        else
                            return (side == 0 ? kown : koth)[e * kTrgswNtt64 + 64 * r + lane];
                    };
                    uint64_t k_own, k_oth;
                    if constexpr (KB) {
                        const uint64_t* const pw = K.tb.pw;
                        const uint64_t xa0 = gold::sub(pw[(b0 + 128u * br5 * a0) & (2 * N - 1)], 1);
                        const uint64_t xa1 = gold::sub(pw[(b1 + 128u * br5 * a1) & (2 * N - 1)], 1);
                        const uint64_t xa01 = gold::sub(pw[(b01 + 128u * br5 * a01) & (2 * N - 1)], 1);
                        // one + BK2 X(a1) + BK1 X(a0) + BK0 X(a0 + a1)
                        k_own = gold::add(one, gold::mul(kw(0, 2), xa1));
                        k_own = gold::add(k_own, gold::mul(kw(0, 1), xa0));
                        k_own = gold::add(k_own, gold::mul(kw(0, 0), xa01));
                        k_oth = gold::mul(kw(1, 2), xa1);
                        k_oth = gold::add(k_oth, gold::mul(kw(1, 1), xa0));
                        k_oth = gold::add(k_oth, gold::mul(kw(1, 0), xa01));
                    }
                    else {
                        k_own = kw(0, 0);
                        k_oth = kw(1, 0);
                    }
                    sp[r] = gold::add(sp[r], gold::add(gold::mul(d_own, k_own), gold::mul(d_oth, k_oth)));
                });
            });
        }
        __syncthreads();  // the partner has read this wave's spectrum
    }
    if constexpr (GHOST) return;
    ntt64_inverse(sp, my_buf, K.tb, lane);
#pragma unroll
    for (int r = 0; r < 32; r++) {
        const int e = lane + 64 * r;
        const uint64_t t = gold_to_torus64(sp[r]);
        acc.lds[e] = KB ? t : acc.lds[e] + t;
    }
    wave_lds_sync();
}

// Full rotation of the (pre-added) lvl0 TLWE tl0 for the wave pair at lvl2.
// GHOST: the loop nest and barriers only (blind_rotate_ntt64_ghost).
template <bool KB, bool GHOST = false>
__device__ __forceinline__ void blind_rotate_ntt64(const uint32_t* tl0, uint64_t mu, AccStore<lvl2param>& acc,
                                                   int j, uint64_t* my_buf, const uint64_t* partner_buf,
                                                   const Ntt64Keys& K)
{
    using P = lvl2param;
    constexpr int N = P::n, n0 = lvl0param::n;
    const int lane = threadIdx.x & 63;
    if constexpr (!GHOST) {
        const uint32_t bar = __builtin_amdgcn_readfirstlane(2 * N - mod_switch<P>(tl0[n0]));
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int i = lane + 64 * r;
            acc.lds[i] = j == 0 ? 0 : test_vector_coef<P>(i, bar, mu);
        }
        wave_lds_sync();
    }
    constexpr uint32_t rnd = kMaskRound<P>;
    auto abar = [&](int i) -> uint32_t {
        if constexpr (GHOST) return 0u;
        return __builtin_amdgcn_readfirstlane(mod_switch<P>(tl0[i] + rnd));
    };
    if constexpr (KB) {
#pragma unroll 1
        for (int it = 0; it < n0 / 2; it++) {
            const uint32_t a0 = abar(2 * it), a1 = abar(2 * it + 1);
            br_step_ntt64<true, GHOST>(acc, j, a0, a1,
                                       GHOST ? nullptr : K.bk + static_cast<size_t>(3 * it) * kTrgswNtt64,
                                       my_buf, partner_buf, K);
        }
    }
    else {
#pragma unroll 1
        for (int it = 0; it < n0; it++) {
            br_step_ntt64<false, GHOST>(acc, j, abar(it), 0,
                                        GHOST ? nullptr : K.bk + static_cast<size_t>(it) * kTrgswNtt64,
                                        my_buf, partner_buf, K);
        }
    }
}

template <bool KB>
__device__ __forceinline__ void blind_rotate_ntt64_ghost(const Ntt64Keys& K)
{
    AccStore<lvl2param> none;
    none.lds = nullptr;
    blind_rotate_ntt64<KB, true>(nullptr, 0, none, 0, nullptr, nullptr, K);
}

}  // namespace cufhe_amd
