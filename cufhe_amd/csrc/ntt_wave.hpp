// Wavefront-level negacyclic NTT over Z_P, P = 2147473409 (N = 1024).
//
// Reference: include/ntt_small_modulus.cuh:38-74 (the prime), :154-282
// (modular ops), :394-509 (SmallForwardNTT/SmallInverseNTT, CT/GS). The
// results of the NTT backend are exact residues, so only the transform's
// mathematics must match, not its butterfly order or roots: any primitive
// 2N-th root psi gives the same negacyclic products bit for bit.
//
// Arithmetic: Montgomery multiplication with R = 2^32 (P < 2^31), operands
// canonical in [0, P). Twiddles and keys are stored in Montgomery form
// (x R mod P), data in normal form, so mont(data, twiddle) stays normal.
//
// Transform (one wave per polynomial, like fft_wave.hpp): coefficient
// i = lane + 64 r (r < 16) in VGPRs; Z[k] = sum_i x_i psi^(i (2k+1)):
// twist psi^i = psi^lane * psi^(64 r) (lane factor merged into the step-2
// twiddle), 16-point DFT over registers, twiddle, 64-point DFT over lanes as
// 8 x 8 with two LDS transposes laid out so that both access patterns of
// each transpose hit 64 distinct banks with 4-byte elements. Output: lane
// L = c + 8q, register g*8+d holds Z[k], k = (q+8g) + 16(c+8d), stored at
// p = (g*8+d)*64 + L.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "fft_wave.hpp"  // wave_lds_sync, static_for

namespace cufhe_amd {
namespace ntt31 {

constexpr uint32_t P = 2147473409u;  // 1048571 * 2^11 + 1
constexpr uint32_t HALF_P = P / 2;

constexpr uint64_t powmod(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    b %= P;
    while (e) {
        if (e & 1) r = r * b % P;
        b = b * b % P;
        e >>= 1;
    }
    return r;
}
// -P^-1 mod 2^32 (Newton iteration)
constexpr uint32_t neg_inv()
{
    uint32_t x = 1;
    for (int i = 0; i < 5; i++) x *= 2u - P * x;
    return 0u - x;
}
constexpr uint32_t PINV = neg_inv();
constexpr uint32_t R1 = static_cast<uint32_t>((1ULL << 32) % P);               // R mod P
constexpr uint32_t R2 = static_cast<uint32_t>((uint64_t)R1 * R1 % P);           // R^2 mod P
constexpr uint32_t to_mont_c(uint32_t x) { return static_cast<uint32_t>(((uint64_t)x << 32) % P); }

// Primitive 2N-th root of unity (N = 1024): g^((P-1)/2048) for the smallest
// quadratic non-residue g.
constexpr uint32_t find_psi()
{
    uint64_t g = 2;
    while (powmod(g, (P - 1) / 2) != P - 1) g++;
    return static_cast<uint32_t>(powmod(g, (P - 1) / 2048));
}
constexpr uint32_t PSI = find_psi();
constexpr uint32_t PSI_INV = static_cast<uint32_t>(powmod(PSI, 2047));
static_assert(powmod(PSI, 1024) == P - 1, "psi must be a primitive 2048-th root");

// root_pow(e) = psi^e (normal form), e taken mod 2048
constexpr uint32_t psi_pow(int64_t e) { return static_cast<uint32_t>(powmod(PSI, ((e % 2048) + 2048) % 2048)); }

}  // namespace ntt31

__device__ __forceinline__ uint32_t mod_add(uint32_t a, uint32_t b)
{
    const uint32_t s = a + b;  // < 2P < 2^32
    return s >= ntt31::P ? s - ntt31::P : s;
}
__device__ __forceinline__ uint32_t mod_sub(uint32_t a, uint32_t b)
{
    const uint32_t d = a + ntt31::P - b;
    return d >= ntt31::P ? d - ntt31::P : d;
}
// r mod P for r < 2P: min(r, r - P) (r - P wraps above r when r < P). Written
// as an unsigned min so it stays two VALU ops; a compare on the 64-bit
// Montgomery sum (what `r >= P ? r - P : r` folds into) costs a 64-bit
// compare, a select and a VCC hazard wait.
__device__ __forceinline__ uint32_t reduce_2p(uint32_t r)
{
    return __builtin_elementwise_min(r, r - ntt31::P);
}
// a * b * R^-1 mod P, canonical (a, b < P)
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b)
{
    const uint64_t t = static_cast<uint64_t>(a) * b;
    const uint32_t m = static_cast<uint32_t>(t) * ntt31::PINV;
    const uint64_t u = t + static_cast<uint64_t>(m) * ntt31::P;  // < 2^63 + 2^63
    return reduce_2p(static_cast<uint32_t>(u >> 32));            // u >> 32 < 2P
}

// (a1 b1 + a2 b2) R^-1 mod P with one Montgomery reduction: the 64-bit sum
// is below 2 P^2, so t + m P < 2^64 and the result is below 2P before the
// final subtraction. Equal to mod_add(mont_mul(a1, b1), mont_mul(a2, b2))
// for canonical inputs (every value here is a residue in [0, P)).
__device__ __forceinline__ uint32_t mont_dot2(uint32_t a1, uint32_t b1, uint32_t a2, uint32_t b2)
{
    const uint64_t t = static_cast<uint64_t>(a1) * b1 + static_cast<uint64_t>(a2) * b2;
    const uint32_t m = static_cast<uint32_t>(t) * ntt31::PINV;
    const uint64_t u = t + static_cast<uint64_t>(m) * ntt31::P;
    return reduce_2p(static_cast<uint32_t>(u >> 32));
}

// Multiply by the compile-time twiddle W_M^E (forward: omega_M^E, inverse:
// omega_M^-E; omega_M = psi^(2048/M)), data in normal form.
template <int M, int E, bool INV>
__device__ __forceinline__ uint32_t ntt_twiddle_const(uint32_t v)
{
    constexpr int e = ((E % M) + M) % M;
    if constexpr (e == 0) {
        return v;
    }
    else if constexpr (2 * e == M) {
        return mod_sub(0u, v);
    }
    else {
        constexpr int64_t ex = static_cast<int64_t>(2048 / M) * e * (INV ? -1 : 1);
        constexpr uint32_t w = ntt31::to_mont_c(ntt31::psi_pow(ex));
        return mont_mul(v, w);
    }
}

// In-register M-point cyclic DFT mod P (radix-2 DIF, natural order).
template <int M, bool INV>
__device__ __forceinline__ void ntt_dft_reg(uint32_t* x)
{
    static_for<4>([&](auto sc) {
        constexpr int half = (M >> 1) >> decltype(sc)::value;
        if constexpr (half >= 1) {
#pragma unroll
            for (int start = 0; start < M; start += 2 * half) {
                static_for<half>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const uint32_t u = x[start + j];
                    const uint32_t v = x[start + j + half];
                    x[start + j] = mod_add(u, v);
                    x[start + j + half] =
                        ntt_twiddle_const<M, j*(M / (2 * half)), INV>(mod_sub(u, v));
                });
            }
        }
    });
    uint32_t tmp[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
        int r = 0;
#pragma unroll
        for (int b = 1, s = M / 2; b < M; b <<= 1, s >>= 1)
            if (i & b) r |= s;
        tmp[r] = x[i];
    }
#pragma unroll
    for (int i = 0; i < M; i++) x[i] = tmp[i];
}

// Device tables (Montgomery form), built at init:
//   twl[k2*64 + l]  = psi^(l (1 + 2 k2))        forward step-2 twiddle x lane twist
//   twli[k2*64 + l] = psi^-(l (1 + 2 k2))       inverse
//   tw64[c*8 + a]   = omega^(16 a c)            (omega = psi^2; inverse: conjugate power)
//   tw64i[c*8 + a]  = omega^-(16 a c)
//   cr[r]  = psi^(64 r), cri[r] = psi^(-64 r) * N^-1   (r < 16)
//   pw[t]  = psi^t (t < 2048): evaluation-point powers for X^a - 1
struct NttTables {
    const uint32_t* twl;
    const uint32_t* twli;
    const uint32_t* tw64;
    const uint32_t* tw64i;
    const uint32_t* cr;
    const uint32_t* cri;
    const uint32_t* pw;
};

constexpr int kNttLdsWords = 16 * 72;  // per-wave scratch (transpose 1 needs 16 x 72)

// Transpose addresses (4-byte elements, 64 banks; both sides conflict-free).
__device__ __forceinline__ int ntt_t1(int k2, int l) { return k2 * 72 + l; }
__device__ __forceinline__ int ntt_t2(int q, int g, int a, int c)
{
    return (q * 2 + g) * 64 + 8 * ((a + q) & 7) + ((c + q) & 7);
}

// Forward negacyclic NTT of x[r] = coefficient (lane + 64 r), in [0, P).
__device__ __forceinline__ void ntt_forward(uint32_t (&x)[16], uint32_t* lds, const NttTables& tb)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 1; r < 16; r++) x[r] = mont_mul(x[r], tb.cr[r]);
    ntt_dft_reg<16, false>(x);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) x[k2] = mont_mul(x[k2], tb.twl[k2 * 64 + lane]);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) lds[ntt_t1(k2, lane)] = x[k2];
    wave_lds_sync();
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
        for (int b = 0; b < 8; b++) x[g * 8 + b] = lds[ntt_t1(g * 8 + q, 8 * b + lo)];
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 2; g++) {
        ntt_dft_reg<8, false>(&x[g * 8]);
#pragma unroll
        for (int c = 1; c < 8; c++) x[g * 8 + c] = mont_mul(x[g * 8 + c], tb.tw64[c * 8 + lo]);
#pragma unroll
        for (int c = 0; c < 8; c++) lds[ntt_t2(q, g, lo, c)] = x[g * 8 + c];
    }
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
        for (int a = 0; a < 8; a++) x[g * 8 + a] = lds[ntt_t2(q, g, a, lo)];
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 2; g++) ntt_dft_reg<8, false>(&x[g * 8]);
}

// Inverse (incl. N^-1): spectrum in the forward's register layout ->
// coefficients x[r] = coefficient (lane + 64 r), canonical.
__device__ __forceinline__ void ntt_inverse(uint32_t (&x)[16], uint32_t* lds, const NttTables& tb)
{
    const int lane = threadIdx.x & 63;
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < 2; g++) {
        ntt_dft_reg<8, true>(&x[g * 8]);
#pragma unroll
        for (int a = 0; a < 8; a++) lds[ntt_t2(q, g, a, lo)] = x[g * 8 + a];
    }
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
        for (int c = 0; c < 8; c++) x[g * 8 + c] = lds[ntt_t2(q, g, lo, c)];
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 2; g++) {
#pragma unroll
        for (int c = 1; c < 8; c++) x[g * 8 + c] = mont_mul(x[g * 8 + c], tb.tw64i[c * 8 + lo]);
        ntt_dft_reg<8, true>(&x[g * 8]);
#pragma unroll
        for (int b = 0; b < 8; b++) lds[ntt_t1(g * 8 + q, 8 * b + lo)] = x[g * 8 + b];
    }
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) x[k2] = lds[ntt_t1(k2, lane)];
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) x[k2] = mont_mul(x[k2], tb.twli[k2 * 64 + lane]);
    ntt_dft_reg<16, true>(x);
#pragma unroll
    for (int r = 0; r < 16; r++) x[r] = mont_mul(x[r], tb.cri[r]);
}

// Reference conversions (include/ntt_small_modulus.cuh:284-361), lvl1 prime.
__device__ __forceinline__ uint32_t torus_to_ntt31(uint32_t a)
{
    // floor((a P + 2^31) / 2^32) in [0, P]; P itself is the residue 0
    const uint64_t v = (static_cast<uint64_t>(a) * ntt31::P + (1ULL << 31)) >> 32;
    return v >= ntt31::P ? static_cast<uint32_t>(v - ntt31::P) : static_cast<uint32_t>(v);
}
// v mod P for |v| < P: as u32 a negative v is above 2^31 > P and v + P
// wraps to P - |v|, so the unsigned min picks the residue either way
// (two VALU ops, no compare and select).
__device__ __forceinline__ uint32_t signed_to_ntt31(int32_t v)
{
    const uint32_t u = static_cast<uint32_t>(v);
    return __builtin_elementwise_min(u, u + ntt31::P);
}
__device__ __forceinline__ uint32_t ntt31_to_torus32(uint32_t val)
{
    constexpr uint64_t inv_mul = (1ULL << 63) / ntt31::P;
    const bool neg = val > ntt31::HALF_P;
    const uint32_t mag = neg ? ntt31::P - val : val;
    const uint64_t temp = static_cast<uint64_t>(mag) * inv_mul;
    const uint32_t t32 = static_cast<uint32_t>((temp + (1ULL << 30)) >> 31);
    return neg ? 0u - t32 : t32;
}

}  // namespace cufhe_amd
