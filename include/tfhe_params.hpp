// Parameter sets and ciphertext/key layouts for TFHE gate bootstrapping.
//
// The reference consumes these symbols from TFHEpp (absent here); the names
// and memory layouts follow the index arithmetic of the reference kernels:
//   BootstrappingTRGSWRows   include/ntt_small_modulus.cuh:30-32
//   TRGSW row/poly layout    src/bootstrap_gpu.cu:70-75
//   key-bundle triple 3i+j   include/gatebootstrapping_gpu.cuh:1373-1375
//   KSK [j][digit][|v|-1][i] include/keyswitch_gpu.cuh:106-113
// Values that only TFHEpp fixes (noise widths, key-switch base/length, the
// block-binary set) are pinned here by this project; README.md:39-42 of the
// reference fixes n=636, N=1024, l=2, Bg=256 and N=2048, l=4, Bg=1024.
#pragma once

#include <array>
#include <cstdint>
#include <type_traits>
#include <vector>

namespace TFHEpp {

// lvl0: binary LWE key of dimension 636, Torus32.
struct lvl0param {
    static constexpr std::uint32_t n = 636;
    static constexpr std::uint32_t k = 1;
    using T = std::uint32_t;
    static constexpr std::make_signed_t<T> μ = 1 << 29;  // 1/8
    static constexpr std::make_signed_t<T> mu = μ;
    static constexpr double α = 0.0000925119974676756;  // 2^-13.4
    static constexpr double alpha = α;
};

// lvl1: ring degree 1024, k=1, l=2, Bg=2^8, Torus32.
struct lvl1param {
    static constexpr std::uint32_t nbit = 10;
    static constexpr std::uint32_t n = 1u << nbit;
    static constexpr std::uint32_t k = 1;
    static constexpr std::uint32_t l = 2;
    static constexpr std::uint32_t Bgbit = 8;
    static constexpr std::uint32_t Bg = 1u << Bgbit;
    static constexpr std::uint32_t lₐ = l;
    static constexpr std::uint32_t Bgₐbit = Bgbit;
    static constexpr std::uint32_t Bgₐ = Bg;
    static constexpr std::uint32_t l̅ = 1;
    static constexpr std::uint32_t l̅ₐ = 1;
    using T = std::uint32_t;
    static constexpr std::make_signed_t<T> μ = 1 << 29;
    static constexpr std::make_signed_t<T> mu = μ;
    static constexpr double α = 0.0000000342338787018369;  // 2^-24.8
    static constexpr double alpha = α;
};

// lvl2: ring degree 2048, k=1, l=4, Bg=2^10, Torus64.
struct lvl2param {
    static constexpr std::uint32_t nbit = 11;
    static constexpr std::uint32_t n = 1u << nbit;
    static constexpr std::uint32_t k = 1;
    static constexpr std::uint32_t l = 4;
    static constexpr std::uint32_t Bgbit = 10;
    static constexpr std::uint32_t Bg = 1u << Bgbit;
    static constexpr std::uint32_t lₐ = l;
    static constexpr std::uint32_t Bgₐbit = Bgbit;
    static constexpr std::uint32_t Bgₐ = Bg;
    static constexpr std::uint32_t l̅ = 1;
    static constexpr std::uint32_t l̅ₐ = 1;
    using T = std::uint64_t;
    static constexpr std::make_signed_t<T> μ = 1LL << 61;
    static constexpr std::make_signed_t<T> mu = μ;
    static constexpr double α = 0.0000000000000568434188608080;  // 2^-44
    static constexpr double alpha = α;
};

// Blind rotation lvl0 -> lvl1 (Addends = key-bundle width when enabled).
struct lvl01param {
    using domainP = lvl0param;
    using targetP = lvl1param;
    static constexpr std::uint32_t Addends = 2;
};
// Blind rotation lvl0 -> lvl2.
struct lvl02param {
    using domainP = lvl0param;
    using targetP = lvl2param;
    static constexpr std::uint32_t Addends = 2;
};
// Identity key switch lvl1 -> lvl0: t digits of basebit bits.
struct lvl10param {
    static constexpr std::uint32_t t = 7;
    static constexpr std::uint32_t basebit = 2;
    static constexpr double α = lvl0param::α;
    using domainP = lvl1param;
    using targetP = lvl0param;
};
// Identity key switch lvl2 -> lvl0.
struct lvl20param {
    static constexpr std::uint32_t t = 7;
    static constexpr std::uint32_t basebit = 2;
    static constexpr double α = lvl0param::α;
    using domainP = lvl2param;
    using targetP = lvl0param;
};

// Block-binary lvl0 key: n/ell blocks, each with at most one bit set, and
// the lvl1 key starts with the lvl0 key (subset key switching).
struct lvl0bbparam : lvl0param {
    static constexpr std::uint32_t ell = 4;  // 636 = 159 blocks of 4
};

template <class P>
using Key = std::array<typename P::T, P::n>;
template <class P>
using TLWE = std::array<typename P::T, P::k * P::n + 1>;
template <class P>
using Polynomial = std::array<typename P::T, P::n>;
template <class P>
using TRLWE = std::array<Polynomial<P>, P::k + 1>;

template <class P>
inline constexpr std::uint32_t BootstrappingTRGSWRows =
    P::k * P::lₐ * P::l̅ₐ + P::l * P::l̅;

template <class P>
using TRGSW = std::array<TRLWE<P>, BootstrappingTRGSWRows<P>>;
// l TRLWE rows: the gadget multiples of one polynomial added to b.
template <class P>
using HalfTRGSW = std::array<TRLWE<P>, P::l>;

// Words of one TRGSW in the polynomial domain.
template <class P>
inline constexpr std::size_t kTRGSWWords =
    static_cast<std::size_t>(BootstrappingTRGSWRows<P>) * (P::k + 1) * P::n;

// Number of TRGSWs in a bootstrapping key (plain or 2-bit key bundle).
template <class brP>
inline constexpr std::size_t kBKTRGSWs(bool key_bundle)
{
    return key_bundle ? (brP::domainP::k * brP::domainP::n / 2) * 3
                      : brP::domainP::k * brP::domainP::n;
}

// Words of a key-switching key [domain coefficient][digit][|digit|-1][target].
template <class ksP>
inline constexpr std::size_t kKSKWords =
    static_cast<std::size_t>(ksP::domainP::k * ksP::domainP::n) * ksP::t *
    (1u << (ksP::basebit - 1)) * (ksP::targetP::k * ksP::targetP::n + 1);

}  // namespace TFHEpp
