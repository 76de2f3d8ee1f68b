// C++ drop-in for cuFHEpp's annihilation key switching (reference
// include/annihilate_gpu.cuh:15-54, src/annihilate_gpu.cu:605-815): the same
// names, key types and argument meaning over the C ABI of libcufhe_amd.so
// (cufhe_annihilate_*, tfhe_annihilate_key_gen). Instantiated for
// TFHEpp::lvl1param and TFHEpp::lvl2param, the two rings the reference
// instantiates besides its AH* parameter sets (src/annihilate_gpu.cu:805-815).
//
// Differences a caller can see: the FFT backend at both levels, the NTT
// backend at lvl2param as a run-time choice instead of a build option; the kernel
// keeps the TRLWE in LDS, so the workspace arguments of the *WithWorkspace
// forms are accepted and not used, and AnnihilateKeySwitching allocates
// nothing; one GPU per process (gpuNum must be 0 or 1 as a count, 0 as an
// index).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "cufhe_amd.h"
#include "tfhe_host.hpp"

namespace cufhe {

template <class P>
using EvalAutoKeyPolynomial = std::array<TFHEpp::HalfTRGSW<P>, P::k>;
template <class P>
using AnnihilateKeyPolynomial = std::array<EvalAutoKeyPolynomial<P>, P::nbit>;

namespace detail_ah {
template <class P>
constexpr int level()
{
    static_assert(std::is_same_v<P, TFHEpp::lvl1param> || std::is_same_v<P, TFHEpp::lvl2param>,
                  "annihilation is provided for lvl1param and lvl2param");
    return std::is_same_v<P, TFHEpp::lvl1param> ? 1 : 2;
}
inline void check(int rc, const char* what)
{
    if (rc != CUFHE_OK) {
        std::fprintf(stderr, "cufhe: %s failed (%d): %s\n", what, rc, cufhe_last_error());
        std::exit(-1);
    }
}
template <class P>
constexpr std::size_t words()
{
    return sizeof(AnnihilateKeyPolynomial<P>) / sizeof(typename P::T);
}
}  // namespace detail_ah

// For bit < nbit and d = 2^(bit+1) + 1: the half TRGSW of tau_d(key) under
// key, at the level's noise (src/annihilate_gpu.cu:605-622).
template <class P>
void AnnihilateKeyPolynomialGen(AnnihilateKeyPolynomial<P>& ahk, const TFHEpp::Key<P>& key)
{
    static_assert(P::k == 1, "k = 1 rings");
    static_assert(detail_ah::words<P>() == static_cast<std::size_t>(P::nbit) * P::l * 2 * P::n,
                  "AnnihilateKeyPolynomial is the flat [bit][row][a, b][N] layout");
    const uint64_t seed = TFHEpp::detail::encrypt_seed() ^ (0x5a17ULL << 48) ^
                          TFHEpp::detail::encrypt_counter()++;
    detail_ah::check(tfhe_annihilate_key_gen(detail_ah::level<P>(), seed, key.data(), P::α,
                                             &ahk[0][0][0][0][0], detail_ah::words<P>()),
                     "AnnihilateKeyPolynomialGen");
}
template <class P>
void AnnihilateKeyPolynomialGen(AnnihilateKeyPolynomial<P>& ahk, const TFHEpp::SecretKey& sk)
{
    AnnihilateKeyPolynomialGen<P>(ahk, sk.key.get<P>());
}

// Transforms the key on the GPU and keeps it for the level (replacing a
// previous one). backend: CUFHE_BACKEND_FFT (the reference's default build)
// or, for lvl2param, CUFHE_BACKEND_NTT (its USE_FFT=OFF build).
template <class P>
void AnnihilateKeyPolynomialToDevice(const AnnihilateKeyPolynomial<P>& ahk, const int gpuNum,
                                     const int backend = CUFHE_BACKEND_FFT)
{
    if (gpuNum != 1) {
        std::fprintf(stderr, "cufhe: AnnihilateKeyPolynomialToDevice: one GPU per process\n");
        std::exit(-1);
    }
    detail_ah::check(cufhe_annihilate_key_load_ex(detail_ah::level<P>(), &ahk[0][0][0][0][0],
                                                  detail_ah::words<P>(), backend),
                     "AnnihilateKeyPolynomialToDevice");
}

template <class P>
void DeleteAnnihilateKey(const int)
{
    detail_ah::check(cufhe_annihilate_key_delete(detail_ah::level<P>()), "DeleteAnnihilateKey");
}

// out = annihilation key switch of the device TRLWE in (asynchronous on st).
template <class P>
void AnnihilateKeySwitching(typename P::T* const out, const typename P::T* const in,
                            const hipStream_t st, const int gpuNum)
{
    if (gpuNum != 0) {
        std::fprintf(stderr, "cufhe: AnnihilateKeySwitching: GPU index must be 0\n");
        std::exit(-1);
    }
    detail_ah::check(cufhe_annihilate_key_switch(detail_ah::level<P>(), out, in, st),
                     "AnnihilateKeySwitching");
}
template <class P>
void AnnihilateKeySwitchingWithWorkspace(typename P::T* const out, const typename P::T* const in,
                                         typename P::T* const /*evaledauto: not needed*/,
                                         const hipStream_t st, const int gpuNum)
{
    AnnihilateKeySwitching<P>(out, in, st, gpuNum);
}
template <class P>
void AnnihilateKeySwitchingBatchWithWorkspace(typename P::T* const out, const std::size_t out_stride,
                                              const typename P::T* const in,
                                              const std::size_t in_stride,
                                              typename P::T* const /*evaledauto: not needed*/,
                                              const std::size_t /*evaledauto_stride*/,
                                              const std::size_t batch_count, const hipStream_t st,
                                              const int gpuNum)
{
    if (gpuNum != 0) {
        std::fprintf(stderr, "cufhe: AnnihilateKeySwitchingBatchWithWorkspace: GPU index must be 0\n");
        std::exit(-1);
    }
    detail_ah::check(cufhe_annihilate_key_switch_batch(detail_ah::level<P>(), out, out_stride, in,
                                                       in_stride, batch_count, st),
                     "AnnihilateKeySwitchingBatchWithWorkspace");
}

}  // namespace cufhe
