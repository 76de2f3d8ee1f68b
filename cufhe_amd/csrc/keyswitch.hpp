// Identity key switching lvl1 -> lvl0 and lvl2 -> lvl0 with the gate's
// linear combination fused in front.
//
// Reference: include/keyswitch_gpu.cuh:13-23 (iksoffsetgen), :33-50
// (KeySwitchTorusConvert), :128-187 (KeySwitchFromTLWE), :189-256
// (IdentityKeySwitchPreAdd). For each output coefficient i in [0, n0]:
//   res = (i == n0) ? conv(ca*in0[dn] + cb*in1[dn] + offset)
//                   : (subset && i < n0 ? conv(ca*in0[i] + cb*in1[i]) : 0)
//   for each switched input coefficient j (all, or j >= n0 when subset):
//     tmp = ca*in0[j] + cb*in1[j] + decomp_offset + round_offset
//     for each of t signed digits v of basebit bits (most significant first):
//       res -= sign(v) * KSK[j'][digit][|v|-1][i]      (j' = j - first)
// KSK layout: [switched j][t][2^(basebit-1)][n0+1] uint32.
#pragma once

#include "device_common.hpp"

namespace cufhe_amd {

template <class KS>
struct KsTraits {
    using DT = typename KS::domainP::T;
    static constexpr int W = kTorusBits<DT>;
    static constexpr uint32_t dn = KS::domainP::k * KS::domainP::n;
    static constexpr uint32_t n0 = KS::targetP::k * KS::targetP::n;
    static constexpr uint32_t numbase = 1u << (KS::basebit - 1);
    static constexpr DT decomp_offset()
    {
        DT o = 0;
        for (uint32_t i = 1; i <= KS::t; i++)
            o += static_cast<DT>(numbase) << (W - i * KS::basebit);
        return o;
    }
    static constexpr DT kRound =
        (KS::basebit * KS::t) < static_cast<uint32_t>(W)
            ? static_cast<DT>(1) << (W - (1 + KS::basebit * KS::t))
            : 0;
    static constexpr DT kOffset = decomp_offset() + kRound;
    __device__ __forceinline__ static uint32_t convert(DT v)
    {
        if constexpr (W == 32)
            return static_cast<uint32_t>(v);
        else
            return static_cast<uint32_t>((v + (static_cast<DT>(1) << 31)) >> 32);
    }
    static constexpr size_t row_words = n0 + 1;
    static constexpr size_t digit_words = numbase * row_words;
    static constexpr size_t coef_words = KS::t * digit_words;
};

// Key switch of the combination ca*in0 + cb*in1 (+offset on the body) for
// one job, spread over NW waves of a workgroup (one job's key switch is a
// long dependent gather loop over dn x t KSK rows; split NW ways its latency
// is hidden): wave w takes the switched coefficients j in its 1/NW share,
// each lane accumulates output columns lane + 64 c (c < 10), the partial
// sums go through red[NW][n0 + 1] in LDS, and threads [0, 64 NW) sum them
// into out. All lanes of a wave see the same j and digit sequence, so
// branches are wave-uniform and each KSK row read is a coalesced run over
// columns. Integer sums mod 2^32, so the split does not change a bit.
// `wave` is the wave's index within the NW waves of this job; every thread
// of the workgroup must call it (one __syncthreads inside).
template <class KS, bool SUBSET, int NW>
__device__ __forceinline__ void key_switch_waves(uint32_t* out, const typename KS::domainP::T* in0,
                                                 const typename KS::domainP::T* in1, int32_t ca,
                                                 int32_t cb, typename KS::domainP::T offset,
                                                 const uint32_t* ksk, uint32_t* red, int wave,
                                                 int lane)
{
    using TR = KsTraits<KS>;
    using DT = typename TR::DT;
    constexpr uint32_t first = SUBSET ? TR::n0 : 0;
    constexpr int NC = (TR::n0 + 1 + 63) / 64;
    constexpr uint32_t share = (TR::dn - first + NW - 1) / NW;
    const DT ua = static_cast<DT>(static_cast<std::make_signed_t<DT>>(ca));
    const DT ub = static_cast<DT>(static_cast<std::make_signed_t<DT>>(cb));
    uint32_t res[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = lane + 64 * c;
        uint32_t r = 0;
        if (wave == 0) {  // base terms once
            if (i == static_cast<int>(TR::n0)) {
                DT v = ua * in0[TR::dn] + offset;
                if (in1) v += ub * in1[TR::dn];
                r = TR::convert(v);
            }
            else if (SUBSET && i < static_cast<int>(TR::n0)) {
                DT v = ua * in0[i];
                if (in1) v += ub * in1[i];
                r = TR::convert(v);
            }
        }
        res[c] = r;
    }
    const uint32_t j0 = first + static_cast<uint32_t>(wave) * share;
    const uint32_t j1 = j0 + share < TR::dn ? j0 + share : TR::dn;
#pragma unroll 4
    for (uint32_t j = j0; j < j1; j++) {
        DT tmp = ua * in0[j] + TR::kOffset;
        if (in1) tmp += ub * in1[j];
        const uint32_t* kj = ksk + static_cast<size_t>(j - first) * TR::coef_words;
#pragma unroll
        for (uint32_t k = 0; k < KS::t; k++) {
            const int32_t val =
                static_cast<int32_t>((tmp >> (TR::W - (k + 1) * KS::basebit)) &
                                     ((static_cast<DT>(1) << KS::basebit) - 1)) -
                static_cast<int32_t>(TR::numbase);
            if (val == 0) continue;
            const uint32_t* row = kj + k * TR::digit_words + (abs(val) - 1) * TR::row_words;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int i = lane + 64 * c;
                if (i <= static_cast<int>(TR::n0)) {
                    const uint32_t x = row[i];
                    res[c] = val > 0 ? res[c] - x : res[c] + x;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = lane + 64 * c;
        if (i <= static_cast<int>(TR::n0)) red[wave * (TR::n0 + 1) + i] = res[c];
    }
    __syncthreads();
    const int tid = wave * 64 + lane;
    for (int i = tid; i <= static_cast<int>(TR::n0); i += 64 * NW) {
        uint32_t r = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) r += red[w * (TR::n0 + 1) + i];
        out[i] = r;
    }
}

}  // namespace cufhe_amd
