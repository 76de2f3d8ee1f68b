// Device-side parameter traits, gadget decomposition and torus conversions
// shared by the blind-rotation and key-switching kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

#include "../../include/tfhe_params.hpp"

namespace cufhe_amd {

using TFHEpp::lvl0param;
using TFHEpp::lvl1param;
using TFHEpp::lvl2param;

template <class T>
inline constexpr int kTorusBits = std::numeric_limits<T>::digits;

// Gadget decomposition of a target-level coefficient (reference
// include/gatebootstrapping_gpu.cuh:20-29 offsetgen and :123-129 / :931-948):
//   digit_q = ((c + offset + round) >> (w - (q+1)*Bgbit)) & (Bg-1) - Bg/2
template <class P>
struct Gadget {
    using T = typename P::T;
    static constexpr int W = kTorusBits<T>;
    static constexpr T offset()
    {
        T o = 0;
        for (uint32_t i = 1; i <= P::l; i++)
            o += static_cast<T>(P::Bg / 2) << (W - i * P::Bgbit);
        return o;
    }
    static constexpr T kOffset = offset() + (static_cast<T>(1) << (W - P::l * P::Bgbit - 1));
    __device__ __forceinline__ static int32_t digit(T c, int q)
    {
        const T t = c + kOffset;
        return static_cast<int32_t>((t >> (W - (q + 1) * P::Bgbit)) & (P::Bg - 1)) -
               static_cast<int32_t>(P::Bg / 2);
    }
    // h_q / 2^w as a real: the gadget entry of row digit q, in torus units.
    __host__ __device__ static constexpr double h_real(int q)
    {
        double v = 1.0;
        for (uint32_t i = 0; i < (q + 1) * P::Bgbit; i++) v *= 0.5;
        return v;
    }
};

// Mod switch of a lvl0 Torus32 value to [0, 2N) (reference
// include/gatebootstrapping_gpu.cuh:10-18): phase >> (32 - 1 - nbit).
template <class TP>
__device__ __forceinline__ uint32_t mod_switch(uint32_t phase)
{
    return phase >> (32 - 1 - TP::nbit);
}
// Rounding offset the reference adds to the mask coefficients
// (include/gatebootstrapping_gpu.cuh:840-841, :1364-1367).
template <class TP>
inline constexpr uint32_t kMaskRound = 1u << (32 - 2 - TP::nbit);

// Double (already scaled to torus units * 2^w) to a Torus32/Torus64 word,
// rounding to nearest and wrapping modulo 2^w (reference uses llrint for
// Torus32 and include/utils_gpu.cuh:48-65 for Torus64).
template <class T>
__device__ __forceinline__ T double_to_torus(double v)
{
    if constexpr (sizeof(T) == 4) {
        // rint(v) mod 2^32 in four instructions: q = rint(v / 2^32) and
        // r = v - q 2^32 (exact: |r| <= 2^31, a multiple of v's ulp), then
        // the low word of r + 1.5 * 2^52 is rint(r) mod 2^32 (round to nearest
        // even at the unit position; |r| < 2^51). rint(v) = q 2^32 + rint(r)
        // with q 2^32 even, so the result equals (uint32_t)(int64_t)rint(v).
        const double q = rint(v * (1.0 / 4294967296.0));
        const double r = fma(-q, 4294967296.0, v);
        const double t = r + 6755399441055744.0;
        return static_cast<T>(static_cast<uint64_t>(__double_as_longlong(t)));
    }
    else {
        const double r = rint(v);
        if (fabs(r) < 9223372036854775808.0) return static_cast<T>(static_cast<int64_t>(r));
        const uint64_t bits = static_cast<uint64_t>(__double_as_longlong(r));
        const int ex = static_cast<int>((bits >> 52) & 0x7FF) - 1075;
        const uint64_t m = (bits & 0x000FFFFFFFFFFFFFULL) | 0x0010000000000000ULL;
        const uint64_t mag = ex >= 64 ? 0 : (m << ex);  // ex >= 11 here
        return (bits >> 63) ? static_cast<T>(0 - mag) : static_cast<T>(mag);
    }
}

template <class T>
inline constexpr double kTorusScale = sizeof(T) == 4 ? 4294967296.0 : 18446744073709551616.0;

// One blind-rotation job: lvl0 input = ca*in0 + cb*in1 (+offset on b),
// test vector +-mu, output the sample-extracted TLWE of the target level.
struct BrJob {
    const uint32_t* in0;
    const uint32_t* in1;  // may be null (treated as zero)
    void* out;
    int32_t ca;
    int32_t cb;
    uint32_t offset;
    uint32_t flags;  // kBrOutTrlwe: store the rotated TRLWE instead of its sample
};
// BrJob::flags bit: out receives the (k+1) N-word TRLWE (reference
// BootstrapTLWE2TRLWE, src/bootstrap_gpu.cu:2371-2388) instead of the
// extracted TLWE.
constexpr uint32_t kBrOutTrlwe = 1;
static_assert(sizeof(BrJob) == 40, "BrJob layout");

// Compute units of the current device (queried once per device).
inline int device_cu_count()
{
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int v = __atomic_load_n(&cached[dev], __ATOMIC_RELAXED);
    if (v <= 0) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            v <= 0)
            v = 256;
        __atomic_store_n(&cached[dev], v, __ATOMIC_RELAXED);
    }
    return v;
}

// Grid for workgroups of up to c jobs with the jobs spread evenly: the fewest
// whole waves of CUs that hold njobs (never more workgroups than jobs).
inline int spread_grid(int njobs, int c)
{
    const int cus = device_cu_count();
    const int rounds = (njobs + c * cus - 1) / (c * cus);
    return njobs < rounds * cus ? njobs : rounds * cus;
}

// One key-switch job: out_lvl0 = KS(ca*in0 + cb*in1 (+offset on b)).
struct KsJob {
    const void* in0;
    const void* in1;  // may be null
    uint32_t* out;
    int32_t ca;
    int32_t cb;
    uint64_t offset;
};
static_assert(sizeof(KsJob) == 40, "KsJob layout");

// One lvl1 CMUX res = c0 + ExtProd(c1 - c0, cs): TRLWEs of 2 x N words, cs a
// TRGSW in the loaded backend's transform domain. Same layout as the C ABI's
// cufhe_cmux_desc.
struct CmuxDesc {
    uint32_t* res;
    const void* cs;
    const uint32_t* c1;
    const uint32_t* c0;
};
static_assert(sizeof(CmuxDesc) == 32, "CmuxDesc layout");
// A batch of CMUXes: per-CMUX descriptors (device array), or, when descs is
// null, CMUX i at res/c1/c0 + 2N i and cs + cs_bytes i.
struct CmuxBatch {
    const CmuxDesc* descs;
    uint32_t* res;
    const void* cs;
    const uint32_t* c1;
    const uint32_t* c0;
    size_t cs_bytes;
    __device__ __forceinline__ CmuxDesc at(int i) const
    {
        if (descs) return descs[i];
        const size_t o = static_cast<size_t>(i) * 2 * lvl1param::n;
        return CmuxDesc{res + o, static_cast<const char*>(cs) + cs_bytes * i, c1 + o, c0 + o};
    }
};

// Elementwise combination of TLWEs (MUX/NMUX combine, NOT, COPY):
// out[i] = s1*r1[i] + s0*r0[i], plus offset on the body word (i = words-1).
struct LinJob {
    const uint32_t* r1;
    const uint32_t* r0;  // may be null
    uint32_t* out;
    int32_t s1;
    int32_t s0;
    uint32_t offset;
    int32_t words;
};

}  // namespace cufhe_amd
