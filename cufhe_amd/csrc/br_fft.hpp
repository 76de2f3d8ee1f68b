// Blind rotation in the Fourier domain (FP64 FFNT backend).
//
// Reference: include/gatebootstrapping_gpu.cuh:69-221 (Accumulate, CMUX
// form: acc += ExtProd((X^a - 1) acc, BK_i)), :895-1033 (AccumulateKeyBundle:
// acc = ExtProd(acc, one + BK2 (X^a1 - 1) + BK1 (X^a0 - 1) + BK0 (X^a01 - 1))),
// :801-884 and :1335-1429 (BR drivers), src/bootstrap_gpu.cu:1201-1221
// (SampleExtractIndex).
//
// MI355X mapping: a workgroup runs C jobs; each job is TWO wavefronts, wave j
// owning TRLWE component j of the accumulator for the whole rotation. Per
// iteration wave j decomposes its component into l digit polynomials,
// transforms each with a wave-private FFT (fft_wave.hpp), and multiplies
// them by the key rows j*l+q for BOTH output components. The partial for the
// other component is exchanged with the partner wave through LDS (two
// workgroup barriers), then each wave inverse-transforms its own component.
// The C jobs of a workgroup walk the key in lock step, so each key tile is
// fetched from L2 once per workgroup and reused by the other waves from the
// CU's L1.
#pragma once

#include "device_common.hpp"
#include "fft_wave.hpp"

namespace cufhe_amd {

// Registers of key values kept in flight ahead of the multiply-accumulate.
#ifndef CUFHE_BR_PREFETCH
#define CUFHE_BR_PREFETCH 2
#endif

template <class P>
inline constexpr size_t kTrgswFourier =
    static_cast<size_t>(TFHEpp::BootstrappingTRGSWRows<P>) * (P::k + 1) * (P::n / 2);

// Device pointers of one Fourier-domain key set for target level P.
template <class P>
struct FourierKeys {
    const cplx* bk;   // [trgsw][row][out][H] spectra in storage order
    const cplx* e2n;  // [2N] exp(i*pi*t/N): evaluation-point powers
    FFTTables fft;
};

// W_8^m = exp(-2*pi*i*m/8) for m in [0, 8).
__constant__ const double kW8[8][2] = {
    {1.0, 0.0},
    {0.70710678118654752440, -0.70710678118654752440},
    {0.0, -1.0},
    {-0.70710678118654752440, -0.70710678118654752440},
    {-1.0, 0.0},
    {-0.70710678118654752440, 0.70710678118654752440},
    {0.0, 1.0},
    {0.70710678118654752440, 0.70710678118654752440}};

// Spectrum of X^a - 1 at the bins of this lane, without a table of 2N
// spectra. fft_forward computes Z[k] = p(zeta_k) with
// zeta_k = exp(i*pi*(1 - 4k)/N); bin (lane c+8q, register g*8+d) has
// k = (q+8g) + R*(c+8d), so zeta_k^a = E[a*(1-4*kb) mod 2N] * W_8^(a*d mod 8)
// with kb = (q+8g) + R*c and E[t] = exp(i*pi*t/N). One gather per (a, g)
// and a constant twiddle per register give X^a - 1 at all R bins.
template <class P>
struct XaiGen {
    static constexpr int N = P::n, R = N / 128, G = R / 8;
    cplx base[G];
    uint32_t a;
    __device__ __forceinline__ void init(uint32_t a_, const cplx* e2n, int lane)
    {
        a = a_;
        const int c = lane & 7, q = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int kb = (q + 8 * g) + R * c;
            const int t = static_cast<int>(a) * (1 - 4 * kb);
            base[g] = e2n[t & (2 * N - 1)];
        }
    }
    // Forget that base[] is loop-invariant: recomputing at() per digit is a
    // few FMAs, keeping all R values live across the digit loop is 4R VGPRs.
    __device__ __forceinline__ void pin()
    {
#pragma unroll
        for (int g = 0; g < G; g++) asm volatile("" : "+v"(base[g].x), "+v"(base[g].y));
    }
    // value at register index reg = g*8 + d
    __device__ __forceinline__ cplx at(int reg) const
    {
        const int g = reg >> 3, d = reg & 7;
        const int m = static_cast<int>((a * static_cast<uint32_t>(d)) & 7u);
        const cplx w = {kW8[m][0], kW8[m][1]};
        const cplx v = cmul(base[g], w);
        return {v.x - 1.0, v.y};
    }
};

// Per-wave accumulator storage: Torus32 coefficients live in VGPRs, Torus64
// ones in LDS (keeps the lvl2 wave under 256 VGPRs).
template <class P>
struct AccStore {
    using T = typename P::T;
    static constexpr int N = P::n;
    static constexpr int R = N / 128;
    static constexpr bool kLds = sizeof(T) == 8;
    T reg[kLds ? 1 : 2 * R];
    T* lds;  // N words when kLds
    // coefficient index for slot r of this lane
    __device__ __forceinline__ static int coef_index(int r, int lane)
    {
        return r < R ? lane + 64 * r : (N / 2) + lane + 64 * (r - R);
    }
    __device__ __forceinline__ T get(int r, int lane) const
    {
        if constexpr (kLds)
            return lds[coef_index(r, lane)];
        else
            return reg[r];
    }
    __device__ __forceinline__ void set(int r, int lane, T v)
    {
        if constexpr (kLds)
            lds[coef_index(r, lane)] = v;
        else
            reg[r] = v;
    }
};

// Test vector X^bar * (mu, ..., mu) (reference
// include/gatebootstrapping_gpu.cuh:31-53): coefficient i is -mu when
// (i < bar mod N) xor (bar >= N), and every coefficient is mu when bar == 2N.
template <class P>
__device__ __forceinline__ typename P::T test_vector_coef(int i, uint32_t bar, typename P::T mu)
{
    using T = typename P::T;
    if (bar == 2 * P::n) return mu;
    const bool neg = (static_cast<uint32_t>(i) < (bar & (P::n - 1))) ^ ((bar >> P::nbit) != 0);
    return neg ? static_cast<T>(0 - mu) : mu;
}

// X^a * p evaluated at coefficient i (negacyclic): +-p[(i - a) mod N].
template <class P>
__device__ __forceinline__ typename P::T rotated_coef(const typename P::T* p, int i, uint32_t a)
{
    using T = typename P::T;
    const T v = p[(i - static_cast<int>(a)) & (P::n - 1)];
    const bool neg = (static_cast<uint32_t>(i) < (a & (P::n - 1))) ^ ((a >> P::nbit) != 0);
    return neg ? static_cast<T>(0 - v) : v;
}

// LDS shape of one wave: one spectrum buffer of H complex; the gadget
// digits are processed one pass at a time.
template <class P>
struct BrShape {
    static constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l;
    static constexpr int NB = 1;
    static constexpr int kPasses = L;
};

// One external-product step of the blind rotation for the wave owning
// component j. KB: acc_j <- ExtProd(acc, bundle(a0, a1)); otherwise
// acc_j <- acc_j + ExtProd((X^a0 - 1) acc, trgsw).
// Per pass: the wave transforms its NB digit polynomials into my_spec
// (each buffer doubles as that FFT's scratch), the pair synchronises, then
// each wave multiplies the spectra of BOTH components (its own and
// partner_spec) by the key column of its own output component only.
// All 2C waves of the workgroup must call this the same number of times.
// GHOST: the same loop nest and workgroup barriers with no loads, stores or
// arithmetic, for waves of the workgroup that take no part in a rotation
// but must keep step with it (every barrier counts all waves that have not
// ended); none of the pointer arguments is dereferenced then.
// DIRECT (with KB false): acc_j <- ExtProd(acc, trgsw), the external
// product of the accumulator itself by one TRGSW, for CMUX (reference
// src/bootstrap_gpu.cu:700-813 __CMUXNTT__ computes c0 + ExtProd(c1 - c0, C)).
template <class P, bool KB, bool GHOST = false, bool DIRECT = false>
__device__ __forceinline__ void br_step_fft(AccStore<P>& acc, int j, uint32_t a0, uint32_t a1,
                                            const cplx* trgsw0, cplx* my_spec,
                                            const cplx* partner_spec, const FourierKeys<P>& K)
{
    static_assert(!(KB && DIRECT), "DIRECT takes a single TRGSW");
    constexpr bool kSelf = KB || DIRECT;  // digits of acc itself; result assigned
    using T = typename P::T;
    using S = BrShape<P>;
    constexpr int N = S::N, H = S::H, R = S::R, L = S::L, NB = S::NB;
    constexpr size_t TS = kTrgswFourier<P>;
    const int lane = threadIdx.x & 63;

    // Digit source: acc itself (KB, DIRECT) or (X^a0 - 1) acc (plain CMUX).
    T src[(kSelf || AccStore<P>::kLds) ? 1 : 2 * R];
    if constexpr (!GHOST && !kSelf && !AccStore<P>::kLds) {
        T* rot = reinterpret_cast<T*>(my_spec);
#pragma unroll
        for (int r = 0; r < 2 * R; r++) rot[AccStore<P>::coef_index(r, lane)] = acc.reg[r];
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 2 * R; r++) {
            const int i = AccStore<P>::coef_index(r, lane);
            src[r] = rotated_coef<P>(rot, i, a0) - acc.reg[r];
        }
        wave_lds_sync();
    }
    auto src_at = [&](int r) -> T {
        if constexpr (kSelf) {
            return acc.get(r, lane);
        }
        else if constexpr (AccStore<P>::kLds) {
            const int i = AccStore<P>::coef_index(r, lane);
            return rotated_coef<P>(acc.lds, i, a0) - acc.lds[i];
        }
        else {
            return src[r];
        }
    };

    const FFTTables tb = K.fft;
    XaiGen<P> x0g, x1g, x01g;
    if constexpr (KB && !GHOST) {
        x0g.init(a0, K.e2n, lane);
        x1g.init(a1, K.e2n, lane);
        x01g.init((a0 + a1) & (2 * N - 1), K.e2n, lane);
    }
    // key column of output component j: [trgsw][row][j][H]
    const cplx* col0 = GHOST ? nullptr : trgsw0 + static_cast<size_t>(j) * H + lane;

    cplx accj[R];
#pragma unroll
    for (int d = 0; d < R; d++) accj[d] = cplx{0.0, 0.0};

#pragma unroll 1
    for (int ps = 0; ps < S::kPasses; ps++) {
        // Key loads of this pass: column j of rows (comp * L + q), comp = 0, 1.
        // They do not depend on the spectra, so the first PD registers' worth
        // is issued before the pass's forward transform (which hides their
        // latency) and the rest PD registers ahead (L2 latency is hidden by
        // loads in flight, not by occupancy).
        struct KeyVals {
            cplx k[2][KB ? 3 : 1];
        };
        // (Torus64: three registers ahead, measured 3% faster on the lvl02
        // mix than two; four spill)
        constexpr int PD = KB ? (sizeof(T) == 8 ? 3 : CUFHE_BR_PREFETCH) : R;
        const int q = ps * NB;  // NB == 1: one digit per pass
        auto load_keys = [&](int d) -> KeyVals {
            KeyVals v;
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                const cplx* kp = col0 + static_cast<size_t>((comp * L + q) * 2) * H + d * 64;
                v.k[comp][0] = kp[0];
                if constexpr (KB) {
                    v.k[comp][1] = kp[TS];
                    v.k[comp][2] = kp[2 * TS];
                }
            }
            return v;
        };
        KeyVals kv[PD];
        if constexpr (!GHOST) {
#pragma unroll
            for (int d = 0; d < PD; d++) kv[d] = load_keys(d);
        }
        if constexpr (!GHOST) {
            const int qd = ps * NB;
            cplx z[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                z[r] = cplx{static_cast<double>(Gadget<P>::digit(src_at(r), qd)),
                            static_cast<double>(Gadget<P>::digit(src_at(R + r), qd))};
            }
            fft_forward<H>(z, my_spec, tb);
#pragma unroll
            for (int d = 0; d < R; d++) lds_store(&my_spec[d * 64 + lane], z[d]);
        }
        __syncthreads();
        if constexpr (KB && !GHOST) {
            x0g.pin();
            x1g.pin();
            x01g.pin();
        }
        const double hq = ldexp(1.0, -static_cast<int>((q + 1) * P::Bgbit));
#pragma unroll
        for (int d = 0; d < (GHOST ? 0 : R); d++) {
            const KeyVals cur = kv[d % PD];
            if (d + PD < R) kv[d % PD] = load_keys(d + PD);
            cplx x0, x1, x01;
            if constexpr (KB) {
                x0 = x0g.at(d);
                x1 = x1g.at(d);
                x01 = x01g.at(d);
            }
            cplx s = accj[d];
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                const cplx* spec = comp == j ? my_spec : partner_spec;
                const cplx D = lds_load(&spec[d * 64 + lane]);
                cplx k;
                if constexpr (KB) {
                    k = {comp == j ? hq : 0.0, 0.0};
                    cmac(k, cur.k[comp][2], x1);
                    cmac(k, cur.k[comp][1], x0);
                    cmac(k, cur.k[comp][0], x01);
                }
                else {
                    k = cur.k[comp][0];
                }
                cmac(s, D, k);
            }
            accj[d] = s;
        }
        __syncthreads();
    }

    if constexpr (GHOST) return;
    fft_inverse<H>(accj, my_spec, tb);
    constexpr double scale = kTorusScale<T> / H;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const cplx v = accj[r];
        const T lo = double_to_torus<T>(v.x * scale);
        const T hi = double_to_torus<T>(v.y * scale);
        if constexpr (kSelf) {
            acc.set(r, lane, lo);
            acc.set(R + r, lane, hi);
        }
        else {
            acc.set(r, lane, acc.get(r, lane) + lo);
            acc.set(R + r, lane, acc.get(R + r, lane) + hi);
        }
    }
    if constexpr (AccStore<P>::kLds) wave_lds_sync();
}

// Full blind rotation + sample extraction for the wave pair of one job.
// tl0: the (pre-added) lvl0 TLWE in LDS. out: target TLWE (null: no store).
// GHOST: the rotation's loop nest and workgroup barriers only (see
// br_step_fft), for the waves of a workgroup that do not rotate; call it as
// blind_rotate_ghost<P, KB>().
template <class P, bool KB, bool GHOST = false>
__device__ __forceinline__ void blind_rotate_fft(const uint32_t* tl0, typename P::T mu,
                                                 AccStore<P>& acc, int j, cplx* my_spec,
                                                 const cplx* partner_spec,
                                                 const FourierKeys<P>& K)
{
    using T = typename P::T;
    constexpr int N = P::n, R = N / 128;
    constexpr int n0 = lvl0param::n;
    const int lane = threadIdx.x & 63;
    if constexpr (!GHOST) {
        const uint32_t bar = __builtin_amdgcn_readfirstlane(2 * N - mod_switch<P>(tl0[n0]));
#pragma unroll
        for (int r = 0; r < 2 * R; r++) {
            const int i = AccStore<P>::coef_index(r, lane);
            acc.set(r, lane, j == 0 ? static_cast<T>(0) : test_vector_coef<P>(i, bar, mu));
        }
        if constexpr (AccStore<P>::kLds) wave_lds_sync();
    }
    constexpr uint32_t rnd = kMaskRound<P>;
    auto abar = [&](int i) -> uint32_t {
        if constexpr (GHOST) return 0u;
        // the lvl0 TLWE is wave-uniform: keep the mod-switched values in SGPRs
        return __builtin_amdgcn_readfirstlane(mod_switch<P>(tl0[i] + rnd));
    };
    if constexpr (KB) {
        constexpr size_t step = 3 * kTrgswFourier<P>;
#pragma unroll 1
        for (int it = 0; it < n0 / 2; it++) {
            const uint32_t a0 = abar(2 * it), a1 = abar(2 * it + 1);
            br_step_fft<P, true, GHOST>(acc, j, a0, a1, GHOST ? nullptr : K.bk + it * step,
                                        my_spec, partner_spec, K);
        }
    }
    else {
#pragma unroll 1
        for (int it = 0; it < n0; it++) {
            br_step_fft<P, false, GHOST>(acc, j, abar(it), 0,
                                         GHOST ? nullptr : K.bk + it * kTrgswFourier<P>, my_spec,
                                         partner_spec, K);
        }
    }
}

// The workgroup barriers of one blind_rotate_fft<P, KB> without its work:
// a wave of the workgroup that takes no part in the rotation calls this in
// its place to stay in step with the rotating waves.
template <class P, bool KB>
__device__ __forceinline__ void blind_rotate_ghost(const FourierKeys<P>& K)
{
    AccStore<P> none;
    blind_rotate_fft<P, KB, true>(nullptr, typename P::T(0), none, 0, nullptr, nullptr, K);
}

// The rotated TRLWE itself: component j's N coefficients at out + j N.
template <class P>
__device__ __forceinline__ void trlwe_store(const AccStore<P>& acc, int j, typename P::T* out)
{
    constexpr int N = P::n, R = N / 128;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < 2 * R; r++)
        out[j * N + AccStore<P>::coef_index(r, lane)] = acc.get(r, lane);
}

// SampleExtractIndex<0> (reference src/bootstrap_gpu.cu:1201-1221):
// res[0] = a[0], res[i] = -a[N-i] (0 < i < N), res[N] = b[0].
template <class P>
__device__ __forceinline__ void sample_extract_store(const AccStore<P>& acc, int j,
                                                     typename P::T* out)
{
    using T = typename P::T;
    constexpr int N = P::n, R = N / 128;
    const int lane = threadIdx.x & 63;
    if (j == 0) {
#pragma unroll
        for (int r = 0; r < 2 * R; r++) {
            const int i = AccStore<P>::coef_index(r, lane);
            const T v = acc.get(r, lane);
            if (i == 0)
                out[0] = v;
            else
                out[N - i] = static_cast<T>(0 - v);
        }
    }
    else if (lane == 0) {
        out[N] = acc.get(0, 0);
    }
}

}  // namespace cufhe_amd
