// Blind rotation on the small-modulus NTT backend (lvl1, P = 2147473409).
//
// Reference: include/gatebootstrapping_gpu.cuh:522-659 (Accumulate, NTT),
// :1187-1332 (AccumulateKeyBundle, NTT), :661-797 (AccumulateBlockBinary),
// :801-884 (BR driver incl. the block-binary body correction :814-838).
// Every product is an exact residue, so the outputs are bit-exact to any
// exact restatement: ntt_mod_to_torus32( sum_rows digit (*) key mod P ).
//
// Same wave organisation as br_fft.hpp: two waves per job (wave j owns TRLWE
// component j, 16 Torus32 coefficients per lane at i = lane + 64 r), one
// pass per gadget digit, spectra exchanged through LDS, each wave multiplies
// both components' spectra by the key column of its own output component.
// Keys are stored in Montgomery form (u32 per bin, 4 KiB per polynomial).
#pragma once

#include "device_common.hpp"
#include "ntt_wave.hpp"

namespace cufhe_amd {

enum NttMode { NTT_PLAIN = 0, NTT_KB = 1, NTT_BB = 2, NTT_DIRECT = 3 };

// Block-binary lvl0 key: blocks of ell bits, at most one bit set per block.
constexpr int kBlockEll = TFHEpp::lvl0bbparam::ell;

// W_8^m in Montgomery form, W_8 = psi^256 (m < 8).
__constant__ const uint32_t kZ8Mont[8] = {
    ntt31::to_mont_c(ntt31::psi_pow(0)),    ntt31::to_mont_c(ntt31::psi_pow(256)),
    ntt31::to_mont_c(ntt31::psi_pow(512)),  ntt31::to_mont_c(ntt31::psi_pow(768)),
    ntt31::to_mont_c(ntt31::psi_pow(1024)), ntt31::to_mont_c(ntt31::psi_pow(1280)),
    ntt31::to_mont_c(ntt31::psi_pow(1536)), ntt31::to_mont_c(ntt31::psi_pow(1792))};

struct NttKeys {
    const uint32_t* bk;  // [trgsw][row][out][N] Montgomery-form spectra
    NttTables tb;
    uint32_t one_mont[lvl1param::l];  // torus_to_ntt_mod(h_q) * R mod P
};

constexpr size_t kTrgswNtt = static_cast<size_t>(TFHEpp::BootstrappingTRGSWRows<lvl1param>) *
                             (lvl1param::k + 1) * lvl1param::n;

// Spectrum of X^a - 1 at this lane's bins (Montgomery form): bin k of
// register g*8+d is k = kb + 128 d with kb = (q+8g) + 16c, so
// psi^(a(2k+1)) = pw[a(2kb+1) mod 2N] * W_8^(a d mod 8).
struct XaiNtt {
    uint32_t base[2];
    uint32_t a;
    __device__ __forceinline__ void init(uint32_t a_, const uint32_t* pw, int lane)
    {
        a = a_;
        const int c = lane & 7, q = lane >> 3;
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int kb = (q + 8 * g) + 16 * c;
            base[g] = pw[(a * static_cast<uint32_t>(2 * kb + 1)) & 2047u];
        }
    }
    // Keeps the per-bin values from being hoisted out of the digit loop
    // (16 live values per generator would cost 16 VGPRs each).
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(base[0]), "+v"(base[1])); }
    __device__ __forceinline__ uint32_t at(int reg) const
    {
        const uint32_t m = (a * static_cast<uint32_t>(reg & 7)) & 7u;
        return mod_sub(mont_mul(base[reg >> 3], kZ8Mont[m]), ntt31::R1);
    }
};

__device__ __forceinline__ int ntt_coef_index(int r, int lane) { return lane + 64 * r; }

// One blind-rotation step on the NTT backend for the wave owning component j.
//   NTT_PLAIN: acc += ExtProd((X^a[0] - 1) acc, BK)            (trgsw0 = BK_i)
//   NTT_KB   : acc  = ExtProd(acc, one + BK2 X(a1) + BK1 X(a0) + BK0 X(a0+a1))
//   NTT_BB   : acc += ExtProd(acc, sum_l BK_l X(a_l)), l < ell  (trgsw0 = BK_block*ell)
//   NTT_DIRECT: acc = ExtProd(acc, C)  (one TRGSW, for CMUX; reference
//              src/bootstrap_gpu.cu:925-1041 __CMUXNTT__)
// GHOST: the same loop nest and workgroup barriers with no loads, stores or
// arithmetic (no pointer argument is dereferenced), for the waves of a
// workgroup that do not rotate but must keep step with the ones that do.
template <int MODE, bool GHOST = false>
__device__ __forceinline__ void br_step_ntt(uint32_t (&acc)[16], int j, const uint32_t* a,
                                            const uint32_t* trgsw0, uint32_t* my_lds,
                                            const uint32_t* partner_lds, const NttKeys& K)
{
    constexpr int N = lvl1param::n, L = lvl1param::l;
    constexpr size_t TS = kTrgswNtt;
    constexpr int NK = (MODE == NTT_PLAIN || MODE == NTT_DIRECT) ? 1 : (MODE == NTT_KB ? 3 : kBlockEll);
    const int lane = threadIdx.x & 63;

    uint32_t src[16];
    if constexpr (GHOST) {
    }
    else if constexpr (MODE == NTT_PLAIN) {
#pragma unroll
        for (int r = 0; r < 16; r++) my_lds[ntt_coef_index(r, lane)] = acc[r];
        wave_lds_sync();
        const uint32_t a0 = a[0];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = ntt_coef_index(r, lane);
            const uint32_t v = my_lds[(i - static_cast<int>(a0)) & (N - 1)];
            const bool neg = (static_cast<uint32_t>(i) < (a0 & (N - 1))) ^ ((a0 >> lvl1param::nbit) != 0);
            src[r] = (neg ? 0u - v : v) - acc[r];
        }
        wave_lds_sync();
    }
    else {
#pragma unroll
        for (int r = 0; r < 16; r++) src[r] = acc[r];
    }

    XaiNtt xg[(MODE == NTT_PLAIN || MODE == NTT_DIRECT) ? 1 : NK];
    if constexpr (GHOST) {
    }
    else if constexpr (MODE == NTT_KB) {
        xg[0].init(a[0], K.tb.pw, lane);
        xg[1].init(a[1], K.tb.pw, lane);
        xg[2].init((a[0] + a[1]) & (2 * N - 1), K.tb.pw, lane);
    }
    else if constexpr (MODE == NTT_BB) {
#pragma unroll
        for (int e = 0; e < NK; e++) xg[e].init(a[e], K.tb.pw, lane);
    }
    const uint32_t* col0 = GHOST ? nullptr : trgsw0 + static_cast<size_t>(j) * N + lane;

    uint32_t s[16];
#pragma unroll
    for (int d = 0; d < 16; d++) s[d] = 0;

#pragma unroll 1
    for (int q = 0; q < L; q++) {
        if constexpr (MODE != NTT_PLAIN && MODE != NTT_DIRECT && !GHOST) {
#pragma unroll
            for (int e = 0; e < NK; e++) xg[e].pin();
        }
        if constexpr (!GHOST) {
            uint32_t x[16];
#pragma unroll
            for (int r = 0; r < 16; r++) x[r] = signed_to_ntt31(Gadget<lvl1param>::digit(src[r], q));
            ntt_forward(x, my_lds, K.tb);
#pragma unroll
            for (int d = 0; d < 16; d++) my_lds[d * 64 + lane] = x[d];
        }
        // key column j of rows (comp*L + q); keys [e] of the bundle / block
        uint32_t kv[2][2][NK];  // [slot][comp][e]
        auto load_keys = [&](int slot, int d) {
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                const uint32_t* kp = col0 + static_cast<size_t>((comp * L + q) * 2) * N + d * 64;
#pragma unroll
                for (int e = 0; e < NK; e++) kv[slot][comp][e] = kp[e * TS];
            }
        };
        // gadget identity term of the bundle (uniform; no dynamic indexing)
        uint32_t one_q = 0;
#pragma unroll
        for (int qq = 0; qq < L; qq++)
            if (qq == q) one_q = K.one_mont[qq];
        if constexpr (!GHOST) load_keys(0, 0);
        __syncthreads();
#pragma unroll
        for (int d = 0; d < (GHOST ? 0 : 16); d++) {
            if (d + 1 < 16) load_keys((d + 1) & 1, d + 1);
            const int sl = d & 1;
            // products summed in pairs before one Montgomery reduction
            // (mont_dot2): the same residues as reducing every product
            uint32_t D[2], kc[2];
#pragma unroll
            for (int comp = 0; comp < 2; comp++) {
                D[comp] = (comp == j ? my_lds : partner_lds)[d * 64 + lane];
                uint32_t k;
                if constexpr (MODE == NTT_PLAIN || MODE == NTT_DIRECT) {
                    k = kv[sl][comp][0];
                }
                else if constexpr (MODE == NTT_KB) {
                    k = comp == j ? one_q : 0u;
                    k = mod_add(k, mont_dot2(xg[1].at(d), kv[sl][comp][2], xg[0].at(d), kv[sl][comp][1]));
                    k = mod_add(k, mont_mul(xg[2].at(d), kv[sl][comp][0]));
                }
                else {
                    k = 0;
#pragma unroll
                    for (int e = 0; e + 1 < NK; e += 2)
                        k = mod_add(k, mont_dot2(xg[e].at(d), kv[sl][comp][e], xg[e + 1].at(d),
                                                 kv[sl][comp][e + 1]));
                    if constexpr (NK % 2) k = mod_add(k, mont_mul(xg[NK - 1].at(d), kv[sl][comp][NK - 1]));
                }
                kc[comp] = k;
            }
            s[d] = mod_add(s[d], mont_dot2(D[0], kc[0], D[1], kc[1]));
        }
        __syncthreads();
    }
    if constexpr (GHOST) return;
    ntt_inverse(s, my_lds, K.tb);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint32_t t = ntt31_to_torus32(s[r]);
        acc[r] = (MODE == NTT_KB || MODE == NTT_DIRECT) ? t : acc[r] + t;
    }
}

// Full rotation of the (pre-added) lvl0 TLWE tl0 (LDS) for the wave pair.
// GHOST: the loop nest and workgroup barriers only (see br_step_ntt); call
// it as blind_rotate_ntt_ghost<MODE>().
template <int MODE, bool GHOST = false>
__device__ __forceinline__ void blind_rotate_ntt(const uint32_t* tl0, uint32_t mu, uint32_t (&acc)[16],
                                                 int j, uint32_t* my_lds, const uint32_t* partner_lds,
                                                 const NttKeys& K)
{
    using P = lvl1param;
    constexpr int N = P::n, n0 = lvl0param::n;
    constexpr uint32_t rnd = kMaskRound<P>;
    const int lane = threadIdx.x & 63;
    if constexpr (!GHOST) {
        uint32_t body = tl0[n0];
        if constexpr (MODE == NTT_BB) {
            // reference :814-838: correction = sum_i (a_i - round(a_i)), the body is
            // shifted by correction/2 (signed, truncating) and rounded
            constexpr int shift = 32 - 1 - P::nbit;
            uint32_t part = 0;
            for (int i = lane; i < n0; i += 64) {
                const uint32_t moded = (tl0[i] + rnd) >> shift;
                part += tl0[i] - (moded << shift);
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) part += __shfl_xor(part, m);
            const int32_t correction = static_cast<int32_t>(part);
            body = body - static_cast<uint32_t>(correction / 2) + rnd;
        }
        const uint32_t bar = __builtin_amdgcn_readfirstlane(2 * N - mod_switch<P>(body));
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = ntt_coef_index(r, lane);
            if (j == 0) {
                acc[r] = 0;
            }
            else if (bar == 2 * N) {
                acc[r] = mu;
            }
            else {
                const bool neg = (static_cast<uint32_t>(i) < (bar & (N - 1))) ^ ((bar >> P::nbit) != 0);
                acc[r] = neg ? 0u - mu : mu;
            }
        }
    }
    auto abar = [&](int i) -> uint32_t {
        if constexpr (GHOST) return 0u;
        return __builtin_amdgcn_readfirstlane(mod_switch<P>(tl0[i] + rnd));
    };
    auto key = [&](size_t trgsw) -> const uint32_t* {
        return GHOST ? nullptr : K.bk + trgsw * kTrgswNtt;
    };
    uint32_t a[kBlockEll > 3 ? kBlockEll : 3];
    if constexpr (MODE == NTT_KB) {
#pragma unroll 1
        for (int it = 0; it < n0 / 2; it++) {
            a[0] = abar(2 * it);
            a[1] = abar(2 * it + 1);
            br_step_ntt<NTT_KB, GHOST>(acc, j, a, key(static_cast<size_t>(it) * 3), my_lds,
                                       partner_lds, K);
        }
    }
    else if constexpr (MODE == NTT_BB) {
#pragma unroll 1
        for (int blk = 0; blk < n0 / kBlockEll; blk++) {
#pragma unroll
            for (int e = 0; e < kBlockEll; e++) a[e] = abar(blk * kBlockEll + e);
            br_step_ntt<NTT_BB, GHOST>(acc, j, a, key(static_cast<size_t>(blk) * kBlockEll), my_lds,
                                       partner_lds, K);
        }
    }
    else {
#pragma unroll 1
        for (int it = 0; it < n0; it++) {
            a[0] = abar(it);
            br_step_ntt<NTT_PLAIN, GHOST>(acc, j, a, key(static_cast<size_t>(it)), my_lds,
                                          partner_lds, K);
        }
    }
}

// The workgroup barriers of one blind_rotate_ntt<MODE> without its work: a
// wave of the workgroup that takes no part in the rotation calls this in its
// place to stay in step with the rotating waves.
template <int MODE>
__device__ __forceinline__ void blind_rotate_ntt_ghost(const NttKeys& K)
{
    uint32_t none[16];
    blind_rotate_ntt<MODE, true>(nullptr, 0u, none, 0, nullptr, nullptr, K);
}

// The rotated TRLWE itself: component j's N coefficients at out + j N.
__device__ __forceinline__ void trlwe_store_ntt(const uint32_t (&acc)[16], int j, uint32_t* out)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < 16; r++) out[j * lvl1param::n + ntt_coef_index(r, lane)] = acc[r];
}

// SampleExtractIndex<0> for the NTT coefficient layout.
__device__ __forceinline__ void sample_extract_ntt(const uint32_t (&acc)[16], int j, uint32_t* out)
{
    constexpr int N = lvl1param::n;
    const int lane = threadIdx.x & 63;
    if (j == 0) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = ntt_coef_index(r, lane);
            if (i == 0)
                out[0] = acc[r];
            else
                out[N - i] = 0u - acc[r];
        }
    }
    else if (lane == 0) {
        out[N] = acc[0];
    }
}

}  // namespace cufhe_amd
