// NTT-backend blind rotation with the bootstrapping key shared by the C jobs
// of a workgroup (the batched form's NTT kernel).
//
// Arithmetic: exactly br_ntt.hpp's br_step_ntt (reference
// include/gatebootstrapping_gpu.cuh:522-797, :1187-1332; drivers :801-884)
// for the three modes -- plain CMUX, 2-bit key bundle, block-binary -- so the
// outputs are bit-identical to k_br_ntt (and exact residues throughout).
//
// Key sharing as in br_fft_kbs.hpp: all jobs of the workgroup run the same
// step sequence, so the key words of one (step, digit, register group of
// 4) -- 2 output columns x 2 components x NK keys x 256 words = NK x 4 KiB, a
// chunk of 4 NK one-KiB slices -- are fetched once per workgroup (wave w
// loads slices w, w + 2C, ... PD chunks ahead into VGPRs), written into a
// two-chunk LDS ring one chunk ahead, and read by every wave from there. The
// transform tables, the X^a - 1 roots and the mod-switched lvl0 masks also
// live in LDS, so the key slices are the only vector-memory loads of the
// loop. One workgroup barrier per chunk plus one per digit pass.
#pragma once

#include "br_ntt.hpp"

namespace cufhe_amd {

#ifndef CUFHE_NTT_SHARED_PD
#define CUFHE_NTT_SHARED_PD 2  // chunks in flight per wave (divides 4)
#endif

// native 128-bit vector (a plain LLVM vector load/store; HIP's uint4 struct
// is copied with memcpy, which keeps an array of them out of registers)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int C>
struct NttShared {
    static constexpr int n0 = lvl0param::n;
    uint32_t lds[2 * C][kNttLdsWords];  // per wave: spectrum + transform scratch
    uint32_t ring[2][16][256];          // key chunks: [buffer][slice][4 registers x 64 lanes]
    uint32_t twl[16 * 64], twli[16 * 64], tw64[64], tw64i[64], cr[16], cri[16];
    uint32_t pw[2 * lvl1param::n];
    uint16_t abar[C][n0 + 4];  // mod-switched lvl0 mask
    uint32_t bar[C];
    uint32_t corr[C][2];  // block-binary correction partial sums (one per wave)
};

template <int MODE>
struct NttModeShape {
    static constexpr int NK = MODE == NTT_PLAIN ? 1 : (MODE == NTT_KB ? 3 : kBlockEll);
    static constexpr int STEP_TRGSW = MODE == NTT_PLAIN ? 1 : (MODE == NTT_KB ? 3 : kBlockEll);
    static constexpr int STEP_BITS = MODE == NTT_PLAIN ? 1 : (MODE == NTT_KB ? 2 : kBlockEll);
    static constexpr int STEPS = lvl0param::n / STEP_BITS;
    static constexpr int NS = 4 * NK;  // slices per chunk
};

// Slice s of chunk (it, q, c4): s = (col*2 + comp)*NK + e; words c4*256 + [0, 256)
// of key e of the step, row comp*l + q, output column col.
template <int MODE>
__device__ __forceinline__ const uint32_t* ntt_slice_ptr(const uint32_t* bk, int it, int q, int c4,
                                                         int s, int lane)
{
    using S = NttModeShape<MODE>;
    constexpr int N = lvl1param::n, L = lvl1param::l;
    const int col = s / (2 * S::NK), comp = (s / S::NK) & 1, e = s % S::NK;
    return bk + (static_cast<size_t>(it) * S::STEP_TRGSW + e) * kTrgswNtt +
           static_cast<size_t>(((comp * L + q) * 2 + col) * N) + c4 * 256 + lane * 4;
}

template <int MODE, int C>
__device__ __forceinline__ void blind_rotate_ntt_shared(NttShared<C>& sm, int slot, int j,
                                                        uint32_t mu, uint32_t (&acc)[16],
                                                        const NttKeys& K)
{
    using P = lvl1param;
    using S = NttModeShape<MODE>;
    constexpr int N = P::n, L = P::l, NK = S::NK, NS = S::NS;
    constexpr int W = 2 * C;
    constexpr int SPW = (NS + W - 1) / W;
    constexpr int PD = CUFHE_NTT_SHARED_PD;
    constexpr int CHUNKS = S::STEPS * L * 4;  // chunk g = (it * L + q) * 4 + c4
    static_assert(4 % PD == 0, "prefetch distance must divide 4");
    const int lane = threadIdx.x & 63;
    const int wave = slot * 2 + j;
    uint32_t* const my_lds = sm.lds[wave];
    const uint32_t* const partner_lds = sm.lds[wave ^ 1];
    const NttTables tb{sm.twl, sm.twli, sm.tw64, sm.tw64i, sm.cr, sm.cri, sm.pw};

    const uint32_t bar = sm.bar[slot];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = ntt_coef_index(r, lane);
        if (j == 0)
            acc[r] = 0;
        else if (bar == 2 * N)
            acc[r] = mu;
        else
            acc[r] = ((static_cast<uint32_t>(i) < (bar & (N - 1))) ^ ((bar >> P::nbit) != 0)) ? 0u - mu
                                                                                          : mu;
    }

    u32x4 kv[PD][SPW];
    auto load_chunk = [&](int g, u32x4 (&dst)[SPW]) {
        const int c4 = g & 3, q = (g >> 2) % L, it = g / (4 * L);
        static_for<SPW>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int s = wave + k * W;
            if (s < NS && g < CHUNKS)
                dst[k] = *reinterpret_cast<const u32x4*>(ntt_slice_ptr<MODE>(K.bk, it, q, c4, s, lane));
        });
    };
    auto store_chunk = [&](int g, const u32x4 (&src)[SPW]) {
        static_for<SPW>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int s = wave + k * W;
            if (s < NS && g < CHUNKS) *reinterpret_cast<u32x4*>(&sm.ring[g & 1][s][lane * 4]) = src[k];
        });
    };
#pragma unroll
    for (int p = 0; p < PD; p++) load_chunk(p, kv[p]);
    store_chunk(0, kv[0]);
    load_chunk(PD, kv[0]);

#pragma unroll 1
    for (int it = 0; it < S::STEPS; it++) {
        uint32_t a[NK > 3 ? NK : 3];
        if constexpr (MODE == NTT_PLAIN) {
            a[0] = __builtin_amdgcn_readfirstlane(sm.abar[slot][it]);
        }
        else if constexpr (MODE == NTT_KB) {
            a[0] = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it]);
            a[1] = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it + 1]);
            a[2] = (a[0] + a[1]) & (2 * N - 1);
        }
        else {
#pragma unroll
            for (int e = 0; e < NK; e++)
                a[e] = __builtin_amdgcn_readfirstlane(sm.abar[slot][it * NK + e]);
        }
        // digit source: acc (KB, BB) or (X^a - 1) acc (plain)
        uint32_t src[16];
        if constexpr (MODE == NTT_PLAIN) {
#pragma unroll
            for (int r = 0; r < 16; r++) my_lds[ntt_coef_index(r, lane)] = acc[r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int i = ntt_coef_index(r, lane);
                const uint32_t v = my_lds[(i - static_cast<int>(a[0])) & (N - 1)];
                const bool neg =
                    (static_cast<uint32_t>(i) < (a[0] & (N - 1))) ^ ((a[0] >> P::nbit) != 0);
                src[r] = (neg ? 0u - v : v) - acc[r];
            }
            wave_lds_sync();
        }
        else {
#pragma unroll
            for (int r = 0; r < 16; r++) src[r] = acc[r];
        }
        XaiNtt xg[MODE == NTT_PLAIN ? 1 : NK];
        if constexpr (MODE != NTT_PLAIN) {
#pragma unroll
            for (int e = 0; e < NK; e++) xg[e].init(a[e], sm.pw, lane);
        }
        uint32_t s[16];
#pragma unroll
        for (int d = 0; d < 16; d++) s[d] = 0;
#pragma unroll 1
        for (int q = 0; q < L; q++) {
            if constexpr (MODE != NTT_PLAIN) {
#pragma unroll
                for (int e = 0; e < NK; e++) xg[e].pin();
            }
            {
                uint32_t x[16];
#pragma unroll
                for (int r = 0; r < 16; r++) x[r] = signed_to_ntt31(Gadget<P>::digit(src[r], q));
                ntt_forward(x, my_lds, tb);
#pragma unroll
                for (int d = 0; d < 16; d++) my_lds[d * 64 + lane] = x[d];
            }
            uint32_t one_q = 0;
#pragma unroll
            for (int qq = 0; qq < L; qq++)
                if (qq == q) one_q = K.one_mont[qq];
            const int g0 = (it * L + q) * 4;
            // compile-time chunk index within the pass (keeps kv[] in VGPRs)
            static_for<4>([&](auto c4c) {
                constexpr int c4 = decltype(c4c)::value;
                const int g = g0 + c4;
                __syncthreads();  // chunk g in the ring; chunk g-1 consumed; spectra written
                const uint32_t (*ring)[256] = sm.ring[g & 1];
#pragma unroll
                for (int dd = 0; dd < 4; dd++) {
                    const int d = c4 * 4 + dd;
                    // products summed in pairs before one reduction (mont_dot2)
                    uint32_t D[2], kc[2];
#pragma unroll
                    for (int comp = 0; comp < 2; comp++) {
                        D[comp] = (comp == j ? my_lds : partner_lds)[d * 64 + lane];
                        const int s0 = (j * 2 + comp) * NK;
                        uint32_t k;
                        if constexpr (MODE == NTT_PLAIN) {
                            k = ring[s0][dd * 64 + lane];
                        }
                        else if constexpr (MODE == NTT_KB) {
                            k = comp == j ? one_q : 0u;
                            k = mod_add(k, mont_dot2(xg[1].at(d), ring[s0 + 2][dd * 64 + lane],
                                                     xg[0].at(d), ring[s0 + 1][dd * 64 + lane]));
                            k = mod_add(k, mont_mul(xg[2].at(d), ring[s0 + 0][dd * 64 + lane]));
                        }
                        else {
                            k = 0;
#pragma unroll
                            for (int e = 0; e + 1 < NK; e += 2)
                                k = mod_add(k, mont_dot2(xg[e].at(d), ring[s0 + e][dd * 64 + lane],
                                                         xg[e + 1].at(d),
                                                         ring[s0 + e + 1][dd * 64 + lane]));
                            if constexpr (NK % 2)
                                k = mod_add(k, mont_mul(xg[NK - 1].at(d),
                                                        ring[s0 + NK - 1][dd * 64 + lane]));
                        }
                        kc[comp] = k;
                    }
                    uint32_t acc_d = mod_add(s[d], mont_dot2(D[0], kc[0], D[1], kc[1]));
                    // keep this chunk's arithmetic ahead of the next barrier
                    asm volatile("" : "+v"(acc_d));
                    s[d] = acc_d;
                }
                store_chunk(g + 1, kv[(c4 + 1) % PD]);
                load_chunk(g + 1 + PD, kv[(c4 + 1) % PD]);
            });
            __syncthreads();  // partner done reading this wave's spectrum
        }
        ntt_inverse(s, my_lds, tb);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t t = ntt31_to_torus32(s[r]);
            acc[r] = MODE == NTT_KB ? t : acc[r] + t;
        }
    }
}

}  // namespace cufhe_amd
