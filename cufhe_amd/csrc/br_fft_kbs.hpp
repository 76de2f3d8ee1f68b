// Key-bundle blind rotation at lvl1 with the bootstrapping key shared by the
// C jobs of a workgroup (the batched form's main kernel).
//
// Arithmetic: exactly br_fft.hpp's key-bundle step (reference
// include/gatebootstrapping_gpu.cuh:895-1033 AccumulateKeyBundle and
// :1335-1429 the drivers): per step the accumulator is decomposed, each
// digit polynomial transformed, multiplied by
//   one + BK2 X(a1) + BK1 X(a0) + BK0 X(a0+a1)      (X(a) = X^a - 1)
// and the product replaces the accumulator. Results are bit-identical to
// k_br_fft<lvl1param, true, C> (same operations in the same order).
//
// What changes is where the key comes from. All jobs of a workgroup run the
// same step sequence, so they need the same key words at the same time. The
// key of one (step, digit, register d) -- 2 output columns x 2 components x 3
// bundle members x 64 bins x 16 B = 12 KiB, a "chunk" of 12 one-KiB slices --
// is fetched ONCE per workgroup: wave w of the 2C waves loads slices
// w, w + 2C, ... into VGPRs PD chunks ahead, writes them into a two-chunk
// LDS ring one chunk ahead, and every wave reads the key values of its own
// output column from the ring. L2 -> CU key traffic drops C-fold and a wave
// holds PD x 4 VGPRs of key in flight instead of a register copy per use.
// Everything else a step reads sits in LDS (twiddles, X^a - 1 roots, the
// mod-switched lvl0 mask), so the only vector-memory loads in the loop are
// the key slices and the compiler's in-order vmcnt waits never stall on
// anything but the oldest slice.
//
// Synchronisation per chunk: one workgroup barrier (chunk g visible in the
// ring, every wave done with chunk g-1 whose buffer is refilled next), plus
// one barrier after each digit pass (the partner may still read this wave's
// spectrum when the next forward transform overwrites it).
#pragma once

#include "br_fft.hpp"

namespace cufhe_amd {

#ifndef CUFHE_KBS_PD
#define CUFHE_KBS_PD 2  // key chunks in flight per wave (divides 8)
#endif
#ifndef CUFHE_KBS_SWZ
#define CUFHE_KBS_SWZ 0  // bank-conflict-free transpose slots in the FFTs
#endif

// LDS of one workgroup of C jobs.
template <int C>
struct KbsShared {
    static constexpr int N = lvl1param::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    cplx spec[2 * C][H];    // per wave: spectrum + transform scratch
    cplx ring[2][12][64];   // key chunks
    cplx twl[64];       // transform tables (copies of FFTTables; row 0 of twl)
    cplx tws[64];
    cplx tw64[64];
    cplx cr[R];
    cplx e2n_hi[2 * N / 64];  // exp(i pi 64 t / N)
    cplx e2n_lo[64];          // exp(i pi t / N)
    uint16_t abar[C][n0 + 4];  // mod-switched lvl0 mask
    uint32_t bar[C];
};

// X^a - 1 at this lane's bins (as XaiGen) with E[t] = hi[t >> 6] * lo[t & 63]
// read from LDS instead of a 2N-entry table in global memory.
struct XaiGenShared {
    static constexpr int N = lvl1param::n, R = N / 128;
    cplx base;
    uint32_t a;
    __device__ __forceinline__ void init(uint32_t a_, const cplx* hi, const cplx* lo, int lane)
    {
        a = a_;
        const int c = lane & 7, q = lane >> 3;
        const int kb = q + R * c;
        const int t = (static_cast<int>(a) * (1 - 4 * kb)) & (2 * N - 1);
        base = cmul(hi[t >> 6], lo[t & 63]);
    }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(base.x), "+v"(base.y)); }
    __device__ __forceinline__ cplx at(int d) const
    {
        const int m = static_cast<int>((a * static_cast<uint32_t>(d)) & 7u);
        const cplx w = {kW8[m][0], kW8[m][1]};
        const cplx v = cmul(base, w);
        return {v.x - 1.0, v.y};
    }
};

// Key slice s of chunk (it, q, d): slice s = (col*2 + comp)*3 + e holds the
// 64 bins d*64 + [0, 64) of bundle member e, row comp*l + q, output column col.
__device__ __forceinline__ const cplx* kbs_slice_ptr(const cplx* bk, int it, int q, int d, int s,
                                                     int lane)
{
    using P = lvl1param;
    constexpr int H = P::n / 2, L = P::l;
    constexpr size_t TS = kTrgswFourier<P>;
    const int col = s / 6, comp = (s / 3) & 1, e = s % 3;
    return bk + (3 * static_cast<size_t>(it) + e) * TS +
           static_cast<size_t>(((comp * L + q) * 2 + col) * H) + d * 64 + lane;
}

// The blind rotation of one job's component j (wave `wave` of the workgroup).
template <int C>
__device__ __forceinline__ void blind_rotate_kbs(KbsShared<C>& sm, int slot, int j, uint32_t mu,
                                                 uint32_t (&acc)[16], const cplx* bk)
{
    using P = lvl1param;
    constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l, n0 = lvl0param::n;
    constexpr int W = 2 * C;             // waves per workgroup
    constexpr int SPW = (12 + W - 1) / W;  // key slices loaded per wave and chunk
    constexpr int PD = CUFHE_KBS_PD;
    constexpr int STEPS = n0 / 2;
    constexpr int CHUNKS = STEPS * L * R;  // chunk g = (it * L + q) * R + d
    static_assert(R == 8 && 8 % PD == 0, "chunk pipeline assumes 8 registers per lane");
    const int lane = threadIdx.x & 63;
    const int wave = slot * 2 + j;
    cplx* const my_spec = sm.spec[wave];
    const cplx* const partner_spec = sm.spec[wave ^ 1];
    const FFTTables tb{sm.twl, sm.tw64, sm.cr, sm.tws};

    // test vector
    const uint32_t bar = sm.bar[slot];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = AccStore<P>::coef_index(r, lane);
        acc[r] = j == 0 ? 0u : test_vector_coef<P>(i, bar, mu);
    }

    // key pipeline: slices of chunk g are held in kv[g % PD] until written
    cplx kv[PD][SPW];
    auto load_chunk = [&](int g, cplx (&dst)[SPW]) {
        const int d = g % R, q = (g / R) % L, it = g / (R * L);
#pragma unroll
        for (int k = 0; k < SPW; k++) {
            const int s = wave + k * W;
            if (s < 12 && g < CHUNKS) dst[k] = *kbs_slice_ptr(bk, it, q, d, s, lane);
        }
    };
    auto store_chunk = [&](int g, const cplx (&src)[SPW]) {
#pragma unroll
        for (int k = 0; k < SPW; k++) {
            const int s = wave + k * W;
            if (s < 12 && g < CHUNKS) lds_store(&sm.ring[g & 1][s][lane], src[k]);
        }
    };
#pragma unroll
    for (int p = 0; p < PD; p++) load_chunk(p, kv[p]);
    store_chunk(0, kv[0]);
    load_chunk(PD, kv[0]);

#pragma unroll 1
    for (int it = 0; it < STEPS; it++) {
        const uint32_t a0 = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it]);
        const uint32_t a1 = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it + 1]);
        XaiGenShared x0g, x1g, x01g;
        x0g.init(a0, sm.e2n_hi, sm.e2n_lo, lane);
        x1g.init(a1, sm.e2n_hi, sm.e2n_lo, lane);
        x01g.init((a0 + a1) & (2 * N - 1), sm.e2n_hi, sm.e2n_lo, lane);
        cplx accj[R];
#pragma unroll
        for (int d = 0; d < R; d++) accj[d] = cplx{0.0, 0.0};
#pragma unroll 1
        for (int q = 0; q < L; q++) {
            {
                cplx z[R];
#pragma unroll
                for (int r = 0; r < R; r++)
                    z[r] = cplx{static_cast<double>(Gadget<P>::digit(acc[r], q)),
                                static_cast<double>(Gadget<P>::digit(acc[R + r], q))};
                fft_forward<H, CUFHE_KBS_SWZ != 0, true>(z, my_spec, tb);
#pragma unroll
                for (int d = 0; d < R; d++) lds_store(&my_spec[d * 64 + lane], z[d]);
            }
            x0g.pin();
            x1g.pin();
            x01g.pin();
            const double hq = ldexp(1.0, -static_cast<int>((q + 1) * P::Bgbit));
            const int g0 = (it * L + q) * R;
#pragma unroll
            for (int d = 0; d < R; d++) {
                const int g = g0 + d;
                __syncthreads();  // chunk g in the ring; chunk g-1 consumed; spectra written
                const cplx (*ring)[64] = sm.ring[g & 1];
                const cplx x0 = x0g.at(d), x1 = x1g.at(d), x01 = x01g.at(d);
                cplx sacc = accj[d];
#pragma unroll
                for (int comp = 0; comp < 2; comp++) {
                    const cplx* spec = comp == j ? my_spec : partner_spec;
                    const cplx D = lds_load(&spec[d * 64 + lane]);
                    const int s0 = (j * 2 + comp) * 3;
                    cplx k = {comp == j ? hq : 0.0, 0.0};
                    cmac(k, lds_load(&ring[s0 + 2][lane]), x1);
                    cmac(k, lds_load(&ring[s0 + 1][lane]), x0);
                    cmac(k, lds_load(&ring[s0 + 0][lane]), x01);
                    cmac(sacc, D, k);
                }
                // finish this chunk's arithmetic before the next barrier: the
                // compiler would otherwise sink it past all eight barriers and
                // keep every chunk's LDS operands alive (spilled) until then
                asm volatile("" : "+v"(sacc.x), "+v"(sacc.y));
                accj[d] = sacc;
                // refill: chunk g+1 into the other buffer, fetch chunk g+1+PD
                store_chunk(g + 1, kv[(d + 1) % PD]);
                load_chunk(g + 1 + PD, kv[(d + 1) % PD]);
            }
            __syncthreads();  // partner done reading this wave's spectrum
        }
        fft_inverse<H, CUFHE_KBS_SWZ != 0, true>(accj, my_spec, tb);
        constexpr double scale = kTorusScale<uint32_t> / H;
#pragma unroll
        for (int r = 0; r < R; r++) {
            acc[r] = double_to_torus<uint32_t>(accj[r].x * scale);
            acc[R + r] = double_to_torus<uint32_t>(accj[r].y * scale);
        }
    }
}

}  // namespace cufhe_amd
