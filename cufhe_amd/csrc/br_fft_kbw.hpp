// Key-bundle blind rotation at lvl1 with ONE wave per job and the key shared
// by the C jobs of a workgroup (the batched form's main kernel at large
// batches).
//
// Arithmetic: exactly br_fft.hpp's key-bundle step (reference
// include/gatebootstrapping_gpu.cuh:895-1033 AccumulateKeyBundle and
// :1335-1429 the drivers), operation for operation, so the outputs are
// bit-identical to k_br_fft / k_br_fft_kbs and the plain per-gate kernels.
//
// Organisation (versus br_fft_kbs.hpp, two waves per job):
//  * a wave holds both TRLWE components of its job's accumulator, transforms
//    both digit polynomials of a gadget digit as one interleaved pair
//    (fft_forward2: the two dependency chains overlap) and keeps the
//    spectra itself (component 0 in VGPRs, component 1 in its idle transform
//    scratch), so no wave waits for a partner's spectrum;
//  * the multiply-accumulate produces both output columns (16 complex
//    accumulators) and the two inverse transforms again run as a pair;
//  * C = 8 jobs per workgroup at 2 waves/SIMD (256 VGPRs per wave): a batch
//    of 4096 is exactly two rounds of workgroups on 256 CUs;
//  * the key of one (step, digit, register pair) -- 2 registers x 2 columns x
//    2 components x 3 bundle members x 1 KiB = 24 KiB, a chunk of 24 slices
//    -- is fetched once per workgroup (wave w loads slices w, w+8, w+16 one
//    chunk ahead into VGPRs), written into a two-chunk LDS ring and read by
//    every wave: one workgroup barrier per chunk, 8 per step, no other
//    barrier in the loop; within a chunk the (register, column) blocks are
//    software pipelined (the next block's six slices are loaded while the
//    current block is multiplied).
#pragma once

#include "br_fft_kbs.hpp"

namespace cufhe_amd {

#ifndef CUFHE_KBW_PD
#define CUFHE_KBW_PD 1  // key chunks in flight per wave
#endif

// LDS of one workgroup of C jobs. The arrays every wave addresses at fixed
// offsets (key ring, tables) come first, below 64 KiB, so one lane-address
// register plus the 16-bit instruction offset reaches all of them.
// Registers of the key chunk: 2 for 8 jobs per workgroup (24 slices, 3 per
// wave), 1 for 4 (12 slices, 3 per wave; two workgroups share a CU).
template <int C>
constexpr int kKbwRpc = C >= 8 ? 2 : 1;

template <int C>
struct KbwShared {
    static constexpr int N = lvl1param::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    cplx ring[2][kKbwRpc<C>][12][64];  // key chunks: [buffer][register][slice][bin]
    cplx twl[64];         // transform tables (copies of FFTTables; row 0 of twl)
    cplx tws[64];
    cplx tw64[64];
    cplx cr[R];
    cplx e2n_hi[2 * N / 64];  // exp(i pi 64 t / N)
    cplx e2n_lo[64];          // exp(i pi t / N)
    cplx scratch[C][fft_stage::kScratchPad(H)];  // per wave: transform scratch (padded rows)
    uint16_t abar[C][n0 + 4];  // mod-switched lvl0 mask
    uint32_t bar[C];
};

// The workgroup's key pipeline as seen by one wave: the slices it loads of
// every key chunk (t = slot, slot + C, ...; chunk g = (step * l + digit) *
// (R / RPC) + register group, slice t = (register, (column * 2 + component)
// * 3 + bundle member)), held in VGPRs PD chunks ahead and written into the
// two-chunk LDS ring one chunk ahead.
template <int C>
struct KbwKeyPipe {
    using P = lvl1param;
    static constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l, n0 = lvl0param::n;
    static constexpr int W = C;
    static constexpr int RPC = kKbwRpc<C>;
    static constexpr int NS = 12 * RPC;
    static constexpr int SPW = (NS + W - 1) / W;
    static constexpr int PD = CUFHE_KBW_PD;
    static constexpr int CPD = R / RPC;
    static constexpr int STEPS = n0 / 2;
    static constexpr int CHUNKS = STEPS * L * CPD;
    KbwShared<C>& sm;
    const int slot, lane;
    const cplx* bk;
    cplx kv[PD][SPW];
    __device__ __forceinline__ KbwKeyPipe(KbwShared<C>& s, int sl, int ln, const cplx* k)
        : sm(s), slot(sl), lane(ln), bk(k)
    {
    }
    __device__ __forceinline__ void load(int g, cplx (&dst)[SPW])
    {
        const int pair = g % CPD, q = (g / CPD) % L, it = g / (CPD * L);
#pragma unroll
        for (int k = 0; k < SPW; k++) {
            const int t = slot + k * W;
            if ((SPW * W == NS || t < NS) && g < CHUNKS)
                dst[k] = *kbs_slice_ptr(bk, it, q, RPC * pair + t / 12, t % 12, lane);
        }
    }
    __device__ __forceinline__ void store(int g, const cplx (&src)[SPW])
    {
#pragma unroll
        for (int k = 0; k < SPW; k++) {
            const int t = slot + k * W;
            if ((SPW * W == NS || t < NS) && g < CHUNKS)
                lds_store(&sm.ring[g & 1][t / 12][t % 12][lane], src[k]);
        }
    }
    // before the first barrier: chunk 0 in the ring, chunks 1..PD in flight
    __device__ __forceinline__ void start()
    {
#pragma unroll
        for (int p = 0; p < PD; p++) load(p, kv[p]);
        store(0, kv[0]);
        load(PD, kv[0]);
    }
    // after the barrier of chunk g (pair = g % CPD): chunk g+1 into the other
    // buffer, fetch chunk g+1+PD
    __device__ __forceinline__ void advance(int g, int pair)
    {
        store(g + 1, kv[(pair + 1) % PD]);
        load(g + 1 + PD, kv[(pair + 1) % PD]);
    }
    // A wave without a job: the same loads, stores and barriers as a
    // rotating wave (one barrier per chunk, nothing else), so a workgroup can
    // run fewer jobs than waves at the speed of its rotating waves.
    __device__ __forceinline__ void run_loader_only()
    {
        start();
#pragma unroll 1
        for (int gq = 0; gq < STEPS * L; gq++) {
#pragma unroll
            for (int pair = 0; pair < CPD; pair++) {
                __syncthreads();  // the rotating waves' chunk barrier
                advance(gq * CPD + pair, pair);
            }
        }
    }
};

// The blind rotation of one job (wave `slot` of the workgroup); acc[j] holds
// component j in AccStore's register order.
template <int C>
__device__ __forceinline__ void blind_rotate_kbw(KbwShared<C>& sm, int slot, uint32_t mu,
                                                 uint32_t (&acc)[2][16], const cplx* bk)
{
    using P = lvl1param;
    constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l, n0 = lvl0param::n;
    constexpr int RPC = kKbwRpc<C>;          // registers per chunk
    constexpr int STEPS = n0 / 2;
    constexpr int CPD = R / RPC;             // chunks per digit pass (g = (it * L + q) * CPD + pair)
    static_assert(R == 8 && P::Bgbit == 8 && L == 2, "lvl1 shapes");
    const int lane = threadIdx.x & 63;
    cplx* const scratch = sm.scratch[slot];
    // Transform tables behind an offset the compiler cannot see through, so
    // their LDS loads stay inside the loop instead of being hoisted out of it
    // (60 VGPRs of twiddles held across the step would spill).
    auto tables = [&]() {
        int zo = 0;
        asm volatile("" : "+s"(zo));
        return FFTTables{sm.twl + zo, sm.tw64 + zo, sm.cr + zo, sm.tws + zo};
    };

    // test vector
    const uint32_t bar = sm.bar[slot];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = AccStore<P>::coef_index(r, lane);
        acc[0][r] = 0u;
        acc[1][r] = test_vector_coef<P>(i, bar, mu);
    }

    KbwKeyPipe<C> pipe(sm, slot, lane, bk);
    pipe.start();

#pragma unroll 1
    for (int it = 0; it < STEPS; it++) {
        const uint32_t a0 = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it]);
        const uint32_t a1 = __builtin_amdgcn_readfirstlane(sm.abar[slot][2 * it + 1]);
        cplx accj[2][R];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int d = 0; d < R; d++) accj[j][d] = cplx{0.0, 0.0};

        auto digit = [](uint32_t byte) { return static_cast<double>(static_cast<int32_t>(byte) - 128); };

        // Digits are bytes of c + kOffset (Bgbit = 8, l = 2): digit 0 is byte
        // 3, digit 1 byte 2. Keep bytes 3..2 of coefficients r and R + r of a
        // component in one register (8 per component instead of 16).
        uint32_t dg[2][R];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t lo = acc[j][r] + Gadget<P>::kOffset;
                const uint32_t hi = acc[j][R + r] + Gadget<P>::kOffset;
                dg[j][r] = (lo >> 16) | (hi & 0xffff0000u);
            }
#pragma unroll 1
        for (int q = 0; q < L; q++) {
            const uint32_t sh = 8 * (1 - q);  // digit q of the low half
            cplx z0[R], z1[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                z0[r] = cplx{digit((dg[0][r] >> sh) & 255u), digit((dg[0][r] >> (sh + 16)) & 255u)};
                z1[r] = cplx{digit((dg[1][r] >> sh) & 255u), digit((dg[1][r] >> (sh + 16)) & 255u)};
            }
            fft_forward2<H, 2, true>(z0, z1, scratch, tables());
            // component 1's spectrum waits in the (now idle) transform
            // scratch: 32 fewer VGPRs held through the multiply-accumulate
#pragma unroll
            for (int d = 0; d < R; d++) lds_store(&scratch[d * 64 + lane], z1[d]);
            XaiGenShared x0g, x1g, x01g;
            x0g.init(a0, sm.e2n_hi, sm.e2n_lo, lane);
            x1g.init(a1, sm.e2n_hi, sm.e2n_lo, lane);
            x01g.init((a0 + a1) & (2 * N - 1), sm.e2n_hi, sm.e2n_lo, lane);
            const double hq = ldexp(1.0, -static_cast<int>((q + 1) * P::Bgbit));
            const int g0 = (it * L + q) * CPD;
#pragma unroll
            for (int pair = 0; pair < CPD; pair++) {
                const int g = g0 + pair;
                __syncthreads();  // chunk g in the ring; chunk g-1 consumed
                const cplx (*ring)[12][64] = sm.ring[g & 1];
                // W_8^(a d) of this chunk's registers, fetched (scalar loads)
                // before any LDS load of the chunk: scalar loads complete out
                // of order, so a wait for them would also drain every LDS
                // load in flight
                cplx w8v[RPC][3];
#pragma unroll
                for (int dd = 0; dd < RPC; dd++) {
                    const uint32_t d = RPC * pair + dd;
                    const uint32_t m[3] = {(x0g.a * d) & 7u, (x1g.a * d) & 7u, (x01g.a * d) & 7u};
#pragma unroll
                    for (int e = 0; e < 3; e++) w8v[dd][e] = cplx{kW8[m[e]][0], kW8[m[e]][1]};
                }
#pragma unroll
                for (int dd = 0; dd < RPC; dd++)
#pragma unroll
                    for (int e = 0; e < 3; e++) asm volatile("" : "+s"(w8v[dd][e].x), "+s"(w8v[dd][e].y));
                // (register, column) blocks of 6 key slices, software
                // pipelined: the next block's LDS loads are issued before the
                // current block's arithmetic, so LDS and ALUs overlap
                constexpr int NB = RPC * 2;
                cplx kb[2][6];
                cplx D1[RPC];
                auto load_block = [&](int blk, cplx (&dst)[6]) {
                    const int dd = blk >> 1, col = blk & 1;
                    if (col == 0) D1[dd] = lds_load(&scratch[(RPC * pair + dd) * 64 + lane]);
#pragma unroll
                    for (int t = 0; t < 6; t++) dst[t] = lds_load(&ring[dd][col * 6 + t][lane]);
                };
                auto xai = [](const cplx& base, const cplx& w) {
                    const cplx v = cmul(base, w);  // as XaiGenShared::at
                    return cplx{v.x - 1.0, v.y};
                };
                load_block(0, kb[0]);
                cplx x0, x1, x01;
#pragma unroll
                for (int blk = 0; blk < NB; blk++) {
                    const int dd = blk >> 1, col = blk & 1;
                    const int d = RPC * pair + dd;
                    if (blk + 1 < NB) load_block(blk + 1, kb[(blk + 1) & 1]);
                    if (col == 0) {
                        x0 = xai(x0g.base, w8v[dd][0]);
                        x1 = xai(x1g.base, w8v[dd][1]);
                        x01 = xai(x01g.base, w8v[dd][2]);
                    }
                    const cplx (&kr)[6] = kb[blk & 1];
                    cplx sacc = accj[col][d];
#pragma unroll
                    for (int comp = 0; comp < 2; comp++) {
                        const cplx D = comp == 0 ? z0[d] : D1[dd];
                        const int s0 = comp * 3;
                        cplx k = {comp == col ? hq : 0.0, 0.0};
                        cmac(k, kr[s0 + 2], x1);
                        cmac(k, kr[s0 + 1], x0);
                        cmac(k, kr[s0 + 0], x01);
                        cmac(sacc, D, k);
                    }
                    // keep this chunk's arithmetic ahead of the next barrier
                    asm volatile("" : "+v"(sacc.x), "+v"(sacc.y));
                    accj[col][d] = sacc;
                    // the scheduler may not pull a later block's loads
                    // ahead of this block's arithmetic
                    __builtin_amdgcn_sched_barrier(0);
                }
                pipe.advance(g, pair);
            }
        }
        fft_inverse2<H, 2, true>(accj[0], accj[1], scratch, tables());
        constexpr double scale = kTorusScale<uint32_t> / H;
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < R; r++) {
                acc[j][r] = double_to_torus<uint32_t>(accj[j][r].x * scale);
                acc[j][R + r] = double_to_torus<uint32_t>(accj[j][r].y * scale);
            }
    }
}

}  // namespace cufhe_amd
