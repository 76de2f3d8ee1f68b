// Batched identity key switching on the matrix cores: a hand-written int8
// MFMA kernel (v_mfma_i32_32x32x32_i8) with the digit matrix expanded in
// registers and the byte planes recombined in the epilogue.
//
// Reference semantics (include/keyswitch_gpu.cuh:189-256,
// IdentityKeySwitchPreAdd; :128-187 KeySwitchFromTLWE): for every job and
// output coefficient i,
//   res_i = base_i - sum_{j, k} sgn(v_jk) * KSK[j][k][|v_jk| - 1][i]
// with v_jk the k-th signed base-2^basebit digit of ca*in0[j] + cb*in1[j]
// (+ offsets). For a batch of M jobs this is a matrix product
//   RES = BASE + S x KSK,   S[job][(j, k, h)] = -sgn(v_jk) if |v_jk| = h + 1 else 0,
// S in {-1, 0, +1} (M x K, K = switched coefficients x t x 2^(basebit-1)).
// Every u32 KSK word w is written as four balanced bytes,
//   w = sum_b sb_b 2^(8b)  (mod 2^32),  sb_b in [-128, 127],
// so sum_r S[r] w_r = sum_b 2^(8b) G_b (mod 2^32) with G_b = sum_r S[r] sb_b(r)
// an int8 x int8 -> int32 product. |G_b| <= K * 128 < 2^22 (K <= 28,672),
// so the int32 accumulation is exact and RES is bit-identical to the gather
// form (key_switch_waves / k_ks_batch).
//
// Two launches per batch:
//  1. k_ks_codes: the digits of every job as 2-bit codes (code = v mod 4),
//     16 codes per u32 word, one word per 32-deep MFMA step, stored
//     [step / 4][job][step % 4] so a lane loads four steps at once; the same
//     kernel writes each job's base terms into its output TLWE.
//  2. k_ks_mfma: one wave owns MM blocks of 32 jobs x 32 output coefficients
//     x the four byte planes (4 MM accumulators of 32x32 int32) over a range
//     of steps. Per step it expands its code words to the S fragment (two
//     shifts/masks and four v_perm byte lookups per 16 bytes) and loads the
//     four plane fragments of the key, which is stored pre-tiled in the MFMA
//     register order (1 KiB per fragment, one dwordx4 per lane). The
//     epilogue forms sum_b G_b << 8b per coefficient in registers and adds
//     it to the output (atomically when the steps are split over waves).
//
// The pairing of S and key bytes only requires that the A and B operands of
// the MFMA use the same lane/byte -> k assignment; which k each byte is
// does not matter for the sum, so both sides follow one logical order:
// lane half g, byte e = 4 part + q  <->  within-step pair i = 4q + 2g + (part >> 1),
// digit row h = part & 1.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fft_wave.hpp"  // wave_lds_sync
#include "keyswitch.hpp"
#include "ks_gemm.hpp"

namespace cufhe_amd {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kOutCols = lvl0param::n + 1;     // 637 output words
constexpr int kCoefBlocks = (kOutCols + 31) / 32;  // 20 blocks of 32
constexpr int kFragBytes = 64 * 16;            // one MFMA operand fragment

// Static shape of one key type.
template <class KS, bool SUBSET>
struct KsShape {
    using TR = KsTraits<KS>;
    static constexpr uint32_t first = SUBSET ? TR::n0 : 0;
    static constexpr uint32_t switched = TR::dn - first;          // 1024, 388, 2048
    static constexpr uint32_t pairs = switched * KS::t;           // (j, k) digit pairs
    static constexpr uint32_t steps = ((pairs + 15) / 16 + 7) / 8 * 8;  // 32-deep MFMA steps (x8)
    static_assert(KS::basebit == 2 && TR::numbase == 2, "2-bit digits with two magnitudes");
    static_assert(KS::t == 7, "64 coefficients give exactly 28 code words");
    static_assert((switched + 63) / 64 * 28 >= steps, "every code word is written by some wave");
};

// Balanced byte b of w: w = sum_b sb_b 2^(8b) mod 2^32, sb_b in [-128, 127].
__device__ __forceinline__ int8_t balanced_byte(uint32_t w, int b)
{
    int64_t x = w;
    for (int i = 0; i < b; i++) {
        const int64_t sb = static_cast<int8_t>(static_cast<uint8_t>(x & 0xff));
        x = (x - sb) >> 8;
    }
    return static_cast<int8_t>(static_cast<uint8_t>(x & 0xff));
}

// Pre-tiled key: frag[cb][step][plane][lane][16 bytes], one dword per thread.
template <class KS, bool SUBSET>
__global__ __launch_bounds__(256) void k_ksk_frags(const uint32_t* __restrict__ ksk,
                                                    uint32_t* __restrict__ frags)
{
    using SH = KsShape<KS, SUBSET>;
    constexpr size_t total = static_cast<size_t>(kCoefBlocks) * SH::steps * 4 * 64 * 4;
    for (size_t x = blockIdx.x * 256ull + threadIdx.x; x < total; x += gridDim.x * 256ull) {
        const int part = static_cast<int>(x & 3);
        const int lane = static_cast<int>((x >> 2) & 63);
        const int plane = static_cast<int>((x >> 8) & 3);
        const size_t rest = x >> 10;
        const uint32_t s = static_cast<uint32_t>(rest % SH::steps);
        const int cb = static_cast<int>(rest / SH::steps);
        const int col = cb * 32 + (lane & 31);
        const int g = lane >> 5;
        uint32_t word = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t p = 16 * s + 4 * q + 2 * g + (part >> 1);
            const int h = part & 1;
            uint8_t byte = 0;
            if (p < SH::pairs && col < kOutCols) {
                const size_t row = (static_cast<size_t>(p) * KsTraits<KS>::numbase + h);
                byte = static_cast<uint8_t>(balanced_byte(ksk[row * kOutCols + col], plane));
            }
            word |= static_cast<uint32_t>(byte) << (8 * q);
        }
        frags[x] = word;
    }
}

// Code words and base terms. A wave takes 64 consecutive switched
// coefficients of one job (loads coalesced over j), each lane forming the 7
// codes (14 bits) of its coefficient; the bits are or-ed into the wave's 28
// words in LDS and stored by 28 lanes. The last wave of each job writes the
// job's base terms into its output TLWE. 4 waves per workgroup, wave id =
// job * (waves per job) + part.
template <class KS, bool SUBSET>
__global__ __launch_bounds__(256) void k_ks_codes(const KsJob* __restrict__ jobs, int njobs,
                                                   uint32_t* __restrict__ codes)
{
    using SH = KsShape<KS, SUBSET>;
    using TR = KsTraits<KS>;
    using DT = typename TR::DT;
    constexpr int parts = static_cast<int>((SH::switched + 63) / 64);  // 16, 7, 32
    constexpr int words = 64 * 7 / 16;                                // 28 per wave
    __shared__ uint32_t wbuf[4][words];
    const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
    const long wid = static_cast<long>(blockIdx.x) * 4 + wl;
    if (wid >= static_cast<long>(njobs) * (parts + 1)) return;  // whole wave; no barriers below
    const int job = static_cast<int>(wid / (parts + 1));
    const int part = static_cast<int>(wid % (parts + 1));
    const KsJob jb = jobs[job];
    const DT* in0 = static_cast<const DT*>(jb.in0);
    const DT* in1 = static_cast<const DT*>(jb.in1);
    const DT ua = static_cast<DT>(static_cast<std::make_signed_t<DT>>(jb.ca));
    const DT ub = static_cast<DT>(static_cast<std::make_signed_t<DT>>(jb.cb));
    if (part == parts) {
        for (int i = lane; i < kOutCols; i += 64) {
            uint32_t r = 0;
            if (i == static_cast<int>(TR::n0)) {
                DT v = ua * in0[TR::dn] + static_cast<DT>(jb.offset);
                if (in1) v += ub * in1[TR::dn];
                r = TR::convert(v);
            }
            else if (SUBSET) {
                DT v = ua * in0[i];
                if (in1) v += ub * in1[i];
                r = TR::convert(v);
            }
            jb.out[i] = r;
        }
        return;
    }
    uint32_t* wb = wbuf[wl];
    if (lane < words) wb[lane] = 0;
    wave_lds_sync();
    const uint32_t jr = static_cast<uint32_t>(part) * 64 + lane;  // switched coefficient
    if (jr < SH::switched) {
        const uint32_t j = jr + SH::first;
        DT tmp = ua * in0[j] + TR::kOffset;
        if (in1) tmp += ub * in1[j];
        uint32_t bits = 0;  // the 7 codes, code = (digit - 2) mod 4
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const uint32_t digit =
                static_cast<uint32_t>(tmp >> (TR::W - (k + 1) * KS::basebit)) & 3u;
            bits |= (digit ^ 2u) << (2 * k);
        }
        const uint32_t pos = static_cast<uint32_t>(lane) * 14;  // bit offset in the wave's words
        atomicOr(&wb[pos >> 5], bits << (pos & 31));
        if ((pos & 31) > 18) atomicOr(&wb[(pos >> 5) + 1], bits >> (32 - (pos & 31)));
    }
    wave_lds_sync();
    if (lane < words) {
        const uint32_t s = static_cast<uint32_t>(part) * words + lane;
        if (s < SH::steps) codes[(static_cast<size_t>(s >> 2) * njobs + job) * 4 + (s & 3)] = wb[lane];
    }
}

// S fragment of one 32-deep step for lane half g from the step's code word.
__device__ __forceinline__ v4i expand_codes(uint32_t w, int g)
{
    constexpr uint32_t kLutH0 = 0x0100ff00u;  // code 1 -> -1, code 3 -> +1
    constexpr uint32_t kLutH1 = 0x00010000u;  // code 2 -> +1
    const uint32_t x0 = (w >> (4 * g)) & 0x03030303u;
    const uint32_t x1 = (w >> (4 * g + 2)) & 0x03030303u;
    v4i a;
    a[0] = static_cast<int>(__builtin_amdgcn_perm(kLutH0, kLutH0, x0));
    a[1] = static_cast<int>(__builtin_amdgcn_perm(kLutH1, kLutH1, x0));
    a[2] = static_cast<int>(__builtin_amdgcn_perm(kLutH0, kLutH0, x1));
    a[3] = static_cast<int>(__builtin_amdgcn_perm(kLutH1, kLutH1, x1));
    return a;
}

// Tiles: t = (coefficient block cb, step split z), T = kCoefBlocks x split
// of them, each covering all jobs in mgroups workgroups of WPB waves (MM x
// 32 jobs per wave). Workgroups are dealt to the 8 XCDs round-robin by
// linear id, so the id is decoded to put every workgroup of one tile on the
// same XCD, consecutively: that XCD's L2 then fetches the tile's key
// fragments once for all jobs. Ids past the last tile exit at once.
// Outputs hold the base terms on entry (k_ks_codes); each wave adds its
// share.
//
// Key fragments and code words go through LDS in chunks of CS steps (4 CS
// fragments): each wave loads its share into registers at the start of a
// chunk, runs the chunk's MFMAs from the other LDS buffer, then stores the
// loads into the idle buffer; one workgroup barrier per chunk. Waves past
// the last job still load and keep step with the barriers, without MFMAs or
// stores.
template <class KS, bool SUBSET, int MM, int WPB, int CS>
__global__ __launch_bounds__(64 * WPB) void k_ks_mfma(const KsJob* __restrict__ jobs, int njobs,
                                                       const uint32_t* __restrict__ codes,
                                                       const v4i* __restrict__ frags, int split,
                                                       int mgroups)
{
    using SH = KsShape<KS, SUBSET>;
    constexpr int S = SH::steps;
    constexpr int NF = 4 * CS;       // fragments per chunk
    constexpr int FW = NF / WPB;     // fragments each wave moves per chunk
    constexpr int CQ = CS / 4;       // uint4 code words per job block per chunk
    static_assert(NF % WPB == 0 && CS % 4 == 0 && S % CS == 0, "chunk shape");
    __shared__ v4i chunk[2][NF][64];                // [buffer][fragment = 4 step + plane][lane]
    __shared__ uint4 cwords[2][WPB][MM * CQ][64];  // each wave's own code words
    const int lane = threadIdx.x & 63, g = lane >> 5, wl = threadIdx.x >> 6;
    const int xcd = static_cast<int>(blockIdx.x & 7);
    const int r = static_cast<int>(blockIdx.x >> 3);
    const int mt = r % mgroups;
    const int t = (r / mgroups) * 8 + xcd;
    if (t >= kCoefBlocks * split) return;  // whole workgroup
    const int cb = t % kCoefBlocks, z = t / kCoefBlocks;
    const int job0 = (mt * WPB + wl) * MM * 32;
    const bool active = job0 < njobs;  // wave-uniform
    // chunks of this split
    const int c0 = (S / CS) * z / split;
    const int c1 = (S / CS) * (z + 1) / split;

    v16i acc[MM][4];
#pragma unroll
    for (int m = 0; m < MM; m++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[m][b] = v16i{};

    const v4i* fb = frags + (static_cast<size_t>(cb) * S * 4) * 64 + lane;
    const uint4* cw = reinterpret_cast<const uint4*>(codes);
    // rows past the batch read the last job's codes and use zero codes
    int rowjob[MM];
    bool live[MM];
#pragma unroll
    for (int m = 0; m < MM; m++) {
        const int j = job0 + m * 32 + (lane & 31);
        live[m] = j < njobs;
        rowjob[m] = live[m] ? j : njobs - 1;
    }
    auto load = [&](int c, v4i (&rg)[FW], uint4 (&w)[MM * CQ]) {
#pragma unroll
        for (int i = 0; i < FW; i++) rg[i] = fb[static_cast<size_t>(c * NF + wl * FW + i) * 64];
#pragma unroll
        for (int m = 0; m < MM; m++)
#pragma unroll
            for (int k = 0; k < CQ; k++)
                w[m * CQ + k] = cw[static_cast<size_t>(c * CQ + k) * njobs + rowjob[m]];
    };
    auto store = [&](int buf, const v4i (&rg)[FW], const uint4 (&w)[MM * CQ]) {
#pragma unroll
        for (int i = 0; i < FW; i++) chunk[buf][wl * FW + i][lane] = rg[i];
#pragma unroll
        for (int i = 0; i < MM * CQ; i++) cwords[buf][wl][i][lane] = w[i];
    };
    // One chunk's MFMAs from LDS buffer buf: the four plane fragments of
    // step s + 1 are read while step s's MFMAs run.
    auto compute = [&](int buf) {
        v4i bq[2][4];
#pragma unroll
        for (int b = 0; b < 4; b++) bq[0][b] = chunk[buf][b][lane];
#pragma unroll
        for (int k = 0; k < CQ; k++) {
            uint4 wc[MM];
#pragma unroll
            for (int m = 0; m < MM; m++) wc[m] = cwords[buf][wl][m * CQ + k][lane];
#pragma unroll
            for (int sub = 0; sub < 4; sub++) {
                const int st = k * 4 + sub;
                if (st + 1 < CS) {
#pragma unroll
                    for (int b = 0; b < 4; b++) bq[(st + 1) & 1][b] = chunk[buf][(st + 1) * 4 + b][lane];
                }
                v4i a[MM];
#pragma unroll
                for (int m = 0; m < MM; m++) {
                    const uint32_t w = sub == 0 ? wc[m].x : sub == 1 ? wc[m].y : sub == 2 ? wc[m].z : wc[m].w;
                    a[m] = expand_codes(live[m] ? w : 0u, g);
                }
                // keep the next step's LDS reads ahead of this step's MFMAs
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int m = 0; m < MM; m++)
                        acc[m][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[m], bq[st & 1][b], acc[m][b], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };

    if (c0 < c1) {
        v4i rg[FW];
        uint4 wn[MM * CQ];
        load(c0, rg, wn);
        store(0, rg, wn);
        __syncthreads();
#pragma unroll 1
        for (int c = c0; c < c1; c++) {
            const int buf = (c - c0) & 1;
            const bool more = c + 1 < c1;  // workgroup-uniform
            if (more) load(c + 1, rg, wn);
            __builtin_amdgcn_sched_barrier(0);
            if (active) compute(buf);
            __builtin_amdgcn_sched_barrier(0);
            if (more) {
                // an empty statement that claims to update the loaded
                // registers here: the stores (and their wait on the loads)
                // cannot be hoisted above this chunk's MFMAs
#pragma unroll
                for (int i = 0; i < FW; i++) asm volatile("" : "+v"(rg[i]));
#pragma unroll
                for (int i = 0; i < MM * CQ; i++)
                    asm volatile("" : "+v"(wn[i].x), "+v"(wn[i].y), "+v"(wn[i].z), "+v"(wn[i].w));
                store(buf ^ 1, rg, wn);
            }
            __syncthreads();
        }
    }
    if (!active) return;

    // 32x32 accumulator layout: lane holds column lane & 31, rows
    // 8 (r / 4) + 4 g + (r % 4) for r < 16. Each lane loads the output
    // pointer of job row lane & 31 of every block once; row r's pointer is
    // then read from lane `row`. All reads are issued before any write.
    const int col = cb * 32 + (lane & 31);
    const bool colok = col < kOutCols;
    uint32_t plo[MM], phi[MM];
#pragma unroll
    for (int m = 0; m < MM; m++) {
        const int j = job0 + m * 32 + (lane & 31);
        const uint64_t p = j < njobs ? reinterpret_cast<uint64_t>(jobs[j].out) : 0;
        plo[m] = static_cast<uint32_t>(p);
        phi[m] = static_cast<uint32_t>(p >> 32);
    }
    uint32_t* dst[MM][16];
    uint32_t val[MM][16];
#pragma unroll
    for (int m = 0; m < MM; m++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * g + (r & 3);
            const uint64_t p = (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(phi[m]), row))) << 32) |
                               static_cast<uint32_t>(__shfl(static_cast<int>(plo[m]), row));
            dst[m][r] = (p != 0 && colok) ? reinterpret_cast<uint32_t*>(p) + col : nullptr;
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) v += static_cast<uint32_t>(acc[m][b][r]) << (8 * b);
            val[m][r] = v;
        }
    if (split == 1) {
        uint32_t old[MM][16];
#pragma unroll
        for (int m = 0; m < MM; m++)
#pragma unroll
            for (int r = 0; r < 16; r++) old[m][r] = dst[m][r] ? *dst[m][r] : 0u;
#pragma unroll
        for (int m = 0; m < MM; m++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                if (dst[m][r]) *dst[m][r] = old[m][r] + val[m][r];
    }
    else {
#pragma unroll
        for (int m = 0; m < MM; m++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                if (dst[m][r]) atomicAdd(dst[m][r], val[m][r]);
    }
}

template <class KS, bool SUBSET>
constexpr int ks_rows()
{
    return static_cast<int>(KsShape<KS, SUBSET>::pairs * KsTraits<KS>::numbase);
}

template <class KS, bool SUBSET>
size_t frag_bytes()
{
    return static_cast<size_t>(kCoefBlocks) * KsShape<KS, SUBSET>::steps * 4 * kFragBytes;
}

int g_split_override = 0;  // 0: pick per batch

template <class KS, bool SUBSET, int MM>
int run_shape(const KsGemmKey& key, const KsJob* d_jobs, int njobs, void* ws, hipStream_t st)
{
    using SH = KsShape<KS, SUBSET>;
    uint32_t* codes = static_cast<uint32_t*>(ws);
    constexpr int parts = static_cast<int>((SH::switched + 63) / 64);
    const size_t waves = static_cast<size_t>(parts + 1) * njobs;
    hipLaunchKernelGGL((k_ks_codes<KS, SUBSET>), dim3(static_cast<unsigned>((waves + 3) / 4)),
                       dim3(256), 0, st, d_jobs, njobs, codes);
    const int waves_m = (njobs + MM * 32 - 1) / (MM * 32);
    // about 3.5 waves per SIMD of the 256-CU chip (measured best for 4096
    // jobs: split 3, 3840 waves; the split's partial sums cost atomics)
    int split = g_split_override;
    if (split <= 0) {
        const int waves = waves_m * kCoefBlocks;
        split = (3584 + waves / 2) / waves;
        if (split < 1) split = 1;
        if (split > 16) split = 16;
    }
    if (split > static_cast<int>(SH::steps / kKsMfmaChunk)) split = SH::steps / kKsMfmaChunk;
    const int mgroups = (waves_m + kKsMfmaWaves - 1) / kKsMfmaWaves;
    const int tiles = kCoefBlocks * split;
    const unsigned grid = static_cast<unsigned>(8 * mgroups * ((tiles + 7) / 8));
    hipLaunchKernelGGL((k_ks_mfma<KS, SUBSET, MM, kKsMfmaWaves, kKsMfmaChunk>), dim3(grid),
                       dim3(64 * kKsMfmaWaves),
                       0, st, d_jobs, njobs, codes, reinterpret_cast<const v4i*>(key.planes), split,
                       mgroups);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

static_assert(KsShape<TFHEpp::lvl10param, false>::steps == kKsLvl10StepsMax &&
                  KsShape<TFHEpp::lvl10param, true>::steps <= kKsLvl10StepsMax,
              "lvl1 key switch workspace bound");

void ks_gemm_set_split(int split) { g_split_override = split; }

int ks_gemm_prepare(KsGemmKey& key, const uint32_t* d_ksk, int domain_level, bool subset,
                    hipStream_t st)
{
    size_t bytes;
    if (domain_level == 2)
        bytes = frag_bytes<TFHEpp::lvl20param, false>();
    else
        bytes = subset ? frag_bytes<TFHEpp::lvl10param, true>() : frag_bytes<TFHEpp::lvl10param, false>();
    if (hipMalloc(reinterpret_cast<void**>(&key.planes), bytes) != hipSuccess) {
        key.planes = nullptr;
        return -1;
    }
    key.rows = domain_level == 2 ? ks_rows<TFHEpp::lvl20param, false>()
                                 : (subset ? ks_rows<TFHEpp::lvl10param, true>()
                                           : ks_rows<TFHEpp::lvl10param, false>());
    key.cols = 4 * kOutCols;
    key.level = domain_level;
    key.subset = subset;
    uint32_t* f = reinterpret_cast<uint32_t*>(key.planes);
    if (domain_level == 2)
        hipLaunchKernelGGL((k_ksk_frags<TFHEpp::lvl20param, false>), dim3(8192), dim3(256), 0, st,
                           d_ksk, f);
    else if (subset)
        hipLaunchKernelGGL((k_ksk_frags<TFHEpp::lvl10param, true>), dim3(8192), dim3(256), 0, st,
                           d_ksk, f);
    else
        hipLaunchKernelGGL((k_ksk_frags<TFHEpp::lvl10param, false>), dim3(8192), dim3(256), 0, st,
                           d_ksk, f);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void ks_gemm_release(KsGemmKey& key)
{
    (void)hipFree(key.planes);
    key = KsGemmKey{};
}

size_t ks_gemm_workspace_bytes(const KsGemmKey& key, int njobs)
{
    uint32_t steps;
    if (key.level == 2)
        steps = KsShape<TFHEpp::lvl20param, false>::steps;
    else
        steps = key.subset ? KsShape<TFHEpp::lvl10param, true>::steps
                           : KsShape<TFHEpp::lvl10param, false>::steps;
    return (static_cast<size_t>(steps) * njobs * 4 + 255) & ~size_t(255);
}

int ks_gemm_run(const KsGemmKey& key, const KsJob* d_jobs, int njobs, void* ws, hipStream_t st)
{
    if (njobs <= 0) return 0;
    if (key.level == 2) return run_shape<TFHEpp::lvl20param, false, kKsMfmaJobBlocks>(key, d_jobs, njobs, ws, st);
    if (key.subset) return run_shape<TFHEpp::lvl10param, true, kKsMfmaJobBlocks>(key, d_jobs, njobs, ws, st);
    return run_shape<TFHEpp::lvl10param, false, kKsMfmaJobBlocks>(key, d_jobs, njobs, ws, st);
}

}  // namespace cufhe_amd
