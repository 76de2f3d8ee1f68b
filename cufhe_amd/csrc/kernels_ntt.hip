// NTT-backend kernels (lvl1, P = 2147473409): key conversion, batched blind
// rotation (plain CMUX, key bundle, block-binary), the plain per-gate form,
// and transform entry points for the parity tests.
#include "br_ntt.hpp"
#include "br_ntt_shared.hpp"
#include "keyswitch.hpp"
#include "launch.hpp"

namespace cufhe_amd {

constexpr int kNttTlwe0Pad = 640;  // lvl0 TLWE (637 words) padded in LDS

#ifndef CUFHE_BR_NTT_WAVES_PER_SIMD
#define CUFHE_BR_NTT_WAVES_PER_SIMD 3
#endif

// ---------------------------------------------------------------------------
// Torus32 polynomials -> NTT spectra in Montgomery form, one wave per
// polynomial (reference src/bootstrap_gpu.cu:426-450 __TRGSW2NTT__:
// torus_to_ntt_mod, then the forward transform). Bins are stored in the
// transform's register order: word reg*64 + lane.
__global__ __launch_bounds__(256) void k_to_ntt(uint32_t* out, const uint32_t* in, int npolys,
                                                NttTables tb)
{
    constexpr int N = lvl1param::n;
    __shared__ uint32_t sc[4][kNttLdsWords];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;  // no workgroup barriers below
    const uint32_t* p = in + static_cast<size_t>(poly) * N;
    uint32_t x[16];
#pragma unroll
    for (int r = 0; r < 16; r++) x[r] = torus_to_ntt31(p[lane + 64 * r]);
    ntt_forward(x, sc[wave], tb);
    uint32_t* o = out + static_cast<size_t>(poly) * N;
#pragma unroll
    for (int d = 0; d < 16; d++) o[d * 64 + lane] = mont_mul(x[d], ntt31::R2);
}

// Transform test entry: residues in [0, P) -> spectrum in natural bin order
// (Z[k] = sum_i x_i psi^(i(2k+1)), normal form), or the inverse of that.
template <bool INV>
__global__ __launch_bounds__(256) void k_ntt_natural(uint32_t* out, const uint32_t* in, int npolys,
                                                     NttTables tb)
{
    constexpr int N = lvl1param::n;
    __shared__ uint32_t sc[4][kNttLdsWords];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;
    const uint32_t* p = in + static_cast<size_t>(poly) * N;
    uint32_t* o = out + static_cast<size_t>(poly) * N;
    const int c = lane & 7, q = lane >> 3;
    uint32_t x[16];
    if constexpr (!INV) {
#pragma unroll
        for (int r = 0; r < 16; r++) x[r] = p[lane + 64 * r];
        ntt_forward(x, sc[wave], tb);
#pragma unroll
        for (int reg = 0; reg < 16; reg++) {
            const int g = reg >> 3, d = reg & 7;
            o[(q + 8 * g) + 16 * (c + 8 * d)] = x[reg];
        }
    }
    else {
#pragma unroll
        for (int reg = 0; reg < 16; reg++) {
            const int g = reg >> 3, d = reg & 7;
            x[reg] = p[(q + 8 * g) + 16 * (c + 8 * d)];
        }
        ntt_inverse(x, sc[wave], tb);
#pragma unroll
        for (int r = 0; r < 16; r++) o[lane + 64 * r] = x[r];
    }
}

// ---------------------------------------------------------------------------
// Batched blind rotation + SampleExtract: C jobs per workgroup, two waves per
// job (wave j owns TRLWE component j). Inputs are pre-added lvl0 TLWEs.
template <int MODE, int C>
__global__ __launch_bounds__(128 * C, CUFHE_BR_NTT_WAVES_PER_SIMD) void k_br_ntt(
    const BrJob* __restrict__ jobs, int njobs, uint32_t mu, NttKeys K)
{
    constexpr int n0 = lvl0param::n;
    __shared__ uint32_t sh_lds[2 * C][kNttLdsWords];
    __shared__ uint32_t sh_tl0[C][kNttTlwe0Pad];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    const int raw = blockIdx.x * C + slot;
    const bool live = raw < njobs;
    const BrJob jb = jobs[live ? raw : njobs - 1];
    for (int i = threadIdx.x & 127; i <= n0; i += 128) {
        uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
        if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
        if (i == n0) v += jb.offset;
        sh_tl0[slot][i] = v;
    }
    __syncthreads();
    uint32_t acc[16];
    blind_rotate_ntt<MODE>(sh_tl0[slot], mu, acc, j, sh_lds[wave], sh_lds[wave ^ 1], K);
    if (live) {
        if (jb.flags & kBrOutTrlwe)
            trlwe_store_ntt(acc, j, static_cast<uint32_t*>(jb.out));
        else
            sample_extract_ntt(acc, j, static_cast<uint32_t*>(jb.out));
    }
}

// ---------------------------------------------------------------------------
// Batched rotation with the key shared by the workgroup's C jobs through an
// LDS ring (br_ntt_shared.hpp); one workgroup of 2C waves per CU.
template <int MODE, int C>
__global__ __launch_bounds__(128 * C, 1) void k_br_ntt_shared(const BrJob* __restrict__ jobs,
                                                               int njobs, uint32_t mu, NttKeys K)
{
    using P = lvl1param;
    constexpr int N = P::n, n0 = lvl0param::n;
    constexpr int NT = 128 * C;
    constexpr uint32_t rnd = kMaskRound<P>;
    constexpr int shift = 32 - 1 - P::nbit;
    __shared__ NttShared<C> sm;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    const int t = threadIdx.x;
    for (int i = t; i < 16 * 64; i += NT) {
        sm.twl[i] = K.tb.twl[i];
        sm.twli[i] = K.tb.twli[i];
    }
    for (int i = t; i < 64; i += NT) {
        sm.tw64[i] = K.tb.tw64[i];
        sm.tw64i[i] = K.tb.tw64i[i];
    }
    for (int i = t; i < 16; i += NT) {
        sm.cr[i] = K.tb.cr[i];
        sm.cri[i] = K.tb.cri[i];
    }
    for (int i = t; i < 2 * N; i += NT) sm.pw[i] = K.tb.pw[i];
    const int raw = blockIdx.x * C + slot;
    const bool live = raw < njobs;
    const BrJob jb = jobs[live ? raw : njobs - 1];
    uint32_t part = 0;
    for (int i = t & 127; i <= n0; i += 128) {
        uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
        if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
        if (i < n0) {
            const uint32_t m = mod_switch<P>(v + rnd);
            sm.abar[slot][i] = static_cast<uint16_t>(m);
            part += v - (m << shift);
        }
        else {
            sm.bar[slot] = v + jb.offset;  // body; turned into bar below
        }
    }
    if constexpr (MODE == NTT_BB) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) part += __shfl_xor(part, m);
        if ((t & 63) == 0) sm.corr[slot][j] = part;
    }
    __syncthreads();
    if ((t & 127) == 0) {
        uint32_t body = sm.bar[slot];
        if constexpr (MODE == NTT_BB) {
            // reference gatebootstrapping_gpu.cuh:814-838
            const int32_t correction = static_cast<int32_t>(sm.corr[slot][0] + sm.corr[slot][1]);
            body = body - static_cast<uint32_t>(correction / 2) + rnd;
        }
        sm.bar[slot] = 2 * N - mod_switch<P>(body);
    }
    __syncthreads();
    uint32_t acc[16];
    blind_rotate_ntt_shared<MODE, C>(sm, slot, j, mu, acc, K);
    if (live) {
        if (jb.flags & kBrOutTrlwe)
            trlwe_store_ntt(acc, j, static_cast<uint32_t*>(jb.out));
        else
            sample_extract_ntt(acc, j, static_cast<uint32_t*>(jb.out));
    }
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl0 API (BR-IKS at lvl1), as k_gate_lvl01_fft.
template <int MODE, bool SUBSET, bool MUX>
__global__ __launch_bounds__(512, 1) void k_gate_lvl01_ntt(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk, NttKeys K)
{
    using P = lvl1param;
    constexpr int C = MUX ? 2 : 1;
    constexpr int N = P::n, n0 = lvl0param::n;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    __shared__ uint32_t sh_lds[2 * C][kNttLdsWords];
    __shared__ uint32_t sh_tl0[C][kNttTlwe0Pad];
    __shared__ uint32_t sh_res[C][N + 4];
    __shared__ uint32_t sh_red[8 * (n0 + 1)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    const int t = threadIdx.x & 127;
    // waves 0 .. 2C-1 rotate; the others keep step through the rotation's
    // barriers and then join the key switch (eight waves)
    const bool rotating = wave < 2 * C;
    if (rotating) {
        const uint32_t* a = MUX ? inc : in0;
        const uint32_t* b = MUX ? (slot == 0 ? in1 : in0) : in1;
        const int32_t sa = MUX ? (slot == 0 ? 1 : -1) : ca;
        const int32_t sb = MUX ? 1 : cb;
        const uint32_t off = MUX ? 0u - mu0 : offset;
        for (int i = t; i <= n0; i += 128) {
            uint32_t v = static_cast<uint32_t>(sa) * a[i] + static_cast<uint32_t>(sb) * b[i];
            if (i == n0) v += off;
            sh_tl0[slot][i] = v;
        }
    }
    __syncthreads();
    if (rotating) {
        uint32_t acc[16];
        blind_rotate_ntt<MODE>(sh_tl0[slot], tv_mu, acc, j, sh_lds[wave], sh_lds[wave ^ 1], K);
        sample_extract_ntt(acc, j, sh_res[slot]);
    }
    else {
        blind_rotate_ntt_ghost<MODE>(K);
    }
    __syncthreads();
    if constexpr (MUX)
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(out, sh_res[0], sh_res[1], mux_sign,
                                                        mux_sign, mux_sign > 0 ? mu : 0u - mu, ksk,
                                                        sh_red, wave, threadIdx.x & 63);
    else
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(out, sh_res[0], nullptr, 1, 0, 0u, ksk,
                                                        sh_red, wave, threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl1 API (IKS-BR order), one gate per launch; MUX runs
// its two rotations on two wave pairs and combines them.
template <int MODE, bool SUBSET, bool MUX>
__global__ __launch_bounds__(512, 1) void k_gate_lvl1_ntt(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, const uint32_t* ksk, NttKeys K)
{
    using P = lvl1param;
    constexpr int C = MUX ? 2 : 1;
    constexpr int N = P::n, n0 = lvl0param::n;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    __shared__ uint32_t sh_lds[2 * C][kNttLdsWords];
    __shared__ uint32_t sh_tl0[C][kNttTlwe0Pad];
    __shared__ uint32_t sh_res[MUX ? 2 : 1][MUX ? N + 1 : 1];
    __shared__ uint32_t sh_red[8 * (n0 + 1)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    // key switch over all eight waves of the workgroup (a single gate's key
    // switch is a long dependent loop), then only the rotation's 2 C waves
    // go on: the others end here, and a workgroup barrier waits only for
    // the waves that have not ended
    if constexpr (MUX) {
        const int job = wave >> 2;
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 4>(
            sh_tl0[job], inc, job == 0 ? in1 : in0, job == 0 ? 1 : -1, 1, 0u - mu, ksk,
            sh_red + job * 4 * (n0 + 1), wave & 3, threadIdx.x & 63);
    }
    else {
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(sh_tl0[0], in0, in1, ca, cb, offset, ksk,
                                                        sh_red, wave, threadIdx.x & 63);
    }
    __syncthreads();
    if (wave >= 2 * C) return;
    uint32_t acc[16];
    blind_rotate_ntt<MODE>(sh_tl0[slot], mu, acc, j, sh_lds[wave], sh_lds[wave ^ 1], K);
    if constexpr (!MUX) {
        sample_extract_ntt(acc, j, out);
    }
    else {
        sample_extract_ntt(acc, j, sh_res[slot]);
        __syncthreads();
        for (int i = threadIdx.x; i <= N; i += 256) {
            uint32_t v = sh_res[0][i] + sh_res[1][i];
            if (i == N) v += mu;
            out[i] = mux_sign > 0 ? v : 0u - v;
        }
    }
}

// Batched lvl1 CMUX on the NTT backend (reference src/bootstrap_gpu.cu:925-1041
// __CMUXNTT__): res = c0 + ExtProd(c1 - c0, C), one workgroup of two waves
// per CMUX; res may alias c0 or c1 (each wave reads its component of both
// before the first barrier and writes it last).
__global__ __launch_bounds__(128, 2) void k_cmux_batch_ntt(CmuxBatch batch, NttKeys K)
{
    constexpr int N = lvl1param::n;
    __shared__ uint32_t sh_lds[2][kNttLdsWords];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const CmuxDesc d = batch.at(blockIdx.x);
    uint32_t acc[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = wave * N + ntt_coef_index(r, lane);
        acc[r] = d.c1[i] - d.c0[i];
    }
    const uint32_t a[3] = {0u, 0u, 0u};
    br_step_ntt<NTT_DIRECT>(acc, wave, a, static_cast<const uint32_t*>(d.cs), sh_lds[wave],
                            sh_lds[wave ^ 1], K);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = wave * N + ntt_coef_index(r, lane);
        d.res[i] = d.c0[i] + acc[r];
    }
}

// ---------------------------------------------------------------------------
// Launchers.
void launch_cmux_batch_ntt(const CmuxBatch& batch, int n, const NttKeys& K, hipStream_t st)
{
    hipLaunchKernelGGL(k_cmux_batch_ntt, dim3(n), dim3(128), 0, st, batch, K);
}

void launch_to_ntt(uint32_t* out, const uint32_t* in, int npolys, const NttTables& tb,
                   hipStream_t st)
{
    hipLaunchKernelGGL(k_to_ntt, dim3((npolys + 3) / 4), dim3(256), 0, st, out, in, npolys, tb);
}

void launch_ntt_natural(bool inverse, uint32_t* out, const uint32_t* in, int npolys,
                        const NttTables& tb, hipStream_t st)
{
    if (inverse)
        hipLaunchKernelGGL(k_ntt_natural<true>, dim3((npolys + 3) / 4), dim3(256), 0, st, out, in,
                           npolys, tb);
    else
        hipLaunchKernelGGL(k_ntt_natural<false>, dim3((npolys + 3) / 4), dim3(256), 0, st, out, in,
                           npolys, tb);
}

template <int MODE>
static void launch_br_ntt_mode(const BrJob* jobs, int njobs, int c, uint32_t mu, const NttKeys& K,
                               hipStream_t st)
{
    if (njobs <= 0) return;
    switch (c) {
        case 2:
            hipLaunchKernelGGL((k_br_ntt<MODE, 2>), dim3((njobs + 1) / 2), dim3(256), 0, st, jobs,
                               njobs, mu, K);
            break;
        case 4:
            hipLaunchKernelGGL((k_br_ntt<MODE, 4>), dim3((njobs + 3) / 4), dim3(512), 0, st, jobs,
                               njobs, mu, K);
            break;
        case 8:
            hipLaunchKernelGGL((k_br_ntt<MODE, 8>), dim3((njobs + 7) / 8), dim3(1024), 0, st, jobs,
                               njobs, mu, K);
            break;
        case 6:
            hipLaunchKernelGGL((k_br_ntt_shared<MODE, 6>), dim3((njobs + 5) / 6), dim3(768), 0, st,
                               jobs, njobs, mu, K);
            break;
        default:
            hipLaunchKernelGGL((k_br_ntt<MODE, 1>), dim3(njobs), dim3(128), 0, st, jobs, njobs, mu,
                               K);
            break;
    }
}

void launch_br_ntt(int mode, const BrJob* jobs, int njobs, int c, uint32_t mu, const NttKeys& K,
                   hipStream_t st)
{
    if (mode == NTT_KB)
        launch_br_ntt_mode<NTT_KB>(jobs, njobs, c, mu, K, st);
    else if (mode == NTT_BB)
        launch_br_ntt_mode<NTT_BB>(jobs, njobs, c, mu, K, st);
    else
        launch_br_ntt_mode<NTT_PLAIN>(jobs, njobs, c, mu, K, st);
}

template <int MODE, bool SUBSET>
static void launch_gate_ntt_mode(bool mux, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                                 const uint32_t* inc, int32_t ca, int32_t cb, uint32_t offset,
                                 int32_t mux_sign, const uint32_t* ksk, const NttKeys& K,
                                 hipStream_t st)
{
    if (mux)
        hipLaunchKernelGGL((k_gate_lvl1_ntt<MODE, SUBSET, true>), dim3(1), dim3(512), 0, st, out,
                           in0, in1, inc, ca, cb, offset, mux_sign, ksk, K);
    else
        hipLaunchKernelGGL((k_gate_lvl1_ntt<MODE, SUBSET, false>), dim3(1), dim3(512), 0, st, out,
                           in0, in1, inc, ca, cb, offset, mux_sign, ksk, K);
}

void launch_gate_lvl1_ntt(int mode, bool subset, bool mux, uint32_t* out, const uint32_t* in0,
                          const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                          uint32_t offset, int32_t mux_sign, const uint32_t* ksk, const NttKeys& K,
                          hipStream_t st)
{
#define CUFHE_GATE_NTT(M, S)                                                                   \
    launch_gate_ntt_mode<M, S>(mux, out, in0, in1, inc, ca, cb, offset, mux_sign, ksk, K, st); \
    return
    if (subset) {
        if (mode == NTT_KB) { CUFHE_GATE_NTT(NTT_KB, true); }
        if (mode == NTT_BB) { CUFHE_GATE_NTT(NTT_BB, true); }
        CUFHE_GATE_NTT(NTT_PLAIN, true);
    }
    if (mode == NTT_KB) { CUFHE_GATE_NTT(NTT_KB, false); }
    if (mode == NTT_BB) { CUFHE_GATE_NTT(NTT_BB, false); }
    CUFHE_GATE_NTT(NTT_PLAIN, false);
#undef CUFHE_GATE_NTT
}

void launch_gate_lvl01_ntt(int mode, bool subset, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
                           const NttKeys& K, hipStream_t st)
{
#define CUFHE_G01N(M, S, X)                                                                     \
    hipLaunchKernelGGL((k_gate_lvl01_ntt<M, S, X>), dim3(1), dim3(512), 0, st, out, in0,          \
                       in1, inc, ca, cb, offset, mux_sign, tv_mu, ksk, K)
#define CUFHE_G01N_MODE(M)                                                \
    if (subset) {                                                         \
        if (mux) CUFHE_G01N(M, true, true); else CUFHE_G01N(M, true, false); \
    }                                                                     \
    else {                                                                \
        if (mux) CUFHE_G01N(M, false, true); else CUFHE_G01N(M, false, false); \
    }
    if (mode == NTT_KB) {
        CUFHE_G01N_MODE(NTT_KB)
    }
    else if (mode == NTT_BB) {
        CUFHE_G01N_MODE(NTT_BB)
    }
    else {
        CUFHE_G01N_MODE(NTT_PLAIN)
    }
#undef CUFHE_G01N_MODE
#undef CUFHE_G01N
}

}  // namespace cufhe_amd
