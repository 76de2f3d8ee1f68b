// FFT-backend kernels: key conversion, batched blind rotation, plain
// per-gate kernels, key switching, MUX combine, and FFNT entry points.
#include "br_fft.hpp"
#include "br_fft_kbs.hpp"
#include "br_fft_kbw.hpp"
#include "keyswitch.hpp"
#include "launch.hpp"

namespace cufhe_amd {

constexpr int kTlwe0Pad = 640;  // lvl0 TLWE (637 words) padded in LDS

// Occupancy target of the batched blind rotation (waves per SIMD); bounds
// the VGPR budget the compiler schedules for (3 -> 168 VGPRs).
#ifndef CUFHE_BR_WAVES_PER_SIMD
#define CUFHE_BR_WAVES_PER_SIMD 3
#endif

// ---------------------------------------------------------------------------
// Polynomial -> Fourier spectra. One wave per polynomial.
// Reference src/bootstrap_gpu.cu:59-112 (__TRGSW2FFT__): signed torus value
// * 2^-w, fold, twist, forward FFT. INTEGER=true skips the 2^-w scaling and
// reads int32 inputs (used for X^a - 1 and test polynomials).
template <class P, bool INTEGER>
__global__ __launch_bounds__(256) void k_to_fourier(cplx* out, const void* in, int npolys,
                                                    FourierKeys<P> K)
{
    using T = typename P::T;
    using ST = std::make_signed_t<T>;
    constexpr int N = P::n, H = N / 2, R = H / 64;
    __shared__ cplx sc[4][H];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;  // no workgroup barriers below
    cplx z[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = lane + 64 * r;
        double re, im;
        if constexpr (INTEGER) {
            const int32_t* p = static_cast<const int32_t*>(in) + static_cast<size_t>(poly) * N;
            re = static_cast<double>(p[m]);
            im = static_cast<double>(p[m + H]);
        }
        else {
            const T* p = static_cast<const T*>(in) + static_cast<size_t>(poly) * N;
            constexpr double s = 1.0 / kTorusScale<T>;
            re = static_cast<double>(static_cast<ST>(p[m])) * s;
            im = static_cast<double>(static_cast<ST>(p[m + H])) * s;
        }
        z[r] = cplx{re, im};
    }
    fft_forward<H>(z, sc[wave], K.fft);
#pragma unroll
    for (int d = 0; d < R; d++) out[static_cast<size_t>(poly) * H + d * 64 + lane] = z[d];
}

// X^a - 1 polynomials for a in [0, 2N) (reference src/ntt_small_modulus.cu
// :928-981 XaittGen): coefficient 0 is -1, coefficient a mod N gets +1
// (a < N) or -1 (a >= N).
template <class P>
__global__ void k_xai_poly(int32_t* out)
{
    constexpr int N = P::n;
    const int a = blockIdx.x;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        int32_t v = (i == 0) ? -1 : 0;
        if (i == (a & (N - 1))) v += (a < N) ? 1 : -1;
        out[static_cast<size_t>(a) * N + i] = v;
    }
}

// Spectra -> real polynomial (inverse FFT, untwist incl. 1/H), as doubles.
template <class P>
__global__ __launch_bounds__(256) void k_from_fourier(double* out, const cplx* in, int npolys,
                                                      FourierKeys<P> K)
{
    constexpr int N = P::n, H = N / 2, R = H / 64;
    __shared__ cplx sc[4][H];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;
    cplx z[R];
#pragma unroll
    for (int d = 0; d < R; d++) z[d] = in[static_cast<size_t>(poly) * H + d * 64 + lane];
    fft_inverse<H>(z, sc[wave], K.fft);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int m = lane + 64 * r;
        const cplx v = cscale(z[r], 1.0 / H);
        out[static_cast<size_t>(poly) * N + m] = v.x;
        out[static_cast<size_t>(poly) * N + m + H] = v.y;
    }
}

// ---------------------------------------------------------------------------
// Batched blind rotation: C jobs per workgroup, two waves per job. The
// workgroups walk the job groups with a grid stride, so a grid that fits on
// the device at once keeps every workgroup on the same rotation step: their
// key reads then meet in L2 instead of drifting apart as later workgroups
// start when earlier ones finish.
template <class P, bool KB, int C>
__global__ __launch_bounds__(128 * C, (sizeof(typename P::T) == 8 && C > 1) ? 1 : CUFHE_BR_WAVES_PER_SIMD) void k_br_fft(const BrJob* __restrict__ jobs, int njobs,
                                                     typename P::T mu, FourierKeys<P> K)
{
    using T = typename P::T;
    constexpr int N = P::n, H = N / 2, n0 = lvl0param::n;
    constexpr bool kL = AccStore<P>::kLds;
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2 * C][NB * H];
    __shared__ uint32_t sh_tl0[C][kTlwe0Pad];
    __shared__ T sh_coef[kL ? 2 * C : 1][kL ? N : 1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    const int ngroups = (njobs + C - 1) / C;
    for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const int raw = grp * C + slot;
        const bool live = raw < njobs;
        const BrJob jb = jobs[live ? raw : njobs - 1];
        // the previous group's rotation has finished reading sh_tl0
        __syncthreads();
        for (int i = threadIdx.x & 127; i <= n0; i += 128) {
            uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
            if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
            if (i == n0) v += jb.offset;
            sh_tl0[slot][i] = v;
        }
        __syncthreads();
        AccStore<P> acc;
        if constexpr (kL) acc.lds = sh_coef[wave];
        blind_rotate_fft<P, KB>(sh_tl0[slot], mu, acc, j, sh_spec[wave], sh_spec[wave ^ 1], K);
        if (live) {
            if (jb.flags & kBrOutTrlwe)
                trlwe_store<P>(acc, j, static_cast<T*>(jb.out));
            else
                sample_extract_store<P>(acc, j, static_cast<T*>(jb.out));
        }
    }
}

// Workgroups of k_br_fft<P, KB, C> resident on the device at once (0 if the
// runtime cannot tell), cached per instance.
template <class P, bool KB, int C>
static int br_fft_resident_groups()
{
    static int cached = -1;
    int v = __atomic_load_n(&cached, __ATOMIC_RELAXED);
    if (v < 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_br_fft<P, KB, C>, 128 * C, 0) !=
                hipSuccess ||
            per_cu < 0)
            per_cu = 0;
        v = per_cu * device_cu_count();
        __atomic_store_n(&cached, v, __ATOMIC_RELAXED);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Batched key-bundle rotation at lvl1 with the key shared by the workgroup's
// C jobs through an LDS ring (br_fft_kbs.hpp). One workgroup of 2C waves per
// CU (C = 6: 12 waves, 138 KiB LDS).
template <int C>
__global__ __launch_bounds__(128 * C, 1) void k_br_fft_kbs(const BrJob* __restrict__ jobs,
                                                            int njobs, uint32_t mu,
                                                            FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int N = P::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    constexpr int NT = 128 * C;
    __shared__ KbsShared<C> sm;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    const int t = threadIdx.x;
    for (int i = t; i < 64; i += NT) {
        sm.twl[i] = K.fft.twl[i];
        sm.tws[i] = K.fft.tws[i];
    }
    for (int i = t; i < 64; i += NT) {
        sm.tw64[i] = K.fft.tw64[i];
        sm.e2n_lo[i] = K.e2n[i];
    }
    for (int i = t; i < R; i += NT) sm.cr[i] = K.fft.cr[i];
    for (int i = t; i < 2 * N / 64; i += NT) sm.e2n_hi[i] = K.e2n[64 * i];
    const int raw = blockIdx.x * C + slot;
    const bool live = raw < njobs;
    const BrJob jb = jobs[live ? raw : njobs - 1];
    constexpr uint32_t rnd = kMaskRound<P>;
    for (int i = t & 127; i <= n0; i += 128) {
        uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
        if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
        if (i < n0)
            sm.abar[slot][i] = static_cast<uint16_t>(mod_switch<P>(v + rnd));
        else
            sm.bar[slot] = 2 * N - mod_switch<P>(v + jb.offset);
    }
    __syncthreads();
    AccStore<P> acc;
    blind_rotate_kbs<C>(sm, slot, j, mu, acc.reg, K.bk);
    if (live) {
        if (jb.flags & kBrOutTrlwe)
            trlwe_store<P>(acc, j, static_cast<uint32_t*>(jb.out));
        else
            sample_extract_store<P>(acc, j, static_cast<uint32_t*>(jb.out));
    }
}

// ---------------------------------------------------------------------------
// Batched key-bundle rotation at lvl1, one wave per job, key shared by the
// workgroup's jobs through an LDS ring (br_fft_kbw.hpp). One workgroup of C
// waves per CU (C = 8: 2 waves/SIMD, 141 KiB LDS). The launcher spreads the
// jobs evenly over the fewest whole waves of CUs that hold them (at most C
// per workgroup); waves without a job only stream the key, and fewer
// rotating waves per CU finish a round sooner (one: 3.5 ms, eight: 6.3 ms).
template <int C>
__global__ __launch_bounds__(64 * C, 8 / C) void k_br_fft_kbw(const BrJob* __restrict__ jobs,
                                                           int njobs, uint32_t mu,
                                                           FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int N = P::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    constexpr int NT = 64 * C;
    __shared__ KbwShared<C> sm;
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = threadIdx.x;
    for (int i = t; i < 64; i += NT) {
        sm.twl[i] = K.fft.twl[i];
        sm.tws[i] = K.fft.tws[i];
    }
    for (int i = t; i < 64; i += NT) {
        sm.tw64[i] = K.fft.tw64[i];
        sm.e2n_lo[i] = K.e2n[i];
    }
    for (int i = t; i < R; i += NT) sm.cr[i] = K.fft.cr[i];
    for (int i = t; i < 2 * N / 64; i += NT) sm.e2n_hi[i] = K.e2n[64 * i];
    // jobs spread evenly over the grid: workgroup b takes [b n / G, (b+1) n / G)
    const int first = static_cast<int>((static_cast<int64_t>(blockIdx.x) * njobs) / gridDim.x);
    const int count =
        static_cast<int>((static_cast<int64_t>(blockIdx.x + 1) * njobs) / gridDim.x) - first;
    const bool live = slot < count;
    if (!live) {
        // no job for this wave: stream its share of the key, meet the barriers
        __syncthreads();
        KbwKeyPipe<C> pipe(sm, slot, t & 63, K.bk);
        pipe.run_loader_only();
        return;
    }
    const BrJob jb = jobs[first + slot];
    constexpr uint32_t rnd = kMaskRound<P>;
    for (int i = t & 63; i <= n0; i += 64) {
        uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
        if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
        if (i < n0)
            sm.abar[slot][i] = static_cast<uint16_t>(mod_switch<P>(v + rnd));
        else
            sm.bar[slot] = 2 * N - mod_switch<P>(v + jb.offset);
    }
    __syncthreads();
    uint32_t acc[2][16];
    blind_rotate_kbw<C>(sm, slot, mu, acc, K.bk);
    {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            AccStore<P> a;
#pragma unroll
            for (int r = 0; r < 16; r++) a.reg[r] = acc[j][r];
            if (jb.flags & kBrOutTrlwe)
                trlwe_store<P>(a, j, static_cast<uint32_t*>(jb.out));
            else
                sample_extract_store<P>(a, j, static_cast<uint32_t*>(jb.out));
        }
    }
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl0 API (BR-IKS at lvl1; reference
// src/cufhe_gates_gpu.cu:198-511, src/bootstrap_gpu.cu:1557-1672): lvl0
// inputs pre-added into the rotation, SampleExtract, key switch lvl1 -> lvl0.
// MUX/NMUX rotate c + t - mu and -c + f - mu on two wave pairs and key-switch
// +-(r1 + r0 + mu) in one pass.
template <bool KB, bool MUX, bool SUBSET>
__global__ __launch_bounds__(512, 1) void k_gate_lvl01_fft(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
    FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int C = MUX ? 2 : 1;
    constexpr int N = P::n, H = N / 2, n0 = lvl0param::n;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2 * C][NB * H];
    __shared__ uint32_t sh_tl0[C][kTlwe0Pad];
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
        AccStore<P> acc;
        blind_rotate_fft<P, KB>(sh_tl0[slot], tv_mu, acc, j, sh_spec[wave], sh_spec[wave ^ 1], K);
        sample_extract_store<P>(acc, j, sh_res[slot]);
    }
    else {
        blind_rotate_ghost<P, KB>(K);
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

// SampleExtract + key switch lvl1 -> lvl0 of n TRLWEs (reference SEIandKS,
// src/keyswitch_gpu.cu:61-68, include/keyswitch_gpu.cuh:52-123), one job per
// workgroup.
template <bool SUBSET>
__global__ __launch_bounds__(512) void k_se_ks_lvl10(uint32_t* out, const uint32_t* trlwe, int n,
                                                     const uint32_t* ksk)
{
    constexpr int N = lvl1param::n, n0 = lvl0param::n;
    __shared__ uint32_t tl[N + 4];
    __shared__ uint32_t sh_red[8 * (n0 + 1)];
    const int job = blockIdx.x;
    if (job >= n) return;  // whole workgroup
    const uint32_t* tr = trlwe + static_cast<size_t>(job) * 2 * N;
    for (int i = threadIdx.x; i <= N; i += 512)
        tl[i] = i == 0 ? tr[0] : (i < N ? 0u - tr[N - i] : tr[N]);
    __syncthreads();
    // the input coefficients split over the eight waves (a job's key switch
    // is a long dependent gather loop; eight shorter ones hide its latency)
    key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(out + static_cast<size_t>(job) * (n0 + 1), tl,
                                                    nullptr, 1, 0, 0u, ksk, sh_red,
                                                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6),
                                                    threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl1 API (IKS-BR order), one gate per launch.
// Binary gates: one job; MUX/NMUX: two jobs run concurrently, then combined
// (reference src/bootstrap_gpu.cu:1223-1246, :2069-2197).
template <bool KB, bool MUX>
__global__ __launch_bounds__(512, 1) void k_gate_lvl1_fft(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
    FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int C = MUX ? 2 : 1;
    constexpr int N = P::n, H = N / 2, n0 = lvl0param::n;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2 * C][NB * H];
    __shared__ uint32_t sh_tl0[C][kTlwe0Pad];
    __shared__ uint32_t sh_res[MUX ? 2 : 1][MUX ? N + 1 : 1];
    __shared__ uint32_t sh_red[8 * (n0 + 1)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = wave >> 1, j = wave & 1;
    // key switch over all eight waves of the workgroup (MUX: four per
    // rotation input), then only the rotation's 2 C waves go on; a
    // workgroup barrier waits only for the waves that have not ended
    if constexpr (MUX) {
        // job 0: c + t - mu ; job 1: -c + f - mu (both AND-like, +-mu out)
        const int job = wave >> 2;
        key_switch_waves<TFHEpp::lvl10param, false, 4>(
            sh_tl0[job], inc, job == 0 ? in1 : in0, job == 0 ? 1 : -1, 1, 0u - mu, ksk,
            sh_red + job * 4 * (n0 + 1), wave & 3, threadIdx.x & 63);
    }
    else {
        key_switch_waves<TFHEpp::lvl10param, false, 8>(sh_tl0[0], in0, in1, ca, cb, offset, ksk,
                                                       sh_red, wave, threadIdx.x & 63);
    }
    __syncthreads();
    if (wave >= 2 * C) return;
    AccStore<P> acc;
    blind_rotate_fft<P, KB>(sh_tl0[slot], mu, acc, j, sh_spec[wave], sh_spec[wave ^ 1], K);
    if constexpr (!MUX) {
        sample_extract_store<P>(acc, j, out);
    }
    else {
        sample_extract_store<P>(acc, j, sh_res[slot]);
        __syncthreads();
        for (int i = threadIdx.x; i <= N; i += 256) {
            uint32_t v = sh_res[0][i] + sh_res[1][i];
            if (i == N) v += mu;
            out[i] = mux_sign > 0 ? v : 0u - v;
        }
    }
}

// ---------------------------------------------------------------------------
// Plain per-gate form with the key bundle at lvl1 (IKS then BR): one gate per
// workgroup of eight waves. The gate's rotation (MUX/NMUX: two rotations) runs
// on wave 0 (and 1) with br_fft_kbw.hpp's one-wave-per-job code; the other
// waves only stream the key through the LDS ring, so a single gate gets the
// shared-key pipeline at the latency of one rotating wave. Outputs are
// bit-identical to k_gate_lvl1_fft (same key switch, same rotation
// arithmetic).
template <bool MUX>
__global__ __launch_bounds__(512, 1) void k_gate_lvl1_kbw(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
    FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int C = 8, NJ = MUX ? 2 : 1;
    constexpr int N = P::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    constexpr int NT = 64 * C;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    constexpr uint32_t rnd = kMaskRound<P>;
    __shared__ KbwShared<C> sm;
    __shared__ uint32_t sh_tl0[NJ][kTlwe0Pad];
    __shared__ uint32_t sh_res[MUX ? 2 : 1][MUX ? N + 1 : 1];
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = threadIdx.x;
    for (int i = t; i < 64; i += NT) {
        sm.twl[i] = K.fft.twl[i];
        sm.tws[i] = K.fft.tws[i];
    }
    for (int i = t; i < 64; i += NT) {
        sm.tw64[i] = K.fft.tw64[i];
        sm.e2n_lo[i] = K.e2n[i];
    }
    for (int i = t; i < R; i += NT) sm.cr[i] = K.fft.cr[i];
    for (int i = t; i < 2 * N / 64; i += NT) sm.e2n_hi[i] = K.e2n[64 * i];
    // key switch over all eight waves (the transform scratch is free yet and
    // holds the partial sums)
    uint32_t* const red = reinterpret_cast<uint32_t*>(&sm.scratch[0][0]);
    if constexpr (MUX) {
        // job 0: c + t - mu ; job 1: -c + f - mu (both AND-like, +-mu out)
        const int job = slot >> 2;
        key_switch_waves<TFHEpp::lvl10param, false, 4>(
            sh_tl0[job], inc, job == 0 ? in1 : in0, job == 0 ? 1 : -1, 1, 0u - mu, ksk,
            red + job * 4 * (n0 + 1), slot & 3, t & 63);
    }
    else {
        key_switch_waves<TFHEpp::lvl10param, false, 8>(sh_tl0[0], in0, in1, ca, cb, offset, ksk,
                                                       red, slot, t & 63);
    }
    __syncthreads();
#pragma unroll
    for (int job = 0; job < NJ; job++) {
        for (int i = t; i <= n0; i += NT) {
            if (i < n0)
                sm.abar[job][i] = static_cast<uint16_t>(mod_switch<P>(sh_tl0[job][i] + rnd));
            else
                sm.bar[job] = 2 * N - mod_switch<P>(sh_tl0[job][n0]);
        }
    }
    __syncthreads();
    if (slot < NJ) {
        uint32_t acc[2][16];
        blind_rotate_kbw<C>(sm, slot, mu, acc, K.bk);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            AccStore<P> a;
#pragma unroll
            for (int r = 0; r < 16; r++) a.reg[r] = acc[j][r];
            sample_extract_store<P>(a, j, MUX ? sh_res[slot] : out);
        }
    }
    else {
        KbwKeyPipe<C> pipe(sm, slot, t & 63, K.bk);
        pipe.run_loader_only();
    }
    if constexpr (MUX) {
        __syncthreads();
        for (int i = t; i <= N; i += NT) {
            uint32_t v = sh_res[0][i] + sh_res[1][i];
            if (i == N) v += mu;
            out[i] = mux_sign > 0 ? v : 0u - v;
        }
    }
}

// Plain per-gate form with the key bundle, lvl0 API (BR then IKS), on the
// same eight-wave organisation as k_gate_lvl1_kbw: waves 0 (and 1) rotate,
// the others stream the key; then all eight waves key-switch. Outputs are
// bit-identical to k_gate_lvl01_fft.
template <bool MUX, bool SUBSET>
__global__ __launch_bounds__(512, 1) void k_gate_lvl01_kbw(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
    FourierKeys<lvl1param> K)
{
    using P = lvl1param;
    constexpr int C = 8, NJ = MUX ? 2 : 1;
    constexpr int N = P::n, H = N / 2, R = H / 64, n0 = lvl0param::n;
    constexpr int NT = 64 * C;
    constexpr uint32_t mu = static_cast<uint32_t>(P::μ);
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    constexpr uint32_t rnd = kMaskRound<P>;
    __shared__ KbwShared<C> sm;
    __shared__ uint32_t sh_res[NJ][N + 4];
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = threadIdx.x;
    for (int i = t; i < 64; i += NT) {
        sm.twl[i] = K.fft.twl[i];
        sm.tws[i] = K.fft.tws[i];
    }
    for (int i = t; i < 64; i += NT) {
        sm.tw64[i] = K.fft.tw64[i];
        sm.e2n_lo[i] = K.e2n[i];
    }
    for (int i = t; i < R; i += NT) sm.cr[i] = K.fft.cr[i];
    for (int i = t; i < 2 * N / 64; i += NT) sm.e2n_hi[i] = K.e2n[64 * i];
    // the pre-added lvl0 inputs (as k_gate_lvl01_fft), mod-switched
#pragma unroll
    for (int job = 0; job < NJ; job++) {
        const uint32_t* a = MUX ? inc : in0;
        const uint32_t* b = MUX ? (job == 0 ? in1 : in0) : in1;
        const int32_t sa = MUX ? (job == 0 ? 1 : -1) : ca;
        const int32_t sb = MUX ? 1 : cb;
        const uint32_t off = MUX ? 0u - mu0 : offset;
        for (int i = t; i <= n0; i += NT) {
            uint32_t v = static_cast<uint32_t>(sa) * a[i] + static_cast<uint32_t>(sb) * b[i];
            if (i < n0)
                sm.abar[job][i] = static_cast<uint16_t>(mod_switch<P>(v + rnd));
            else
                sm.bar[job] = 2 * N - mod_switch<P>(v + off);
        }
    }
    __syncthreads();
    if (slot < NJ) {
        uint32_t acc[2][16];
        blind_rotate_kbw<C>(sm, slot, tv_mu, acc, K.bk);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            AccStore<P> r;
#pragma unroll
            for (int k = 0; k < 16; k++) r.reg[k] = acc[j][k];
            sample_extract_store<P>(r, j, sh_res[slot]);
        }
    }
    else {
        KbwKeyPipe<C> pipe(sm, slot, t & 63, K.bk);
        pipe.run_loader_only();
    }
    __syncthreads();
    // key switch over all eight waves, partial sums in the transform scratch
    uint32_t* const red = reinterpret_cast<uint32_t*>(&sm.scratch[0][0]);
    if constexpr (MUX)
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(out, sh_res[0], sh_res[1], mux_sign,
                                                        mux_sign, mux_sign > 0 ? mu : 0u - mu, ksk,
                                                        red, slot, t & 63);
    else
        key_switch_waves<TFHEpp::lvl10param, SUBSET, 8>(out, sh_res[0], nullptr, 1, 0, 0u, ksk,
                                                        red, slot, t & 63);
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl02 API (BR-IKS order): lvl0 in, BR to lvl2
// (Torus64), SE, key switch lvl2 -> lvl0 (reference src/bootstrap_gpu.cu
// :1248-1273, :1305-1331, MUX :1949-2060). MUX runs its two rotations one
// after the other in the same wave pair (LDS holds one lvl2 job).
template <bool KB, bool MUX>
__global__ __launch_bounds__(256, 1) void k_gate_lvl02_fft(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
    FourierKeys<lvl2param> K)
{
    using P = lvl2param;
    using T = uint64_t;
    constexpr int N = P::n, H = N / 2, n0 = lvl0param::n;
    constexpr T mu = static_cast<T>(P::μ);
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2][NB * H];
    __shared__ T sh_coef[2][N];
    __shared__ uint32_t sh_tl0[kTlwe0Pad];
    __shared__ T sh_res[MUX ? 2 : 1][N + 8];
    // waves 0 and 1 rotate; waves 2 and 3 keep step through the rotation's
    // barriers and then join the key switch (four waves: one per SIMD, so
    // the rotating waves keep the whole register file)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = wave & 1;
    const int t = threadIdx.x & 127;
    const bool rotating = wave < 2;
    AccStore<P> acc;
    acc.lds = sh_coef[j];
#pragma unroll 1
    for (int pass = 0; pass < (MUX ? 2 : 1); pass++) {
        // MUX: pass 0 = c + t - mu, pass 1 = -c + f - mu
        const uint32_t* a = MUX ? inc : in0;
        const uint32_t* b = MUX ? (pass == 0 ? in1 : in0) : in1;
        const int32_t sa = MUX ? (pass == 0 ? 1 : -1) : ca;
        const int32_t sb = MUX ? 1 : cb;
        const uint32_t off = MUX ? 0u - static_cast<uint32_t>(lvl0param::μ) : offset;
        __syncthreads();
        if (rotating) {
            for (int i = t; i <= n0; i += 128) {
                uint32_t v = static_cast<uint32_t>(sa) * a[i] + static_cast<uint32_t>(sb) * b[i];
                if (i == n0) v += off;
                sh_tl0[i] = v;
            }
        }
        __syncthreads();
        if (rotating) {
            blind_rotate_fft<P, KB>(sh_tl0, mu, acc, j, sh_spec[j], sh_spec[j ^ 1], K);
            sample_extract_store<P>(acc, j, sh_res[pass]);
        }
        else {
            blind_rotate_ghost<P, KB>(K);
        }
    }
    __syncthreads();
    // partial sums of the key switch in the (now free) transform scratch
    uint32_t* const red = reinterpret_cast<uint32_t*>(&sh_spec[0][0]);
    if constexpr (MUX) {
        key_switch_waves<TFHEpp::lvl20param, false, 4>(out, sh_res[0], sh_res[1], mux_sign,
                                                       mux_sign, mux_sign > 0 ? mu : 0 - mu, ksk,
                                                       red, wave, threadIdx.x & 63);
    }
    else {
        key_switch_waves<TFHEpp::lvl20param, false, 4>(out, sh_res[0], nullptr, 1, 0, 0, ksk, red,
                                                       wave, threadIdx.x & 63);
    }
}

// ---------------------------------------------------------------------------
// Batched key switching (gather form): one job per workgroup of eight waves,
// each wave taking 1/8 of the input coefficients for all output columns,
// partial sums reduced through LDS. One job's key switch is a long dependent
// gather loop (1024 x 7 KSK rows at lvl1, 2048 x 7 at lvl2), so a job on a
// single wave is latency-bound for milliseconds; split eight ways it is not.
template <class KS, bool SUBSET>
__global__ __launch_bounds__(512) void k_ks_batch(const KsJob* __restrict__ jobs, int njobs,
                                                  const uint32_t* ksk)
{
    using DT = typename KS::domainP::T;
    __shared__ uint32_t sh_red[8 * (KsTraits<KS>::n0 + 1)];
    const int job = blockIdx.x;
    if (job >= njobs) return;
    const KsJob jb = jobs[job];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    key_switch_waves<KS, SUBSET, 8>(jb.out, static_cast<const DT*>(jb.in0),
                                    static_cast<const DT*>(jb.in1), jb.ca, jb.cb,
                                    static_cast<DT>(jb.offset), ksk, sh_red, wave,
                                    threadIdx.x & 63);
}

// Elementwise TLWE combinations (MUX/NMUX combine, NOT, COPY).
__global__ __launch_bounds__(256) void k_linear(const LinJob* __restrict__ jobs, int njobs)
{
    const LinJob jb = jobs[blockIdx.x];
    for (int i = threadIdx.x; i < jb.words; i += 256) {
        uint32_t v = static_cast<uint32_t>(jb.s1) * jb.r1[i];
        if (jb.r0) v += static_cast<uint32_t>(jb.s0) * jb.r0[i];
        if (i == jb.words - 1) v += jb.offset;
        jb.out[i] = v;
    }
}

// One CMUX accumulate step on a TRLWE in global memory (reference
// include/gatebootstrapping_gpu.cuh:69-221): acc += ExtProd((X^a-1) acc, C).
template <class P>
__global__ __launch_bounds__(128, 2) void k_cmux_fft(typename P::T* trlwe, const cplx* trgsw,
                                                   uint32_t abar, FourierKeys<P> K)
{
    using T = typename P::T;
    constexpr int N = P::n, H = N / 2, R = N / 128;
    constexpr bool kL = AccStore<P>::kLds;
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2][NB * H];
    __shared__ T sh_coef[kL ? 2 : 1][kL ? N : 1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    AccStore<P> acc;
    if constexpr (kL) acc.lds = sh_coef[wave];
#pragma unroll
    for (int r = 0; r < 2 * R; r++)
        acc.set(r, lane, trlwe[wave * N + AccStore<P>::coef_index(r, lane)]);
    if constexpr (kL) wave_lds_sync();
    br_step_fft<P, false>(acc, wave, abar, 0, trgsw, sh_spec[wave], sh_spec[wave ^ 1], K);
#pragma unroll
    for (int r = 0; r < 2 * R; r++)
        trlwe[wave * N + AccStore<P>::coef_index(r, lane)] = acc.get(r, lane);
}

// Batched lvl1 CMUX (reference src/bootstrap_gpu.cu:700-813 __CMUXNTT__,
// FFT build): res = c0 + ExtProd(c1 - c0, C), one workgroup of two waves per
// CMUX (wave j owns component j; the pair exchanges digit spectra through
// LDS). Each wave reads its own component of c1 and c0 before the first
// barrier and writes its component of res last, so res may alias c0 or c1.
template <class P>
__global__ __launch_bounds__(128, 2) void k_cmux_batch_fft(CmuxBatch batch, FourierKeys<P> K)
{
    static_assert(!AccStore<P>::kLds, "register accumulator (lvl1)");
    constexpr int N = P::n, R = N / 128;
    constexpr int NB = BrShape<P>::NB;
    __shared__ cplx sh_spec[2][NB * (N / 2)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const CmuxDesc d = batch.at(blockIdx.x);
    AccStore<P> acc;
#pragma unroll
    for (int r = 0; r < 2 * R; r++) {
        const int i = wave * N + AccStore<P>::coef_index(r, lane);
        acc.set(r, lane, d.c1[i] - d.c0[i]);
    }
    br_step_fft<P, false, false, true>(acc, wave, 0, 0, static_cast<const cplx*>(d.cs),
                                       sh_spec[wave], sh_spec[wave ^ 1], K);
#pragma unroll
    for (int r = 0; r < 2 * R; r++) {
        const int i = wave * N + AccStore<P>::coef_index(r, lane);
        d.res[i] = d.c0[i] + acc.get(r, lane);
    }
}

// ---------------------------------------------------------------------------
// Launchers

void launch_cmux_batch_fft(const CmuxBatch& batch, int n, const FourierKeys<lvl1param>& K,
                           hipStream_t st)
{
    hipLaunchKernelGGL((k_cmux_batch_fft<lvl1param>), dim3(n), dim3(128), 0, st, batch, K);
}

template <class P>
void launch_to_fourier(cplx* out, const void* in, int npolys, bool integer,
                       const FourierKeys<P>& K, hipStream_t st)
{
    const dim3 grid((npolys + 3) / 4), block(256);
    if (integer)
        hipLaunchKernelGGL((k_to_fourier<P, true>), grid, block, 0, st, out, in, npolys, K);
    else
        hipLaunchKernelGGL((k_to_fourier<P, false>), grid, block, 0, st, out, in, npolys, K);
}
template void launch_to_fourier<lvl1param>(cplx*, const void*, int, bool,
                                           const FourierKeys<lvl1param>&, hipStream_t);
template void launch_to_fourier<lvl2param>(cplx*, const void*, int, bool,
                                           const FourierKeys<lvl2param>&, hipStream_t);

template <class P>
void launch_from_fourier(double* out, const cplx* in, int npolys, const FourierKeys<P>& K,
                         hipStream_t st)
{
    hipLaunchKernelGGL((k_from_fourier<P>), dim3((npolys + 3) / 4), dim3(256), 0, st, out, in,
                       npolys, K);
}
template void launch_from_fourier<lvl1param>(double*, const cplx*, int,
                                             const FourierKeys<lvl1param>&, hipStream_t);
template void launch_from_fourier<lvl2param>(double*, const cplx*, int,
                                             const FourierKeys<lvl2param>&, hipStream_t);

template <class P>
void launch_xai_poly(int32_t* out, hipStream_t st)
{
    hipLaunchKernelGGL((k_xai_poly<P>), dim3(2 * P::n), dim3(256), 0, st, out);
}
template void launch_xai_poly<lvl1param>(int32_t*, hipStream_t);
template void launch_xai_poly<lvl2param>(int32_t*, hipStream_t);

template <class P, bool KB>
void launch_br_fft(const BrJob* jobs, int njobs, int c, typename P::T mu,
                   const FourierKeys<P>& K, hipStream_t st)
{
    if (njobs <= 0) return;
    const dim3 block(128 * c);
    const dim3 grid((njobs + c - 1) / c);
    switch (c) {
        case 1: {
            // one resident grid (grid-stride over the jobs) for the Torus64
            // rotation, whose key (167 MB, KB 250 MB) outgrows the caches
            // once the workgroups drift apart
            dim3 g1 = grid;
            if constexpr (sizeof(typename P::T) == 8) {
                const int cap = br_fft_resident_groups<P, KB, 1>();
                if (cap > 0 && static_cast<int>(g1.x) > cap) g1.x = cap;
            }
            hipLaunchKernelGGL((k_br_fft<P, KB, 1>), g1, block, 0, st, jobs, njobs, mu, K);
            break;
        }
        case 2: {
            // Torus64: two jobs (136 KiB LDS) per workgroup in a resident grid,
            // their key loads issued together (a tuning option; one job per
            // workgroup measured faster on the lvl02 mix)
            dim3 g2 = grid;
            if constexpr (sizeof(typename P::T) == 8) {
                const int cap = br_fft_resident_groups<P, KB, 2>();
                if (cap > 0 && static_cast<int>(g2.x) > cap) g2.x = cap;
            }
            hipLaunchKernelGGL((k_br_fft<P, KB, 2>), g2, block, 0, st, jobs, njobs, mu, K);
            break;
        }
        case 4:
            if constexpr (sizeof(typename P::T) == 4)
                hipLaunchKernelGGL((k_br_fft<P, KB, 4>), grid, block, 0, st, jobs, njobs, mu,
                                   K);
            break;
        case 8:
            // key bundle: one wave per job (k_br_fft_kbw); otherwise two
            if constexpr (KB && sizeof(typename P::T) == 4)
                hipLaunchKernelGGL((k_br_fft_kbw<8>), dim3(spread_grid(njobs, 8)), dim3(64 * 8), 0,
                                   st, jobs, njobs, mu, K);
            else if constexpr (sizeof(typename P::T) == 4)
                hipLaunchKernelGGL((k_br_fft<P, KB, 8>), grid, block, 0, st, jobs, njobs, mu,
                                   K);
            break;
        case 6:
            if constexpr (KB && sizeof(typename P::T) == 4)
                hipLaunchKernelGGL((k_br_fft_kbs<6>), grid, block, 0, st, jobs, njobs, mu, K);
            break;
        default:
            break;
    }
}
template void launch_br_fft<lvl1param, true>(const BrJob*, int, int, uint32_t,
                                             const FourierKeys<lvl1param>&, hipStream_t);
template void launch_br_fft<lvl1param, false>(const BrJob*, int, int, uint32_t,
                                              const FourierKeys<lvl1param>&, hipStream_t);
template void launch_br_fft<lvl2param, true>(const BrJob*, int, int, uint64_t,
                                             const FourierKeys<lvl2param>&, hipStream_t);
template void launch_br_fft<lvl2param, false>(const BrJob*, int, int, uint64_t,
                                              const FourierKeys<lvl2param>&, hipStream_t);

void launch_gate_lvl1_fft(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                          const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                          uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                          const FourierKeys<lvl1param>& K, hipStream_t st)
{
    const dim3 grid(1), block(512);
#define CUFHE_L(KBV, MUXV)                                                                    \
    hipLaunchKernelGGL((k_gate_lvl1_fft<KBV, MUXV>), grid, block, 0, st, out, in0, in1, inc, \
                       ca, cb, offset, mux_sign, ksk, K)
    if (kb) {
        // key bundle: the one-wave rotation with seven key-loading waves
        if (mux)
            hipLaunchKernelGGL((k_gate_lvl1_kbw<true>), grid, dim3(512), 0, st, out, in0, in1, inc,
                               ca, cb, offset, mux_sign, ksk, K);
        else
            hipLaunchKernelGGL((k_gate_lvl1_kbw<false>), grid, dim3(512), 0, st, out, in0, in1,
                               inc, ca, cb, offset, mux_sign, ksk, K);
    }
    else {
        if (mux) CUFHE_L(false, true); else CUFHE_L(false, false);
    }
#undef CUFHE_L
}

void launch_gate_lvl01_fft(bool kb, bool mux, bool subset, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
                           const FourierKeys<lvl1param>& K, hipStream_t st)
{
#define CUFHE_G01(KBV, MUXV, SUB)                                                              \
    hipLaunchKernelGGL((k_gate_lvl01_fft<KBV, MUXV, SUB>), dim3(1), dim3(512), 0, \
                       st, out, in0, in1, inc, ca, cb, offset, mux_sign, tv_mu, ksk, K)
    if (kb) {
        // key bundle: the one-wave rotation with seven key-streaming waves
#define CUFHE_K01(MUXV, SUB)                                                                     \
    hipLaunchKernelGGL((k_gate_lvl01_kbw<MUXV, SUB>), dim3(1), dim3(512), 0, st, out, in0, in1, \
                       inc, ca, cb, offset, mux_sign, tv_mu, ksk, K)
        if (mux) {
            if (subset) CUFHE_K01(true, true); else CUFHE_K01(true, false);
        }
        else {
            if (subset) CUFHE_K01(false, true); else CUFHE_K01(false, false);
        }
#undef CUFHE_K01
    }
    else {
        if (mux) {
            if (subset) CUFHE_G01(false, true, true); else CUFHE_G01(false, true, false);
        }
        else {
            if (subset) CUFHE_G01(false, false, true); else CUFHE_G01(false, false, false);
        }
    }
#undef CUFHE_G01
}

void launch_se_ks_lvl10(bool subset, uint32_t* out, const uint32_t* trlwe, int n,
                        const uint32_t* ksk, hipStream_t st)
{
    if (n <= 0) return;
    if (subset)
        hipLaunchKernelGGL((k_se_ks_lvl10<true>), dim3(n), dim3(512), 0, st, out, trlwe, n, ksk);
    else
        hipLaunchKernelGGL((k_se_ks_lvl10<false>), dim3(n), dim3(512), 0, st, out, trlwe, n, ksk);
}

void launch_gate_lvl02_fft(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                           const FourierKeys<lvl2param>& K, hipStream_t st)
{
    const dim3 grid(1), block(256);
#define CUFHE_L(KBV, MUXV)                                                                     \
    hipLaunchKernelGGL((k_gate_lvl02_fft<KBV, MUXV>), grid, block, 0, st, out, in0, in1, inc, \
                       ca, cb, offset, mux_sign, ksk, K)
    if (kb) {
        if (mux) CUFHE_L(true, true); else CUFHE_L(true, false);
    }
    else {
        if (mux) CUFHE_L(false, true); else CUFHE_L(false, false);
    }
#undef CUFHE_L
}

void launch_ks_batch(int domain_level, bool subset, const KsJob* jobs, int njobs,
                     const uint32_t* ksk, hipStream_t st)
{
    if (njobs <= 0) return;
    const dim3 grid(njobs), block(512);
    if (domain_level == 1) {
        if (subset)
            hipLaunchKernelGGL((k_ks_batch<TFHEpp::lvl10param, true>), grid, block, 0, st, jobs,
                               njobs, ksk);
        else
            hipLaunchKernelGGL((k_ks_batch<TFHEpp::lvl10param, false>), grid, block, 0, st,
                               jobs, njobs, ksk);
    }
    else {
        hipLaunchKernelGGL((k_ks_batch<TFHEpp::lvl20param, false>), grid, block, 0, st, jobs,
                           njobs, ksk);
    }
}

__global__ __launch_bounds__(256) void k_linear_one(LinJob jb)
{
    for (int i = threadIdx.x; i < jb.words; i += 256) {
        uint32_t v = static_cast<uint32_t>(jb.s1) * jb.r1[i];
        if (jb.r0) v += static_cast<uint32_t>(jb.s0) * jb.r0[i];
        if (i == jb.words - 1) v += jb.offset;
        jb.out[i] = v;
    }
}

void launch_linear_one(const LinJob& jb, hipStream_t st)
{
    hipLaunchKernelGGL(k_linear_one, dim3(1), dim3(256), 0, st, jb);
}

void launch_linear(const LinJob* jobs, int njobs, hipStream_t st)
{
    hipLaunchKernelGGL(k_linear, dim3(njobs), dim3(256), 0, st, jobs, njobs);
}

template <class P>
void launch_cmux_fft(typename P::T* trlwe, const cplx* trgsw, uint32_t abar,
                     const FourierKeys<P>& K, hipStream_t st)
{
    hipLaunchKernelGGL((k_cmux_fft<P>), dim3(1), dim3(128), 0, st, trlwe, trgsw, abar, K);
}
template void launch_cmux_fft<lvl1param>(uint32_t*, const cplx*, uint32_t,
                                         const FourierKeys<lvl1param>&, hipStream_t);
template void launch_cmux_fft<lvl2param>(uint64_t*, const cplx*, uint32_t,
                                         const FourierKeys<lvl2param>&, hipStream_t);

}  // namespace cufhe_amd
