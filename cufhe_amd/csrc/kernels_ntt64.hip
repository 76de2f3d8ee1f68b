// lvl2 NTT-backend kernels (Goldilocks prime P = 2^64 - 2^32 + 1): key
// conversion, transform entry points for the parity tests, the batched
// blind rotation (plain CMUX and key bundle) and the plain per-gate lvl02
// form (reference src/bootstrap_gpu.cu:213-340 for the lvl2 NTT key and
// tables, include/gatebootstrapping_gpu.cuh:1187-1332 for the rotation).
#include "br_ntt64.hpp"
#include "keyswitch.hpp"
#include "launch.hpp"

namespace cufhe_amd {

constexpr int kNtt64Tlwe0Pad = 640;  // lvl0 TLWE (637 words) padded in LDS

// ---------------------------------------------------------------------------
// Torus64 polynomials -> Goldilocks spectra (torus_to_ntt_mod, then the
// forward transform), one wave per polynomial.
__global__ __launch_bounds__(256) void k_to_ntt64(uint64_t* out, const uint64_t* in, int npolys,
                                                  Ntt64Tables tb)
{
    constexpr int N = kNtt64N;
    __shared__ uint64_t sc[4][kNtt64Buf];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;  // no workgroup barriers below
    const uint64_t* p = in + static_cast<size_t>(poly) * N;
    uint64_t x[32];
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = torus64_to_gold(p[lane + 64 * r]);
    ntt64_forward(x, sc[wave], tb, lane);
    // word 64 r + lane holds position 32 lane + r: coalesced here and in the rotation
    uint64_t* o = out + static_cast<size_t>(poly) * N;
#pragma unroll
    for (int r = 0; r < 32; r++) o[64 * r + lane] = x[r];
}

// Transform test entry: residues in [0, P) -> spectrum in natural bin order
// (Z[k] = sum_i x_i psi^(i (2k+1)) mod P), or the inverse of that.
template <bool INV>
__global__ __launch_bounds__(256) void k_ntt64_natural(uint64_t* out, const uint64_t* in,
                                                       int npolys, Ntt64Tables tb)
{
    constexpr int N = kNtt64N;
    __shared__ uint64_t sc[4][kNtt64Buf];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int poly = blockIdx.x * 4 + wave;
    if (poly >= npolys) return;
    const uint64_t* p = in + static_cast<size_t>(poly) * N;
    uint64_t* o = out + static_cast<size_t>(poly) * N;
    uint64_t x[32];
    // position q holds bin brv11(q)
    auto bin = [](int q) { return static_cast<int>(__brev(static_cast<uint32_t>(q)) >> 21); };
    if constexpr (!INV) {
#pragma unroll
        for (int r = 0; r < 32; r++) x[r] = p[lane + 64 * r];
        ntt64_forward(x, sc[wave], tb, lane);
#pragma unroll
        for (int r = 0; r < 32; r++) o[bin(32 * lane + r)] = x[r];
    }
    else {
#pragma unroll
        for (int r = 0; r < 32; r++) x[r] = p[bin(32 * lane + r)];
        ntt64_inverse(x, sc[wave], tb, lane);
#pragma unroll
        for (int r = 0; r < 32; r++) o[lane + 64 * r] = x[r];
    }
}

// ---------------------------------------------------------------------------
// Batched lvl2 blind rotation + SampleExtract (or the TRLWE itself): one job
// per workgroup of two waves (wave j owns TRLWE component j; 68 KiB LDS, so
// two workgroups per CU). The workgroups walk the jobs with a grid stride in
// a grid that fits on the device at once, keeping them on the same rotation
// step so their key reads meet in L2.
template <bool KB>
__global__ __launch_bounds__(128) void k_br_ntt64(const BrJob* __restrict__ jobs, int njobs,
                                                  uint64_t mu, Ntt64Keys K)
{
    using P = lvl2param;
    constexpr int N = P::n, n0 = lvl0param::n;
    __shared__ uint64_t sh_buf[2][kNtt64Buf];
    __shared__ uint64_t sh_coef[2][N];
    __shared__ uint32_t sh_tl0[kNtt64Tlwe0Pad];
    const int j = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int job = blockIdx.x; job < njobs; job += gridDim.x) {
        const BrJob jb = jobs[job];
        // the previous job's rotation has finished reading sh_tl0
        __syncthreads();
        for (int i = threadIdx.x; i <= n0; i += 128) {
            uint32_t v = static_cast<uint32_t>(jb.ca) * jb.in0[i];
            if (jb.in1) v += static_cast<uint32_t>(jb.cb) * jb.in1[i];
            if (i == n0) v += jb.offset;
            sh_tl0[i] = v;
        }
        __syncthreads();
        AccStore<P> acc;
        acc.lds = sh_coef[j];
        blind_rotate_ntt64<KB>(sh_tl0, mu, acc, j, sh_buf[j], sh_buf[j ^ 1], K);
        if (jb.flags & kBrOutTrlwe)
            trlwe_store<P>(acc, j, static_cast<uint64_t*>(jb.out));
        else
            sample_extract_store<P>(acc, j, static_cast<uint64_t*>(jb.out));
    }
}

// Workgroups of k_br_ntt64<KB> resident on the device at once (0 if the
// runtime cannot tell), cached per instance.
template <bool KB>
static int br_ntt64_resident_groups()
{
    static int cached = -1;
    int v = __atomic_load_n(&cached, __ATOMIC_RELAXED);
    if (v < 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_br_ntt64<KB>, 128, 0) != hipSuccess ||
            per_cu < 0)
            per_cu = 0;
        v = per_cu * device_cu_count();
        __atomic_store_n(&cached, v, __ATOMIC_RELAXED);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Plain per-gate form, lvl02 API (BR-IKS order) on the NTT backend, as
// k_gate_lvl02_fft: waves 0 and 1 rotate (MUX: its two rotations one after
// the other), waves 2 and 3 keep step through the rotation's barriers, then
// all four waves run the lvl2 -> lvl0 key switch.
template <bool KB, bool MUX>
__global__ __launch_bounds__(256, 1) void k_gate_lvl02_ntt(
    uint32_t* out, const uint32_t* in0, const uint32_t* in1, const uint32_t* inc, int32_t ca,
    int32_t cb, uint32_t offset, int32_t mux_sign, const uint32_t* ksk, Ntt64Keys K)
{
    using P = lvl2param;
    using T = uint64_t;
    constexpr int N = P::n, n0 = lvl0param::n;
    constexpr T mu = static_cast<T>(P::μ);
    __shared__ uint64_t sh_buf[2][kNtt64Buf];
    __shared__ T sh_coef[2][N];
    __shared__ uint32_t sh_tl0[kNtt64Tlwe0Pad];
    __shared__ T sh_res[MUX ? 2 : 1][N + 8];
    static_assert(sizeof(sh_buf) >= 4 * (n0 + 1) * sizeof(uint32_t), "key-switch reduction scratch");
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
            blind_rotate_ntt64<KB>(sh_tl0, mu, acc, j, sh_buf[j], sh_buf[j ^ 1], K);
            sample_extract_store<P>(acc, j, sh_res[pass]);
        }
        else {
            blind_rotate_ntt64_ghost<KB>(K);
        }
    }
    __syncthreads();
    uint32_t* const red = reinterpret_cast<uint32_t*>(&sh_buf[0][0]);
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
void launch_to_ntt64(uint64_t* out, const uint64_t* in, int npolys, const Ntt64Tables& tb,
                     hipStream_t st)
{
    if (npolys <= 0) return;
    hipLaunchKernelGGL(k_to_ntt64, dim3((npolys + 3) / 4), dim3(256), 0, st, out, in, npolys, tb);
}

void launch_ntt64_natural(bool inverse, uint64_t* out, const uint64_t* in, int npolys,
                          const Ntt64Tables& tb, hipStream_t st)
{
    if (npolys <= 0) return;
    const dim3 grid((npolys + 3) / 4), block(256);
    if (inverse)
        hipLaunchKernelGGL(k_ntt64_natural<true>, grid, block, 0, st, out, in, npolys, tb);
    else
        hipLaunchKernelGGL(k_ntt64_natural<false>, grid, block, 0, st, out, in, npolys, tb);
}

void launch_br_ntt64(bool kb, const BrJob* jobs, int njobs, uint64_t mu, const Ntt64Keys& K,
                     hipStream_t st)
{
    if (njobs <= 0) return;
    int grid = njobs;
    const int cap = kb ? br_ntt64_resident_groups<true>() : br_ntt64_resident_groups<false>();
    if (cap > 0 && grid > cap) grid = cap;
    if (kb)
        hipLaunchKernelGGL(k_br_ntt64<true>, dim3(grid), dim3(128), 0, st, jobs, njobs, mu, K);
    else
        hipLaunchKernelGGL(k_br_ntt64<false>, dim3(grid), dim3(128), 0, st, jobs, njobs, mu, K);
}

void launch_gate_lvl02_ntt(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                           const Ntt64Keys& K, hipStream_t st)
{
    const dim3 grid(1), block(256);
#define CUFHE_L(KBV, MUXV)                                                                     \
    hipLaunchKernelGGL((k_gate_lvl02_ntt<KBV, MUXV>), grid, block, 0, st, out, in0, in1, inc, \
                       ca, cb, offset, mux_sign, ksk, K)
    if (kb) {
        if (mux) CUFHE_L(true, true); else CUFHE_L(true, false);
    }
    else {
        if (mux) CUFHE_L(false, true); else CUFHE_L(false, false);
    }
#undef CUFHE_L
}

}  // namespace cufhe_amd
