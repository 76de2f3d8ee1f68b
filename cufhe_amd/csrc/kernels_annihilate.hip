// Annihilation key switching on the FFT backend (reference
// src/annihilate_gpu.cu:676-803 AnnihilateKeySwitching*, with the kernels it
// launches: :180-202 halving and sum, :436-462 / :557-603 EvalAuto, :245-360
// the half external product).
//
// Per TRLWE c = (a, b) of ring size N = 2^nbit, nbit rounds, round `bit`
// with d = 2^(bit+1) + 1:
//   c >>= 1                                  (every torus word, logical shift)
//   c += (0, tau_d(b)) - ExtProd_half(tau_d(a), AHK[bit])
// tau_d is the automorphism p(X) -> p(X^d); the half external product
// decomposes tau_d(a) into l gadget digits and multiplies digit q by the two
// polynomials of the key's row q. After the nbit rounds the phase keeps the
// constant coefficient of the input's phase and sums the others to ~0.
//
// MI355X layout: one workgroup of two waves per TRLWE runs all nbit rounds
// with the TRLWE resident in LDS (the reference launches three kernels per
// round and round-trips the TRLWE through global memory between them).
// Wave w owns output component w. The l digit transforms of a round are split
// over the two waves (digits 2p and 2p+1 transformed side by side, the
// spectra exchanged through LDS), every wave multiplies all l spectra with
// its component's key row and inverse-transforms only its own component.
// tau_d is applied while reading: coefficient j of tau_d(p) is p[u] for
// u = j d^-1 mod 2N below N and -p[u - N] otherwise, so no permuted copy is
// stored. The key spectra (nbit x l x 2 polynomials, 0.33 MB at N = 1024,
// 1.44 MB at N = 2048) stay in L2 across the batch.
//
// k_annihilate_ntt64 is the same schedule on the lvl2 NTT backend (the
// reference's USE_FFT=OFF build, src/annihilate_gpu.cu:362-434): digits as
// signed residues mod the Goldilocks prime, the key as torus_to_ntt_mod of
// its words, products exact mod P, results back through ntt_mod_to_torus64.
#include "br_fft.hpp"
#include "br_ntt64.hpp"
#include "launch.hpp"

namespace cufhe_amd {

template <class P>
__global__ __launch_bounds__(128) void k_annihilate_fft(typename P::T* out, size_t out_stride,
                                                        const typename P::T* in, size_t in_stride,
                                                        const cplx* ahk, FFTTables tb)
{
    using T = typename P::T;
    constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l;
    static_assert(L % 2 == 0, "digits are transformed in pairs, one per wave");
    __shared__ T sc[2 * N];       // the TRLWE: a at [0, N), b at [N, 2N)
    __shared__ cplx spec[2][H];   // per wave: transform scratch, then a digit spectrum
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const size_t job = blockIdx.x;

    const T* src = in + job * in_stride;
    for (int i = threadIdx.x; i < 2 * N; i += 128) sc[i] = src[i];
    __syncthreads();

    constexpr double scale = kTorusScale<T> / H;
    T* const own = sc + wave * N;
#pragma unroll 1
    for (int bit = 0; bit < static_cast<int>(P::nbit); bit++) {
        const uint32_t d = (2u << bit) + 1u;
        // d^-1 mod 2^32 by Newton's iteration (d d = 1 mod 8; each step
        // doubles the correct low bits), then mod 2N
        uint32_t di = d;
#pragma unroll
        for (int k = 0; k < 4; k++) di *= 2u - d * di;
        di &= 2u * N - 1u;
        // coefficient j of tau_d(p >> 1)
        auto tau = [di](const T* p, int j) -> T {
            const uint32_t u = (static_cast<uint32_t>(j) * di) & (2u * N - 1u);
            const T v = p[u & (N - 1)] >> 1;
            return u < static_cast<uint32_t>(N) ? v : static_cast<T>(0 - v);
        };

        cplx acc[R];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] = cplx{0.0, 0.0};
#pragma unroll 1
        for (int pr = 0; pr < L / 2; pr++) {
            const int q = 2 * pr + wave;
            cplx z[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int m = lane + 64 * r;
                z[r] = cplx{static_cast<double>(Gadget<P>::digit(tau(sc, m), q)),
                            static_cast<double>(Gadget<P>::digit(tau(sc, m + H), q))};
            }
            fft_forward<H>(z, spec[wave], tb);
#pragma unroll
            for (int r = 0; r < R; r++) lds_store(&spec[wave][r * 64 + lane], z[r]);
            // the key values of this wave's component for the pair's first
            // digit, loaded while the pair meets at the barrier
            const cplx* key0 =
                ahk + ((static_cast<size_t>(bit) * L + 2 * pr) * 2 + wave) * H + lane;
#pragma unroll
            for (int r = 0; r < R; r++) z[r] = key0[r * 64];
            __syncthreads();  // both digit spectra of the pair written
#pragma unroll
            for (int r = 0; r < R; r++) cmac(acc[r], lds_load(&spec[0][r * 64 + lane]), z[r]);
            {
                const cplx* key1 = key0 + 2 * H;
#pragma unroll
                for (int r = 0; r < R; r++) z[r] = key1[r * 64];
#pragma unroll
                for (int r = 0; r < R; r++)
                    cmac(acc[r], lds_load(&spec[1][r * 64 + lane]), z[r]);
            }
            __syncthreads();  // both waves done reading before the next pair's transforms
        }
        // Every read of the a half (both waves, in the digit loop) precedes the
        // barrier above, and only wave 1 reads or writes the b half from here on.
        fft_inverse<H>(acc, spec[wave], tb);
        T nv[2 * R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int m = lane + 64 * r;
            const T lo = double_to_torus<T>(acc[r].x * scale);
            const T hi = double_to_torus<T>(acc[r].y * scale);
            T vlo = static_cast<T>((own[m] >> 1) - lo);
            T vhi = static_cast<T>((own[m + H] >> 1) - hi);
            if (wave == 1) {
                vlo += tau(own, m);
                vhi += tau(own, m + H);
            }
            nv[r] = vlo;
            nv[R + r] = vhi;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            own[lane + 64 * r] = nv[r];
            own[lane + 64 * r + H] = nv[R + r];
        }
        __syncthreads();  // the round's TRLWE visible to both waves
    }

    T* dst = out + job * out_stride;
    for (int i = threadIdx.x; i < 2 * N; i += 128) dst[i] = sc[i];
}

// lvl1 with ONE wave per TRLWE (the form the launcher uses at lvl1): the wave
// decomposes tau_d(a) into both gadget digits at once, transforms the two
// digit polynomials as an interleaved pair (fft_forward2), multiplies them
// into both output components and inverse-transforms the two as a pair
// (fft_inverse2). No workgroup barrier after the table copy: the waves of a
// workgroup are independent TRLWEs. The transforms and the multiply-
// accumulate order are those of k_annihilate_fft, so the output words are
// identical.
constexpr int kAh1Waves = 4;

struct Ah1Shared {
    static constexpr int N = lvl1param::n, H = N / 2, R = H / 64;
    cplx twl[64];
    cplx tws[64];
    cplx tw64[64];
    cplx cr[R];
    uint32_t sc[kAh1Waves][2 * N];                       // per wave: its TRLWE (a, b)
    cplx scratch[kAh1Waves][fft_stage::kScratchPad(H)];  // per wave: transform scratch
};

__global__ __launch_bounds__(64 * kAh1Waves) void k_annihilate_fft1w(uint32_t* out, size_t out_stride,
                                                                    const uint32_t* in,
                                                                    size_t in_stride, int batch,
                                                                    const cplx* ahk, FFTTables gtb)
{
    using P = lvl1param;
    using T = uint32_t;
    constexpr int N = P::n, H = N / 2, R = H / 64, L = P::l;
    static_assert(L == 2, "both digits as one transform pair");
    __shared__ Ah1Shared sm;
    const int t = threadIdx.x;
    for (int i = t; i < 64; i += 64 * kAh1Waves) {
        sm.twl[i] = gtb.twl[i];
        sm.tws[i] = gtb.tws[i];
        sm.tw64[i] = gtb.tw64[i];
    }
    for (int i = t; i < R; i += 64 * kAh1Waves) sm.cr[i] = gtb.cr[i];
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lane = t & 63;
    const size_t job = static_cast<size_t>(blockIdx.x) * kAh1Waves + wave;
    if (job >= static_cast<size_t>(batch)) return;  // no workgroup barrier below
    const FFTTables tb{sm.twl, sm.tw64, sm.cr, sm.tws};
    T* const sc = sm.sc[wave];
    cplx* const scratch = sm.scratch[wave];

    const T* src = in + job * in_stride;
#pragma unroll
    for (int r = 0; r < 2 * N / 64; r++) sc[lane + 64 * r] = src[lane + 64 * r];
    wave_lds_sync();

    constexpr double scale = kTorusScale<T> / H;
#pragma unroll 1
    for (int bit = 0; bit < static_cast<int>(P::nbit); bit++) {
        const uint32_t d = (2u << bit) + 1u;
        uint32_t di = d;
#pragma unroll
        for (int k = 0; k < 4; k++) di *= 2u - d * di;
        di &= 2u * N - 1u;
        auto tau = [di](const T* p, int j) -> T {
            const uint32_t u = (static_cast<uint32_t>(j) * di) & (2u * N - 1u);
            const T v = p[u & (N - 1)] >> 1;
            return u < static_cast<uint32_t>(N) ? v : static_cast<T>(0 - v);
        };
        // key [bit][row q][component o][H]
        const cplx* key = ahk + static_cast<size_t>(bit) * L * 2 * H + lane;
        cplx x[R], y[R];  // digit 0 and digit 1 of tau_d(a >> 1)
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int m = lane + 64 * r;
            const T lo = tau(sc, m), hi = tau(sc, m + H);
            x[r] = cplx{static_cast<double>(Gadget<P>::digit(lo, 0)),
                        static_cast<double>(Gadget<P>::digit(hi, 0))};
            y[r] = cplx{static_cast<double>(Gadget<P>::digit(lo, 1)),
                        static_cast<double>(Gadget<P>::digit(hi, 1))};
        }
        fft_forward2<H>(x, y, scratch, tb);
        // acc_o = D0 K(0, o) + D1 K(1, o); component c = q * 2 + o of the key
#pragma unroll
        for (int r = 0; r < R; r++) {
            // the register's four key values first, then the products (this
            // order measured 10% faster than loads inside the products)
            cplx k[4];
#pragma unroll
            for (int c = 0; c < 4; c++) k[c] = key[c * H + 64 * r];
            cplx a0 = cplx{0.0, 0.0}, a1 = cplx{0.0, 0.0};
            cmac(a0, x[r], k[0 * 2 + 0]);
            cmac(a0, y[r], k[1 * 2 + 0]);
            cmac(a1, x[r], k[0 * 2 + 1]);
            cmac(a1, y[r], k[1 * 2 + 1]);
            x[r] = a0;
            y[r] = a1;
        }
        fft_inverse2<H>(x, y, scratch, tb);
        T na[2 * R], nb[2 * R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int m = lane + 64 * r;
            na[r] = static_cast<T>((sc[m] >> 1) - double_to_torus<T>(x[r].x * scale));
            na[R + r] = static_cast<T>((sc[m + H] >> 1) - double_to_torus<T>(x[r].y * scale));
            nb[r] = static_cast<T>((sc[N + m] >> 1) + tau(sc + N, m) -
                                   double_to_torus<T>(y[r].x * scale));
            nb[R + r] = static_cast<T>((sc[N + m + H] >> 1) + tau(sc + N, m + H) -
                                       double_to_torus<T>(y[r].y * scale));
        }
        wave_lds_sync();  // every read of the round before the first write
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int m = lane + 64 * r;
            sc[m] = na[r];
            sc[m + H] = na[R + r];
            sc[N + m] = nb[r];
            sc[N + m + H] = nb[R + r];
        }
        wave_lds_sync();
    }
    T* dst = out + job * out_stride;
#pragma unroll
    for (int r = 0; r < 2 * N / 64; r++) dst[lane + 64 * r] = sc[lane + 64 * r];
}

__global__ __launch_bounds__(128) void k_annihilate_ntt64(uint64_t* out, size_t out_stride,
                                                          const uint64_t* in, size_t in_stride,
                                                          const uint64_t* ahk, Ntt64Tables tb)
{
    using P = lvl2param;
    constexpr int N = P::n, L = P::l;
    static_assert(L % 2 == 0, "digits are transformed in pairs, one per wave");
    __shared__ uint64_t sc[2 * N];
    __shared__ uint64_t buf[2][kNtt64Buf];  // per wave: transform scratch, then its digit spectrum
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const size_t job = blockIdx.x;

    const uint64_t* src = in + job * in_stride;
    for (int i = threadIdx.x; i < 2 * N; i += 128) sc[i] = src[i];
    __syncthreads();

    uint64_t* const own = sc + wave * N;
#pragma unroll 1
    for (int bit = 0; bit < static_cast<int>(P::nbit); bit++) {
        const uint32_t d = (2u << bit) + 1u;
        uint32_t di = d;
#pragma unroll
        for (int k = 0; k < 4; k++) di *= 2u - d * di;
        di &= 2u * N - 1u;
        auto tau = [di](const uint64_t* p, int j) -> uint64_t {
            const uint32_t u = (static_cast<uint32_t>(j) * di) & (2u * N - 1u);
            const uint64_t v = p[u & (N - 1)] >> 1;
            return u < static_cast<uint32_t>(N) ? v : 0 - v;
        };
        uint64_t sp[32];
#pragma unroll
        for (int r = 0; r < 32; r++) sp[r] = 0;
#pragma unroll 1
        for (int pr = 0; pr < L / 2; pr++) {
            const int q = 2 * pr + wave;
            uint64_t x[32];
#pragma unroll
            for (int r = 0; r < 32; r++)
                x[r] = signed_to_gold(Gadget<P>::digit(tau(sc, lane + 64 * r), q));
            ntt64_forward(x, buf[wave], tb, lane);
#pragma unroll
            for (int r = 0; r < 32; r++) buf[wave][ntt64_addr(32 * lane + r)] = x[r];
            // key word 64 r + lane holds transform position 32 lane + r; the
            // pair's first digit's words are loaded while the pair meets at
            // the barrier
            const uint64_t* key0 =
                ahk + ((static_cast<size_t>(bit) * L + 2 * pr) * 2 + wave) * N + lane;
#pragma unroll
            for (int r = 0; r < 32; r++) x[r] = key0[64 * r];
            __syncthreads();  // both digit spectra of the pair written
#pragma unroll
            for (int w2 = 0; w2 < 2; w2++) {
                if (w2 == 1) {
#pragma unroll
                    for (int r = 0; r < 32; r++) x[r] = key0[2 * N + 64 * r];
                }
                const uint64_t* spec = buf[w2] + 33 * lane;
#pragma unroll
                for (int r = 0; r < 32; r++) sp[r] = gold::add(sp[r], gold::mul(spec[r], x[r]));
            }
            __syncthreads();  // both waves done reading before the next pair's transforms
        }
        ntt64_inverse(sp, buf[wave], tb, lane);
        // lane holds coefficients lane + 64 r; the a half was last read before
        // the barrier above, the b half is read and written by wave 1 only
        uint64_t nv[32];
#pragma unroll
        for (int r = 0; r < 32; r++) {
            const int e = lane + 64 * r;
            uint64_t v = (own[e] >> 1) - gold_to_torus64(sp[r]);
            if (wave == 1) v += tau(own, e);
            nv[r] = v;
        }
#pragma unroll
        for (int r = 0; r < 32; r++) own[lane + 64 * r] = nv[r];
        __syncthreads();
    }

    uint64_t* dst = out + job * out_stride;
    for (int i = threadIdx.x; i < 2 * N; i += 128) dst[i] = sc[i];
}

void launch_annihilate_ntt64(uint64_t* out, size_t out_stride, const uint64_t* in,
                             size_t in_stride, int batch, const uint64_t* ahk,
                             const Ntt64Tables& tb, hipStream_t st)
{
    if (batch <= 0) return;
    hipLaunchKernelGGL(k_annihilate_ntt64, dim3(batch), dim3(128), 0, st, out, out_stride, in,
                       in_stride, ahk, tb);
}

template <class P>
void launch_annihilate_fft(typename P::T* out, size_t out_stride, const typename P::T* in,
                           size_t in_stride, int batch, const cplx* ahk, const FourierKeys<P>& K,
                           hipStream_t st)
{
    if (batch <= 0) return;
    if constexpr (sizeof(typename P::T) == 4) {
        hipLaunchKernelGGL(k_annihilate_fft1w, dim3((batch + kAh1Waves - 1) / kAh1Waves),
                           dim3(64 * kAh1Waves), 0, st, out, out_stride, in, in_stride, batch, ahk,
                           K.fft);
        return;
    }
    hipLaunchKernelGGL((k_annihilate_fft<P>), dim3(batch), dim3(128), 0, st, out, out_stride, in,
                       in_stride, ahk, K.fft);
}
template void launch_annihilate_fft<lvl1param>(uint32_t*, size_t, const uint32_t*, size_t, int,
                                               const cplx*, const FourierKeys<lvl1param>&,
                                               hipStream_t);
template void launch_annihilate_fft<lvl2param>(uint64_t*, size_t, const uint64_t*, size_t, int,
                                               const cplx*, const FourierKeys<lvl2param>&,
                                               hipStream_t);

}  // namespace cufhe_amd
