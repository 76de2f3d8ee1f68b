// Wavefront-level FP64 FFT for the folded negacyclic transform (FFNT).
//
// Reference behaviour (include/ntt_small_modulus.cuh:827-1476, tables
// src/ntt_small_modulus.cu:207-244): a real polynomial p of degree < N is
// folded into H = N/2 complex values z[m] = (p[m] + i p[m+H]) * twist[m],
// twist[m] = exp(i*pi*m/N), and transformed by an H-point cyclic FFT; the
// inverse applies the inverse FFT and untwist[m] = conj(twist[m]) / H.
// The spectrum order is internal: the key, the X^a-1 table and the digits are
// all transformed by these same routines, so only consistency matters.
//
// MI355X design: ONE wavefront (64 lanes) owns one H-point transform with
// R = H/64 complex values per lane in VGPRs (R=8 for N=1024, R=16 for
// N=2048). H = R x 8 x 8:
//   1. R-point DFT in registers over m2 (m = lane + 64*m2),
//   2. twiddle W_H^(lane*k2),
//   3. 64-point DFT over the lane index as 8x8 with two LDS transposes
//      (XOR-swizzled, bank-conflict-free 16-byte accesses), no workgroup
//      barrier: only the wave itself touches its scratch.
// Output: lane L = c + 8q, register g*8+d holds Z[k] for
//   k = (q + 8g) + R*(c + 8d); it is stored at p = (g*8+d)*64 + L so that a
//   wave reads a spectrum with fully coalesced 1 KiB accesses per register.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace cufhe_amd {

struct cplx {
    double x, y;
};

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return {fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ cplx cmulc(cplx a, cplx b)  // a * conj(b)
{
    return {fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y)};
}
// acc += a * b
__device__ __forceinline__ void cmac(cplx& acc, cplx a, cplx b)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.x * s, a.y * s}; }

// Order LDS traffic of one wavefront: every lane's earlier LDS accesses are
// complete and visible before any later one (LDS ops of a wave execute in
// issue order; the fences stop the compiler from reordering across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hide a uniform pointer's provenance so loads through it are not hoisted
// out of a loop (keeps loop-invariant twiddle tables out of VGPRs; they
// are re-read from L1 instead).
template <class T>
__device__ __forceinline__ T* opaque_ptr(T* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ cplx lds_load(const cplx* p)
{
    const double2 v = *reinterpret_cast<const double2*>(p);
    return {v.x, v.y};
}
__device__ __forceinline__ void lds_store(cplx* p, cplx v)
{
    *reinterpret_cast<double2*>(p) = make_double2(v.x, v.y);
}

// Compile-time loop: f(std::integral_constant<int, I>{}) for I in [0, N).
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// cos(2*pi*e/16), sin(2*pi*e/16) for e in [0, 16).
__host__ __device__ constexpr double cos16(int e)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173,
                     r = 0.70710678118654752440;
    constexpr double t[16] = {1.0, c1, r, s1, 0.0, -s1, -r, -c1,
                              -1.0, -c1, -r, -s1, 0.0, s1, r, c1};
    return t[e & 15];
}
__host__ __device__ constexpr double sin16(int e) { return cos16(e + 12); }

// Multiply by W_M^e = exp(-2*pi*i*e/M) (INV: exp(+2*pi*i*e/M)), e constant,
// M in {2, 4, 8, 16}. Multiples of a quarter turn are exact swaps.
template <int M, int E, bool INV>
__device__ __forceinline__ cplx twiddle_const(cplx v)
{
    constexpr int e16 = (E * (16 / M)) & 15;
    if constexpr (e16 == 0) {
        return v;
    }
    else if constexpr (e16 == 8) {
        return {-v.x, -v.y};
    }
    else if constexpr (e16 == 4) {  // forward: -i
        return INV ? cplx{-v.y, v.x} : cplx{v.y, -v.x};
    }
    else if constexpr (e16 == 12) {  // forward: +i
        return INV ? cplx{v.y, -v.x} : cplx{-v.y, v.x};
    }
    else {
        constexpr double c = cos16(e16);
        constexpr double s = INV ? -sin16(e16) : sin16(e16);
        // (x + iy)(c - i s) = (x c + y s) + i (y c - x s)
        return {fma(v.x, c, v.y * s), fma(v.y, c, -v.x * s)};
    }
}

// In-register M-point DFT (radix-2 DIF, natural order in and out).
template <int M, bool INV>
__device__ __forceinline__ void dft_reg(cplx* x)
{
    static_for<4>([&](auto sc) {
        constexpr int half = (M >> 1) >> decltype(sc)::value;
        if constexpr (half >= 1) {
#pragma unroll
            for (int start = 0; start < M; start += 2 * half) {
                static_for<half>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const cplx u = x[start + j];
                    const cplx v = x[start + j + half];
                    x[start + j] = cadd(u, v);
                    x[start + j + half] = twiddle_const<M, j*(M / (2 * half)), INV>(csub(u, v));
                });
            }
        }
    });
    // bit-reversal permutation (register renaming after unrolling)
    cplx tmp[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
        int r = 0;
#pragma unroll
        for (int b = 1, s = M / 2; b < M; b <<= 1, s >>= 1)
            if (i & b) r |= s;
        tmp[r] = x[i];
    }
#pragma unroll
    for (int i = 0; i < M; i++) x[i] = tmp[i];
}

// Transpose slots in the wave's scratch (16-byte elements). Rows are 64
// elements. SWZ: a 16-lane group of a 128-bit LDS access must touch 16
// distinct 4-bank units; odd rows are swizzled so that the two rows read or
// written by the same 16 lanes (q = 2m, 2m+1) land in opposite bank halves
// (conflict-free, at the price of per-lane address registers):
//   step 1: element (row, col) at row*64 + (col ^ 8*(row&1))
//   step 2: element (row, c, x) at row*64 + 8*(c ^ (row&1)) + x
// Without SWZ the same slots without the XOR on the row parity (2-way
// conflicts between the two rows of a 16-lane group).
template <bool SWZ>
__device__ __forceinline__ int fft_t1(int row, int col)
{
    return row * 64 + (SWZ ? (col ^ ((row & 1) << 3)) : col);
}
template <bool SWZ>
__device__ __forceinline__ int fft_t2(int row, int c, int x)
{
    return row * 64 + ((SWZ ? (c ^ (row & 1)) : c) << 3) + x;
}

// Wave issue priority (s_setprio) for transforms run by many waves that
// meet at workgroup barriers: a transform starts at priority 3 and steps down
// after each transpose, so the waves that are behind get issue slots first
// and the workgroup reaches its next barrier together (the arbiter otherwise
// favours the oldest waves and the youngest set the pace).
template <bool ON>
__device__ __forceinline__ void wave_prio(int)
{
}
template <>
__device__ __forceinline__ void wave_prio<true>(int p)
{
    switch (p) {
        case 3: __builtin_amdgcn_s_setprio(3); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        default: __builtin_amdgcn_s_setprio(0); break;
    }
}

// Tables of the folded transform (device memory, built at init), H = 64 R:
//   twl[k2*64 + l] = exp(i*pi*l/N) * exp(-2*pi*i*l*k2/H)   (lane twist merged
//                    with the step-2 twiddle), k2 < R, l < 64 (the transforms
//                    read row 0 and build the others with tws, below)
//   tws[l]         = exp(-2*pi*i*l/H)
//   tw64[c*8 + a]  = exp(-2*pi*i*a*c/64)
//   cr[r]          = exp(i*pi*64*r/N)                        (r < R)
// twist[m] = exp(i*pi*m/N) factors as exp(i*pi*l/N) * cr[r] for m = l + 64 r;
// the lane factor is constant over the register DFT, so it moves into twl.
struct FFTTables {
    const cplx* twl;
    const cplx* tw64;
    const cplx* cr;
    const cplx* tws;  // tws[l] = exp(-2*pi*i*l/H)
};

// Twiddle sequences computed from one or two table entries by repeated
// complex multiplication (R - 1 products instead of R - 1 table loads; the
// loads, not the products, limit the transforms). Every transform builds them
// in this one order, so all kernels -- and the key spectra -- see identical
// twiddles:
//   stage A: twl[k2*64 + l] = twl[l] * tws[l]^k2   (k2 = 0 .. R-1)
//   stage B: tw64[c*8 + a]  = tw64[8 + a]^c        (c = 1 .. 7)
struct TwiddleSeq {
    cplx t, w;
    __device__ __forceinline__ void next() { t = cmul(t, w); }
};
__device__ __forceinline__ TwiddleSeq stage_a_twiddles(const FFTTables& tb, int lane)
{
    return TwiddleSeq{tb.twl[lane], tb.tws[lane]};
}
__device__ __forceinline__ TwiddleSeq stage_b_twiddles(const FFTTables& tb, int lo)
{
    const cplx w = tb.tw64[8 + lo];
    return TwiddleSeq{w, w};
}

// Stages of the transforms for the paired forms (fft_forward2 /
// fft_inverse2): the operations of fft_forward / fft_inverse split at the
// transposes, so a polynomial's spectrum is bit-identical whichever form
// transformed it (no expression here can be contracted differently).
namespace fft_stage {

// Scratch slot layouts of the transposes (16-byte elements). LAY 0 and 1 are
// fft_t1/fft_t2 without / with the XOR swizzle; LAY 2 pads rows to 72
// elements and step-2 sub-rows to 9, which is conflict-free for both access
// patterns of both transposes (a 16-lane group's two rows land 8 elements =
// half a bank sweep apart) and needs no per-register lane XOR, so one
// address register serves all eight accesses of a pattern. Scratch for LAY 2
// is kScratchPad(H) elements.
constexpr int kPadRow = 72;
__host__ __device__ constexpr int kScratchPad(int H) { return H / 64 * kPadRow; }
// step 1: element (row, col)
template <int LAY>
__device__ __forceinline__ int slot_t1(int row, int col)
{
    if constexpr (LAY == 2)
        return row * kPadRow + col;
    else
        return fft_t1<LAY == 1>(row, col);
}
// step 2: element (row, c, x) of the 8 x 8 transpose of lane group `row`:
// lane x of the group holds register c before, lane c holds register x after
template <int LAY>
__device__ __forceinline__ int slot_t2(int row, int c, int x)
{
    if constexpr (LAY == 2)
        return row * kPadRow + c * 9 + x;
    else
        return fft_t2<LAY == 1>(row, c, x ^ c);
}

template <int R>
__device__ __forceinline__ void fwd_a(cplx (&x)[R], const FFTTables& tb, int lane)
{
#pragma unroll
    for (int r = 1; r < R; r++) x[r] = cmul(x[r], tb.cr[r]);
    dft_reg<R, false>(x);
    TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) {
        x[k2] = cmul(x[k2], ts.t);
        if (k2 + 1 < R) ts.next();
    }
}
// fwd_a of two transforms with each twiddle read once
template <int R>
__device__ __forceinline__ void fwd_a2(cplx (&x)[R], cplx (&y)[R], const FFTTables& tb, int lane)
{
#pragma unroll
    for (int r = 1; r < R; r++) {
        x[r] = cmul(x[r], tb.cr[r]);
        y[r] = cmul(y[r], tb.cr[r]);
    }
    dft_reg<R, false>(x);
    dft_reg<R, false>(y);
    TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) {
        x[k2] = cmul(x[k2], ts.t);
        y[k2] = cmul(y[k2], ts.t);
        if (k2 + 1 < R) ts.next();
    }
}
template <int R, int LAY>
__device__ __forceinline__ void t1_store(const cplx (&x)[R], cplx* lds, int lane)
{
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) lds_store(&lds[slot_t1<LAY>(k2, lane)], x[k2]);
}
template <int R, int LAY>
__device__ __forceinline__ void t1_load(cplx (&x)[R], const cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int b = 0; b < 8; b++) x[g * 8 + b] = lds_load(&lds[slot_t1<LAY>(g * 8 + q, 8 * b + lo)]);
}
template <int R>
__device__ __forceinline__ void fwd_b(cplx (&x)[R], const FFTTables& tb, int lane)
{
    const int lo = lane & 7;
#pragma unroll
    for (int g = 0; g < R / 8; g++) {
        dft_reg<8, false>(&x[g * 8]);
        TwiddleSeq ts = stage_b_twiddles(tb, lo);
#pragma unroll
        for (int c = 1; c < 8; c++) {
            x[g * 8 + c] = cmul(x[g * 8 + c], ts.t);
            if (c < 7) ts.next();
        }
    }
}
template <int R, int LAY>
__device__ __forceinline__ void t2_store_cx(const cplx (&x)[R], cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int c = 0; c < 8; c++) lds_store(&lds[slot_t2<LAY>(g * 8 + q, c, lo)], x[g * 8 + c]);
}
template <int R, int LAY>
__device__ __forceinline__ void t2_load_xc(cplx (&x)[R], const cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int a = 0; a < 8; a++) x[g * 8 + a] = lds_load(&lds[slot_t2<LAY>(g * 8 + q, lo, a)]);
}
template <int R>
__device__ __forceinline__ void fwd_c(cplx (&x)[R])
{
#pragma unroll
    for (int g = 0; g < R / 8; g++) dft_reg<8, false>(&x[g * 8]);
}
template <int R>
__device__ __forceinline__ void inv_a(cplx (&x)[R])
{
#pragma unroll
    for (int g = 0; g < R / 8; g++) dft_reg<8, true>(&x[g * 8]);
}
template <int R, int LAY>
__device__ __forceinline__ void t2_store_xc(const cplx (&x)[R], cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int a = 0; a < 8; a++) lds_store(&lds[slot_t2<LAY>(g * 8 + q, lo, a)], x[g * 8 + a]);
}
template <int R, int LAY>
__device__ __forceinline__ void t2_load_cx(cplx (&x)[R], const cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int c = 0; c < 8; c++) x[g * 8 + c] = lds_load(&lds[slot_t2<LAY>(g * 8 + q, c, lo)]);
}
template <int R>
__device__ __forceinline__ void inv_b(cplx (&x)[R], const FFTTables& tb, int lane)
{
    const int lo = lane & 7;
#pragma unroll
    for (int g = 0; g < R / 8; g++) {
        // undo W_64^(a*c): lane low bits are a here, register index c
        TwiddleSeq ts = stage_b_twiddles(tb, lo);
#pragma unroll
        for (int c = 1; c < 8; c++) {
            x[g * 8 + c] = cmulc(x[g * 8 + c], ts.t);
            if (c < 7) ts.next();
        }
        dft_reg<8, true>(&x[g * 8]);
    }
}
template <int R, int LAY>
__device__ __forceinline__ void t1_store_inv(const cplx (&x)[R], cplx* lds, int lane)
{
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < R / 8; g++)
#pragma unroll
        for (int b = 0; b < 8; b++) lds_store(&lds[slot_t1<LAY>(g * 8 + q, 8 * b + lo)], x[g * 8 + b]);
}
template <int R, int LAY>
__device__ __forceinline__ void t1_load_inv(cplx (&x)[R], const cplx* lds, int lane)
{
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) x[k2] = lds_load(&lds[slot_t1<LAY>(k2, lane)]);
}
template <int R>
__device__ __forceinline__ void inv_c(cplx (&x)[R], const FFTTables& tb, int lane)
{
    {
        TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
        for (int k2 = 0; k2 < R; k2++) {
            x[k2] = cmulc(x[k2], ts.t);
            if (k2 + 1 < R) ts.next();
        }
    }
    dft_reg<R, true>(x);
#pragma unroll
    for (int r = 1; r < R; r++) x[r] = cmulc(x[r], tb.cr[r]);
}
// inv_c of two transforms with each twiddle read once
template <int R>
__device__ __forceinline__ void inv_c2(cplx (&x)[R], cplx (&y)[R], const FFTTables& tb, int lane)
{
    TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) {
        x[k2] = cmulc(x[k2], ts.t);
        y[k2] = cmulc(y[k2], ts.t);
        if (k2 + 1 < R) ts.next();
    }
    dft_reg<R, true>(x);
    dft_reg<R, true>(y);
#pragma unroll
    for (int r = 1; r < R; r++) {
        x[r] = cmulc(x[r], tb.cr[r]);
        y[r] = cmulc(y[r], tb.cr[r]);
    }
}

}  // namespace fft_stage

// Forward folded transform. In: x[r] = p[m] + i p[m+H] for m = lane + 64 r
// (untwisted). Out: spectrum Z[k] = sum_m (p[m] + i p[m+H]) zeta_k^m with
// zeta_k = exp(i*pi*(1-4k)/N), in the register layout described above.
// lds: H complex scratch owned by this wave.
template <int H, bool SWZ = false, bool PRIO = false>
__device__ __forceinline__ void fft_forward(cplx (&x)[H / 64], cplx* lds, const FFTTables& tb)
{
    constexpr int R = H / 64;
    constexpr int G = R / 8;
    const int lane = threadIdx.x & 63;
    wave_prio<PRIO>(3);
#pragma unroll
    for (int r = 1; r < R; r++) x[r] = cmul(x[r], tb.cr[r]);
    dft_reg<R, false>(x);
    {
        TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
        for (int k2 = 0; k2 < R; k2++) {
            x[k2] = cmul(x[k2], ts.t);
            if (k2 + 1 < R) ts.next();
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) lds_store(&lds[fft_t1<SWZ>(k2, lane)], x[k2]);
    wave_lds_sync();
    const int lo = lane & 7, q = lane >> 3;
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int b = 0; b < 8; b++) x[g * 8 + b] = lds_load(&lds[fft_t1<SWZ>(g * 8 + q, 8 * b + lo)]);
    wave_lds_sync();
    wave_prio<PRIO>(2);
#pragma unroll
    for (int g = 0; g < G; g++) {
        dft_reg<8, false>(&x[g * 8]);
        TwiddleSeq ts = stage_b_twiddles(tb, lo);
#pragma unroll
        for (int c = 1; c < 8; c++) {
            x[g * 8 + c] = cmul(x[g * 8 + c], ts.t);
            if (c < 7) ts.next();
        }
#pragma unroll
        for (int c = 0; c < 8; c++) lds_store(&lds[fft_t2<SWZ>(g * 8 + q, c, lo ^ c)], x[g * 8 + c]);
    }
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int a = 0; a < 8; a++) x[g * 8 + a] = lds_load(&lds[fft_t2<SWZ>(g * 8 + q, lo, a ^ lo)]);
    wave_lds_sync();
    wave_prio<PRIO>(1);
#pragma unroll
    for (int g = 0; g < G; g++) dft_reg<8, false>(&x[g * 8]);
    wave_prio<PRIO>(0);
}

// Inverse folded transform without the 1/H factor: spectrum in the layout
// fft_forward produces; result x[r] = H * (p[m] + i p[m+H]), m = lane + 64 r.
template <int H, bool SWZ = false, bool PRIO = false>
__device__ __forceinline__ void fft_inverse(cplx (&x)[H / 64], cplx* lds, const FFTTables& tb)
{
    constexpr int R = H / 64;
    constexpr int G = R / 8;
    const int lane = threadIdx.x & 63;
    const int lo = lane & 7, q = lane >> 3;
    wave_prio<PRIO>(3);
#pragma unroll
    for (int g = 0; g < G; g++) {
        dft_reg<8, true>(&x[g * 8]);
#pragma unroll
        for (int a = 0; a < 8; a++) lds_store(&lds[fft_t2<SWZ>(g * 8 + q, lo, a ^ lo)], x[g * 8 + a]);
    }
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int c = 0; c < 8; c++) x[g * 8 + c] = lds_load(&lds[fft_t2<SWZ>(g * 8 + q, c, lo ^ c)]);
    wave_lds_sync();
    wave_prio<PRIO>(2);
#pragma unroll
    for (int g = 0; g < G; g++) {
        // undo W_64^(a*c): lane low bits are a here, register index c
        TwiddleSeq ts = stage_b_twiddles(tb, lo);
#pragma unroll
        for (int c = 1; c < 8; c++) {
            x[g * 8 + c] = cmulc(x[g * 8 + c], ts.t);
            if (c < 7) ts.next();
        }
        dft_reg<8, true>(&x[g * 8]);
#pragma unroll
        for (int b = 0; b < 8; b++) lds_store(&lds[fft_t1<SWZ>(g * 8 + q, 8 * b + lo)], x[g * 8 + b]);
    }
    wave_lds_sync();
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) x[k2] = lds_load(&lds[fft_t1<SWZ>(k2, lane)]);
    wave_lds_sync();
    wave_prio<PRIO>(1);
    {
        TwiddleSeq ts = stage_a_twiddles(tb, lane);
#pragma unroll
        for (int k2 = 0; k2 < R; k2++) {
            x[k2] = cmulc(x[k2], ts.t);
            if (k2 + 1 < R) ts.next();
        }
    }
    dft_reg<R, true>(x);
#pragma unroll
    for (int r = 1; r < R; r++) x[r] = cmulc(x[r], tb.cr[r]);
    wave_prio<PRIO>(0);
}

// Two forward transforms by one wave through ONE scratch buffer: the
// transposes take turns on the scratch while the other transform's register
// stage runs, so the two dependency chains overlap.
template <int H, int LAY = 2, bool PRIO = false>
__device__ __forceinline__ void fft_forward2(cplx (&x)[H / 64], cplx (&y)[H / 64], cplx* lds,
                                             const FFTTables& tb)
{
    using namespace fft_stage;
    constexpr int R = H / 64;
    const int lane = threadIdx.x & 63;
    wave_prio<PRIO>(3);
    fwd_a2<R>(x, y, tb, lane);
    t1_store<R, LAY>(x, lds, lane);
    wave_lds_sync();
    t1_load<R, LAY>(x, lds, lane);
    wave_lds_sync();
    t1_store<R, LAY>(y, lds, lane);
    fwd_b<R>(x, tb, lane);
    wave_lds_sync();
    t1_load<R, LAY>(y, lds, lane);
    wave_lds_sync();
    wave_prio<PRIO>(2);
    t2_store_cx<R, LAY>(x, lds, lane);
    fwd_b<R>(y, tb, lane);
    wave_lds_sync();
    t2_load_xc<R, LAY>(x, lds, lane);
    wave_lds_sync();
    wave_prio<PRIO>(1);
    t2_store_cx<R, LAY>(y, lds, lane);
    fwd_c<R>(x);
    wave_lds_sync();
    t2_load_xc<R, LAY>(y, lds, lane);
    wave_lds_sync();
    fwd_c<R>(y);
    wave_prio<PRIO>(0);
}

// Two inverse transforms by one wave through one scratch buffer (as above).
template <int H, int LAY = 2, bool PRIO = false>
__device__ __forceinline__ void fft_inverse2(cplx (&x)[H / 64], cplx (&y)[H / 64], cplx* lds,
                                             const FFTTables& tb)
{
    using namespace fft_stage;
    constexpr int R = H / 64;
    const int lane = threadIdx.x & 63;
    wave_prio<PRIO>(3);
    inv_a<R>(x);
    t2_store_xc<R, LAY>(x, lds, lane);
    inv_a<R>(y);
    wave_lds_sync();
    t2_load_cx<R, LAY>(x, lds, lane);
    wave_lds_sync();
    t2_store_xc<R, LAY>(y, lds, lane);
    inv_b<R>(x, tb, lane);
    wave_lds_sync();
    t2_load_cx<R, LAY>(y, lds, lane);
    wave_lds_sync();
    wave_prio<PRIO>(2);
    t1_store_inv<R, LAY>(x, lds, lane);
    inv_b<R>(y, tb, lane);
    wave_lds_sync();
    t1_load_inv<R, LAY>(x, lds, lane);
    wave_lds_sync();
    wave_prio<PRIO>(1);
    t1_store_inv<R, LAY>(y, lds, lane);
    wave_lds_sync();
    t1_load_inv<R, LAY>(y, lds, lane);
    wave_lds_sync();
    inv_c2<R>(x, y, tb, lane);
    wave_prio<PRIO>(0);
}

}  // namespace cufhe_amd
