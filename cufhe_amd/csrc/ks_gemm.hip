// Batched identity key switching as an exact int8 GEMM on the matrix cores.
//
// Reference semantics (include/keyswitch_gpu.cuh:189-256,
// IdentityKeySwitchPreAdd): for every job and output coefficient i,
//   res_i = base_i - sum_{j, k} sgn(v_jk) * KSK[j][k][|v_jk| - 1][i]
// with v_jk the k-th signed base-2^basebit digit of ca*in0[j] + cb*in1[j]
// (+ offsets). For a batch of M jobs this is a matrix product
//   RES = BASE + S x KSK,   S[job][(j, k, h)] = -sgn(v_jk) if |v_jk| = h + 1 else 0,
// with S in {-1, 0, +1} (M x K, K = switched coefficients x t x 2^(basebit-1))
// and KSK viewed as a K x (n0 + 1) matrix of u32 words. Splitting every word
// into four bytes b_0..b_3 (stored as signed b - 128) turns it into one int8
// x int8 -> int32 GEMM with 4 (n0 + 1) columns:
//   sum_r S[r] KSK[r][i] = sum_b 2^(8b) (sum_r S[r] (B_b[r][i] + 128))
//                        = sum_b 2^(8b) (G[b][i] + 128 rowsum(S))   (mod 2^32)
// Every partial sum is below 2^31 in magnitude (K <= 28672 terms of |.| <=
// 255), so the int32 accumulation is exact and RES is bit-identical to the
// gather form (k_ks_batch). The KSK is read once per 4096-job batch by the
// GEMM instead of once per job.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <mutex>

#include "keyswitch.hpp"
#include "ks_gemm.hpp"

namespace cufhe_amd {

namespace {

std::mutex g_blas_mu;
rocblas_handle g_blas = nullptr;

template <class KS>
__global__ __launch_bounds__(256) void k_ksk_planes(const uint32_t* __restrict__ ksk, int rows,
                                                     int8_t* __restrict__ planes)
{
    constexpr int W = KsTraits<KS>::n0 + 1;
    const size_t total = static_cast<size_t>(rows) * W;
    for (size_t x = blockIdx.x * 256ull + threadIdx.x; x < total; x += gridDim.x * 256ull) {
        const size_t r = x / W;
        const int i = static_cast<int>(x - r * W);
        const uint32_t v = ksk[x];
        int8_t* row = planes + r * (4 * W);
#pragma unroll
        for (int b = 0; b < 4; b++) row[b * W + i] = static_cast<int8_t>(static_cast<int>((v >> (8 * b)) & 255u) - 128);
    }
}

// Digit matrix S (row-major [job][K]), the base terms and rowsum(S).
template <class KS, bool SUBSET>
__global__ __launch_bounds__(256) void k_ks_digits(const KsJob* __restrict__ jobs, int njobs,
                                                    int kdim, int8_t* __restrict__ S,
                                                    uint32_t* __restrict__ base,
                                                    int32_t* __restrict__ rowsum)
{
    using TR = KsTraits<KS>;
    using DT = typename TR::DT;
    constexpr uint32_t first = SUBSET ? TR::n0 : 0;
    constexpr int NB = TR::numbase;
    constexpr int W = TR::n0 + 1;
    __shared__ int32_t red[4];
    const int job = blockIdx.x;
    if (job >= njobs) return;  // whole workgroup
    const KsJob jb = jobs[job];
    const DT* in0 = static_cast<const DT*>(jb.in0);
    const DT* in1 = static_cast<const DT*>(jb.in1);
    const DT ua = static_cast<DT>(static_cast<std::make_signed_t<DT>>(jb.ca));
    const DT ub = static_cast<DT>(static_cast<std::make_signed_t<DT>>(jb.cb));
    int8_t* srow = S + static_cast<size_t>(job) * kdim;
    int32_t part = 0;
    for (uint32_t jj = threadIdx.x; jj < TR::dn - first; jj += 256) {
        const uint32_t j = jj + first;
        DT tmp = ua * in0[j] + TR::kOffset;
        if (in1) tmp += ub * in1[j];
        int8_t* dst = srow + static_cast<size_t>(jj) * KS::t * NB;
#pragma unroll
        for (uint32_t k = 0; k < KS::t; k++) {
            const int32_t val =
                static_cast<int32_t>((tmp >> (TR::W - (k + 1) * KS::basebit)) &
                                     ((static_cast<DT>(1) << KS::basebit) - 1)) -
                static_cast<int32_t>(NB);
            const int32_t mag = val < 0 ? -val : val;
            const int8_t s = val > 0 ? -1 : 1;  // res -= KSK for val > 0, += for val < 0
#pragma unroll
            for (int h = 0; h < NB; h++) dst[k * NB + h] = (mag == h + 1) ? s : 0;
            if (mag) part += s;
        }
    }
    // base terms: body with offset, and the copied coefficients of a subset key
    for (int i = threadIdx.x; i < W; i += 256) {
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
        base[static_cast<size_t>(job) * W + i] = r;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) part += __shfl_xor(part, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) rowsum[job] = red[0] + red[1] + red[2] + red[3];
}

template <class KS>
__global__ __launch_bounds__(256) void k_ks_combine(const KsJob* __restrict__ jobs, int njobs,
                                                     const int32_t* __restrict__ G,
                                                     const uint32_t* __restrict__ base,
                                                     const int32_t* __restrict__ rowsum)
{
    constexpr int W = KsTraits<KS>::n0 + 1;
    const int job = blockIdx.x;
    if (job >= njobs) return;
    uint32_t* out = jobs[job].out;
    const int32_t* g = G + static_cast<size_t>(job) * 4 * W;
    const uint32_t bias = static_cast<uint32_t>(rowsum[job]) * 128u;
    for (int i = threadIdx.x; i < W; i += 256) {
        uint32_t r = base[static_cast<size_t>(job) * W + i];
#pragma unroll
        for (int b = 0; b < 4; b++) r += (static_cast<uint32_t>(g[b * W + i]) + bias) << (8 * b);
        out[i] = r;
    }
}

template <class KS>
constexpr int ks_rows(bool subset)
{
    using TR = KsTraits<KS>;
    return static_cast<int>((subset ? TR::dn - TR::n0 : TR::dn) * KS::t * TR::numbase);
}

}  // namespace

int ks_gemm_prepare(KsGemmKey& key, const uint32_t* d_ksk, int domain_level, bool subset,
                    hipStream_t st)
{
    {
        std::lock_guard<std::mutex> g(g_blas_mu);
        if (!g_blas) {
            if (rocblas_create_handle(&g_blas) != rocblas_status_success) {
                g_blas = nullptr;
                return -1;
            }
            // a fixed workspace: the library never allocates inside a
            // captured plan
            static void* ws = nullptr;
            constexpr size_t kBlasWs = 64u << 20;
            if (hipMalloc(&ws, kBlasWs) != hipSuccess ||
                rocblas_set_workspace(g_blas, ws, kBlasWs) != rocblas_status_success)
                return -1;
        }
    }
    const int rows = domain_level == 2 ? ks_rows<TFHEpp::lvl20param>(false)
                                       : ks_rows<TFHEpp::lvl10param>(subset);
    const int cols = 4 * (lvl0param::n + 1);
    if (hipMalloc(reinterpret_cast<void**>(&key.planes), static_cast<size_t>(rows) * cols) != hipSuccess)
        return -1;
    key.rows = rows;
    key.cols = cols;
    key.level = domain_level;
    key.subset = subset;
    if (domain_level == 2)
        hipLaunchKernelGGL(k_ksk_planes<TFHEpp::lvl20param>, dim3(4096), dim3(256), 0, st, d_ksk, rows,
                           key.planes);
    else
        hipLaunchKernelGGL(k_ksk_planes<TFHEpp::lvl10param>, dim3(4096), dim3(256), 0, st, d_ksk,
                           rows, key.planes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void ks_gemm_release(KsGemmKey& key)
{
    (void)hipFree(key.planes);
    key = KsGemmKey{};
}

size_t ks_gemm_workspace_bytes(const KsGemmKey& key, int njobs)
{
    const size_t W = lvl0param::n + 1, n = static_cast<size_t>(njobs);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return up(n * key.rows) + up(n * key.cols * 4) + up(n * W * 4) + up(n * 4);
}

int ks_gemm_run(const KsGemmKey& key, const KsJob* d_jobs, int njobs, void* ws, hipStream_t st)
{
    if (njobs <= 0) return 0;
    const size_t W = lvl0param::n + 1, n = static_cast<size_t>(njobs);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    int8_t* S = static_cast<int8_t*>(ws);
    int32_t* G = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + up(n * key.rows));
    uint32_t* base = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(G) + up(n * key.cols * 4));
    int32_t* rowsum = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(base) + up(n * W * 4));
    if (key.level == 2)
        hipLaunchKernelGGL((k_ks_digits<TFHEpp::lvl20param, false>), dim3(njobs), dim3(256), 0, st,
                           d_jobs, njobs, key.rows, S, base, rowsum);
    else if (key.subset)
        hipLaunchKernelGGL((k_ks_digits<TFHEpp::lvl10param, true>), dim3(njobs), dim3(256), 0, st,
                           d_jobs, njobs, key.rows, S, base, rowsum);
    else
        hipLaunchKernelGGL((k_ks_digits<TFHEpp::lvl10param, false>), dim3(njobs), dim3(256), 0, st,
                           d_jobs, njobs, key.rows, S, base, rowsum);
    // G (column-major cols x njobs) = planes (cols x rows) * S (rows x njobs)
    {
        std::lock_guard<std::mutex> g(g_blas_mu);
        const int32_t alpha = 1, beta = 0;
        if (rocblas_set_stream(g_blas, st) != rocblas_status_success) return -1;
        const rocblas_status s = rocblas_gemm_ex(
            g_blas, rocblas_operation_none, rocblas_operation_none, key.cols, njobs, key.rows, &alpha,
            key.planes, rocblas_datatype_i8_r, key.cols, S, rocblas_datatype_i8_r, key.rows, &beta, G,
            rocblas_datatype_i32_r, key.cols, G, rocblas_datatype_i32_r, key.cols,
            rocblas_datatype_i32_r, rocblas_gemm_algo_standard, 0, 0);
        if (s != rocblas_status_success) return -1;
    }
    if (key.level == 2)
        hipLaunchKernelGGL(k_ks_combine<TFHEpp::lvl20param>, dim3(njobs), dim3(256), 0, st, d_jobs,
                           njobs, G, base, rowsum);
    else
        hipLaunchKernelGGL(k_ks_combine<TFHEpp::lvl10param>, dim3(njobs), dim3(256), 0, st, d_jobs,
                           njobs, G, base, rowsum);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace cufhe_amd
