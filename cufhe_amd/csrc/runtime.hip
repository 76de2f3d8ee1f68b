// Runtime behind the C ABI (include/cufhe_amd.h): device selection,
// transform tables, key upload/conversion, plain per-gate launches and
// batched plans (optionally captured in a HIP graph).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cufhe_amd.h"
#include "../../include/tfhe_params.hpp"
#include "br_fft.hpp"
#include "br_ntt.hpp"
#include "br_ntt64.hpp"
#include "engine.hpp"
#include "ks_gemm.hpp"
#include "launch.hpp"

namespace cufhe_amd {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

#define CUFHE_HIP_TRY(expr)                                                              \
    do {                                                                                 \
        const hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                            \
            return fail(CUFHE_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class P>
struct TableSet {
    cplx* twl = nullptr;
    cplx* tw64 = nullptr;
    cplx* cr = nullptr;
    cplx* tws = nullptr;
    cplx* e2n = nullptr;
};

// NTT-backend tables (Montgomery form), see ntt_wave.hpp.
struct NttTableSet {
    uint32_t* mem = nullptr;  // one allocation holding every table
    NttTables tb{};
};

// lvl2 NTT-backend tables (Goldilocks prime), see br_ntt64.hpp.
struct Ntt64TableSet {
    uint64_t* mem = nullptr;
    Ntt64Tables tb{};
};

template <class P>
struct LevelKeys {
    bool loaded = false;      // both keys present: the level's gates can run
    bool bk_loaded = false;   // the bootstrapping key (bk / bk_ntt / bk_ntt64)
    bool ksk_loaded = false;  // the key-switching key (ksk, ksg)
    uint64_t generation = 0;  // changes on every load: plans and servers check it
    bool kb = false;
    int backend = CUFHE_BACKEND_FFT;
    int variant = CUFHE_BR_PLAIN;
    bool subset = false;
    cplx* bk = nullptr;
    uint32_t* bk_ntt = nullptr;
    uint64_t* bk_ntt64 = nullptr;  // lvl2 NTT backend
    uint32_t* ksk = nullptr;
    KsGemmKey ksg;  // byte planes of ksk for the GEMM key switch
    FourierKeys<P> fk{};
    NttKeys nk{};
    Ntt64Keys n64{};
};

struct State {
    std::mutex mu;
    bool initialized = false;
    uint64_t generations = 0;  // key loads so far
    int device = 0;
    TableSet<lvl1param> t1;
    TableSet<lvl2param> t2;
    NttTableSet tn;
    Ntt64TableSet t64;
    LevelKeys<lvl1param> k1;
    LevelKeys<lvl2param> k2;
    cplx* ahk1 = nullptr;  // annihilation key spectra, level 1 and 2
    cplx* ahk2 = nullptr;
    uint64_t* ahk2_ntt = nullptr;  // level 2 on the NTT backend (exclusive with ahk2)
};

State& state()
{
    static State s;
    return s;
}

// exp(i*pi*t/N) for t in [0, 2N), computed in long double.
std::vector<cplx> e2n_table(int N)
{
    std::vector<cplx> e(2 * N);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < 2 * N; t++) {
        const long double a = pi * static_cast<long double>(t) / static_cast<long double>(N);
        e[t] = cplx{static_cast<double>(cosl(a)), static_cast<double>(sinl(a))};
    }
    // exact values on the axes
    e[0] = {1.0, 0.0};
    e[N / 2] = {0.0, 1.0};
    e[N] = {-1.0, 0.0};
    e[3 * N / 2] = {0.0, -1.0};
    // Store each root as the FP64 product of e[64 (t >> 6)] and e[t & 63],
    // rounded exactly as the device's cmul rounds it: kernels that rebuild
    // e[t] from those two short tables (in LDS) and kernels that gather e[t]
    // then see identical values. Entries t < 64 and t = 64u are unchanged.
    std::vector<cplx> p(2 * N);
    for (int t = 0; t < 2 * N; t++) {
        const cplx a = e[64 * (t >> 6)], b = e[t & 63];
        p[t] = cplx{std::fma(a.x, b.x, -a.y * b.y), std::fma(a.x, b.y, a.y * b.x)};
    }
    return p;
}

template <class P>
int build_tables(TableSet<P>& ts)
{
    constexpr int N = P::n, H = N / 2, R = H / 64;
    const std::vector<cplx> E = e2n_table(N);
    auto at = [&](long long t) { return E[static_cast<size_t>(((t % (2 * N)) + 2 * N) % (2 * N))]; };
    std::vector<cplx> twl(R * 64), tw64(64), cr(R), tws(64);
    // twl[k2*64+l] = exp(i*pi*l/N) * exp(-2*pi*i*l*k2/H) = E[l*(1 - 4*k2)]
    for (int k2 = 0; k2 < R; k2++)
        for (int l = 0; l < 64; l++) twl[k2 * 64 + l] = at(static_cast<long long>(l) * (1 - 4 * k2));
    // tw64[c*8+a] = exp(-2*pi*i*a*c/64) = E[-(2N/64)*a*c]
    for (int c = 0; c < 8; c++)
        for (int a = 0; a < 8; a++) tw64[c * 8 + a] = at(-static_cast<long long>(2 * N / 64) * a * c);
    // tws[l] = exp(-2*pi*i*l/H) = E[-4 l]
    for (int l = 0; l < 64; l++) tws[l] = at(-4LL * l);
    // cr[r] = exp(i*pi*64*r/N) = E[64 r]
    for (int r = 0; r < R; r++) cr[r] = at(64LL * r);
    auto up = [](cplx** dst, const std::vector<cplx>& v) -> hipError_t {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(cplx));
        if (e != hipSuccess) return e;
        return hipMemcpy(*dst, v.data(), v.size() * sizeof(cplx), hipMemcpyHostToDevice);
    };
    CUFHE_HIP_TRY(up(&ts.twl, twl));
    CUFHE_HIP_TRY(up(&ts.tw64, tw64));
    CUFHE_HIP_TRY(up(&ts.cr, cr));
    CUFHE_HIP_TRY(up(&ts.tws, tws));
    CUFHE_HIP_TRY(up(&ts.e2n, E));
    return CUFHE_OK;
}

// Host Montgomery helpers for P = 2147473409.
uint32_t host_mulmod(uint32_t a, uint32_t b)
{
    return static_cast<uint32_t>(static_cast<uint64_t>(a) * b % ntt31::P);
}
uint32_t host_to_mont(uint32_t x) { return static_cast<uint32_t>((static_cast<uint64_t>(x) << 32) % ntt31::P); }
// psi^e for any integer e (psi has order 2048)
uint32_t host_psi_pow(long long e) { return static_cast<uint32_t>(ntt31::psi_pow(e)); }

int build_ntt_tables(NttTableSet& ts)
{
    constexpr int N = lvl1param::n;
    // layout: twl[1024] twli[1024] tw64[64] tw64i[64] cr[16] cri[16] pw[2048]
    std::vector<uint32_t> h(1024 + 1024 + 64 + 64 + 16 + 16 + 2048);
    uint32_t* twl = h.data();
    uint32_t* twli = twl + 1024;
    uint32_t* tw64 = twli + 1024;
    uint32_t* tw64i = tw64 + 64;
    uint32_t* cr = tw64i + 64;
    uint32_t* cri = cr + 16;
    uint32_t* pw = cri + 16;
    for (int k2 = 0; k2 < 16; k2++)
        for (int l = 0; l < 64; l++) {
            const long long e = static_cast<long long>(l) * (1 + 2 * k2);
            twl[k2 * 64 + l] = host_to_mont(host_psi_pow(e));
            twli[k2 * 64 + l] = host_to_mont(host_psi_pow(-e));
        }
    // omega = psi^2, omega^(16 a c) = psi^(32 a c)
    for (int c = 0; c < 8; c++)
        for (int a = 0; a < 8; a++) {
            tw64[c * 8 + a] = host_to_mont(host_psi_pow(32LL * a * c));
            tw64i[c * 8 + a] = host_to_mont(host_psi_pow(-32LL * a * c));
        }
    const uint32_t ninv = static_cast<uint32_t>(ntt31::powmod(N, ntt31::P - 2));
    for (int r = 0; r < 16; r++) {
        cr[r] = host_to_mont(host_psi_pow(64LL * r));
        cri[r] = host_to_mont(host_mulmod(host_psi_pow(-64LL * r), ninv));
    }
    for (int t = 0; t < 2 * N; t++) pw[t] = host_to_mont(host_psi_pow(t));
    CUFHE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&ts.mem), h.size() * sizeof(uint32_t)));
    CUFHE_HIP_TRY(hipMemcpy(ts.mem, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    ts.tb.twl = ts.mem;
    ts.tb.twli = ts.mem + (twli - twl);
    ts.tb.tw64 = ts.mem + (tw64 - twl);
    ts.tb.tw64i = ts.mem + (tw64i - twl);
    ts.tb.cr = ts.mem + (cr - twl);
    ts.tb.cri = ts.mem + (cri - twl);
    ts.tb.pw = ts.mem + (pw - twl);
    return CUFHE_OK;
}

void free_ntt_tables(NttTableSet& ts)
{
    (void)hipFree(ts.mem);
    ts = NttTableSet{};
}

// Goldilocks arithmetic on the host (tables and key constants only).
uint64_t gold_mulmod(uint64_t a, uint64_t b)
{
    return static_cast<uint64_t>(static_cast<unsigned __int128>(a) * b % gold::P);
}
uint64_t gold_powmod(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    for (; e; e >>= 1, b = gold_mulmod(b, b))
        if (e & 1) r = gold_mulmod(r, b);
    return r;
}
// psi: primitive 4096-th root, ROOT_2_32^(2^32 / 4096)
uint64_t gold_psi() { return gold_powmod(gold::ROOT_2_32, 1ULL << 20); }

int build_ntt64_tables(Ntt64TableSet& ts)
{
    constexpr int N = kNtt64N;
    // layout: psi_rev[N] psi_inv_rev[N] pw[2N]
    std::vector<uint64_t> h(4 * N);
    const uint64_t psi = gold_psi(), psi_inv = gold_powmod(psi, gold::P - 2);
    for (int i = 0; i < N; i++) {
        uint32_t b = 0;
        for (int k = 0; k < 11; k++) b |= ((i >> k) & 1u) << (10 - k);
        h[i] = gold_powmod(psi, b);
        h[N + i] = gold_powmod(psi_inv, b);
    }
    for (int e = 0; e < 2 * N; e++) h[2 * N + e] = gold_powmod(psi, e);
    CUFHE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&ts.mem), h.size() * sizeof(uint64_t)));
    CUFHE_HIP_TRY(hipMemcpy(ts.mem, h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    ts.tb.psi_rev = ts.mem;
    ts.tb.psi_inv_rev = ts.mem + N;
    ts.tb.pw = ts.mem + 2 * N;
    ts.tb.n_inv = gold_powmod(N, gold::P - 2);
    return CUFHE_OK;
}

void free_ntt64_tables(Ntt64TableSet& ts)
{
    (void)hipFree(ts.mem);
    ts = Ntt64TableSet{};
}

template <class P>
FourierKeys<P> keys_view(const TableSet<P>& ts, const cplx* bk)
{
    FourierKeys<P> k;
    k.bk = bk;
    k.e2n = ts.e2n;
    k.fft.twl = ts.twl;
    k.fft.tw64 = ts.tw64;
    k.fft.cr = ts.cr;
    k.fft.tws = ts.tws;
    return k;
}

template <class P>
void free_tables(TableSet<P>& ts)
{
    (void)hipFree(ts.twl);
    (void)hipFree(ts.tw64);
    (void)hipFree(ts.cr);
    (void)hipFree(ts.tws);
    (void)hipFree(ts.e2n);
    ts = TableSet<P>{};
}

// One of a level's two keys, the other kept (BootstrappingKeyToNTT /
// KeySwitchingKeyToDevice and their Delete* counterparts).
template <class P>
void free_bk(LevelKeys<P>& k)
{
    (void)hipFree(k.bk);
    (void)hipFree(k.bk_ntt);
    (void)hipFree(k.bk_ntt64);
    k.bk = nullptr;
    k.bk_ntt = nullptr;
    k.bk_ntt64 = nullptr;
    k.fk = FourierKeys<P>{};
    k.nk = NttKeys{};
    k.n64 = Ntt64Keys{};
    k.bk_loaded = false;
    k.loaded = false;
}
template <class P>
void free_ksk(LevelKeys<P>& k)
{
    (void)hipFree(k.ksk);
    k.ksk = nullptr;
    ks_gemm_release(k.ksg);
    k.ksk_loaded = false;
    k.loaded = false;
}

template <class P>
void free_keys(LevelKeys<P>& k)
{
    (void)hipFree(k.bk);
    (void)hipFree(k.bk_ntt);
    (void)hipFree(k.bk_ntt64);
    (void)hipFree(k.ksk);
    ks_gemm_release(k.ksg);
    k = LevelKeys<P>{};
}

// ---- scratch of the entry point that allocates nothing -------------------
// cufhe_identity_key_switch_lvl10_noalloc keeps its job descriptors and the
// MFMA key switch's code words in this file-scope array (one copy per GPU,
// held until the process ends). Host code never uses the array's name as an
// address: the device address is looked up once per device with
// hipGetSymbolAddress (at cufhe_initialize, or at the first call on another
// device) and that pointer is what kernels receive.
constexpr int kNoAllocMaxJobs = 4096;
constexpr size_t kNoAllocJobBytes = (kNoAllocMaxJobs * sizeof(KsJob) + 255) & ~size_t(255);
constexpr size_t kNoAllocScratchBytes =
    kNoAllocJobBytes + static_cast<size_t>(kKsLvl10StepsMax) * kNoAllocMaxJobs * 4;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_on_current_device(void** out)
{
    int dev = 0;
    CUFHE_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CUFHE_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CUFHE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return CUFHE_OK;
}

// Job descriptors of n extracted lvl1 TLWEs key-switched to lvl0.
__global__ __launch_bounds__(256) void k_fill_ks_jobs(KsJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        jobs[i] = KsJob{in + static_cast<size_t>(i) * (lvl1param::n + 1), nullptr,
                        out + static_cast<size_t>(i) * (lvl0param::n + 1), 1, 0, 0};
}

int ensure_init()
{
    State& s = state();
    if (s.initialized) return CUFHE_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(CUFHE_ERR_NO_DEVICE, "no HIP device visible");
    CUFHE_HIP_TRY(hipSetDevice(s.device));
    {
        void* sp = nullptr;
        const int src = scratch_on_current_device(&sp);
        if (src != CUFHE_OK) return src;
    }
    int rc = build_tables(s.t1);
    if (rc != CUFHE_OK) return rc;
    rc = build_tables(s.t2);
    if (rc != CUFHE_OK) return rc;
    rc = build_ntt_tables(s.tn);
    if (rc != CUFHE_OK) return rc;
    rc = build_ntt64_tables(s.t64);
    if (rc != CUFHE_OK) return rc;
    s.initialized = true;
    return CUFHE_OK;
}

hipStream_t as_stream(void* st) { return static_cast<hipStream_t>(st); }

int check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CUFHE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return CUFHE_OK;
}

template <class brP>
int load_bk(LevelKeys<typename brP::targetP>& lk, const TableSet<typename brP::targetP>& ts,
            const typename brP::targetP::T* bk, size_t bk_words, bool kb)
{
    using P = typename brP::targetP;
    const size_t ntrgsw = TFHEpp::kBKTRGSWs<brP>(kb);
    const size_t want = ntrgsw * TFHEpp::kTRGSWWords<P>;
    if (bk_words != want)
        return fail(CUFHE_ERR_SIZE, "bootstrapping key has " + std::to_string(bk_words) +
                                        " words, expected " + std::to_string(want));
    const size_t npolys = ntrgsw * TFHEpp::BootstrappingTRGSWRows<P> * (P::k + 1);
    void* tmp = nullptr;
    CUFHE_HIP_TRY(hipMalloc(&tmp, want * sizeof(typename P::T)));
    hipError_t e = hipMemcpy(tmp, bk, want * sizeof(typename P::T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_HIP, std::string("bk upload: ") + hipGetErrorString(e));
    }
    e = hipMalloc(reinterpret_cast<void**>(&lk.bk), npolys * (P::n / 2) * sizeof(cplx));
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_NOMEM, "bootstrapping key spectra allocation failed");
    }
    lk.fk = keys_view(ts, lk.bk);
    launch_to_fourier<P>(lk.bk, tmp, static_cast<int>(npolys), false, lk.fk, nullptr);
    e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(CUFHE_ERR_HIP, std::string("bk conversion: ") + hipGetErrorString(e));
    lk.kb = kb;
    return CUFHE_OK;
}

// Annihilation key: nbit x l x 2 polynomials -> spectra in dst (allocated on
// the first load of the level).
template <class P>
int load_annihilate_key(cplx*& dst, const TableSet<P>& ts, const void* ahk, size_t words)
{
    constexpr size_t npolys = static_cast<size_t>(P::nbit) * P::l * 2;
    const size_t want = npolys * P::n;
    if (words != want)
        return fail(CUFHE_ERR_SIZE, "annihilation key has " + std::to_string(words) +
                                        " words, expected " + std::to_string(want));
    const size_t bytes = want * sizeof(typename P::T);
    void* tmp = nullptr;
    CUFHE_HIP_TRY(hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpy(tmp, ahk, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && !dst)
        e = hipMalloc(reinterpret_cast<void**>(&dst), npolys * (P::n / 2) * sizeof(cplx));
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_HIP, std::string("annihilation key upload: ") + hipGetErrorString(e));
    }
    launch_to_fourier<P>(dst, tmp, static_cast<int>(npolys), false, keys_view(ts, nullptr), nullptr);
    e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(CUFHE_ERR_HIP, std::string("annihilation key conversion: ") + hipGetErrorString(e));
    return CUFHE_OK;
}

// Annihilation key on the lvl2 NTT backend: torus_to_ntt_mod and the forward
// transform per polynomial (the bootstrapping key's launch_to_ntt64 layout).
int load_annihilate_key_ntt64(uint64_t*& dst, const Ntt64TableSet& ts, const void* ahk, size_t words)
{
    using P = lvl2param;
    constexpr size_t npolys = static_cast<size_t>(P::nbit) * P::l * 2;
    const size_t want = npolys * P::n;
    if (words != want)
        return fail(CUFHE_ERR_SIZE, "annihilation key has " + std::to_string(words) +
                                        " words, expected " + std::to_string(want));
    const size_t bytes = want * sizeof(uint64_t);
    void* tmp = nullptr;
    CUFHE_HIP_TRY(hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpy(tmp, ahk, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && !dst) e = hipMalloc(reinterpret_cast<void**>(&dst), bytes);
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_HIP, std::string("annihilation key upload: ") + hipGetErrorString(e));
    }
    launch_to_ntt64(dst, static_cast<const uint64_t*>(tmp), static_cast<int>(npolys), ts.tb, nullptr);
    e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess)
        return fail(CUFHE_ERR_HIP, std::string("annihilation key conversion: ") + hipGetErrorString(e));
    return CUFHE_OK;
}

// lvl1 key on the NTT backend: torus_to_ntt_mod + forward NTT per polynomial
// (reference src/bootstrap_gpu.cu:212-264, :426-450), stored as u32
// Montgomery residues in the transform's bin order.
int load_bk_ntt(LevelKeys<lvl1param>& lk, const NttTableSet& ts, const uint32_t* bk,
                size_t bk_words, int variant)
{
    using P = lvl1param;
    const size_t ntrgsw = TFHEpp::kBKTRGSWs<TFHEpp::lvl01param>(variant == CUFHE_BR_KEY_BUNDLE);
    const size_t want = ntrgsw * TFHEpp::kTRGSWWords<P>;
    if (bk_words != want)
        return fail(CUFHE_ERR_SIZE, "bootstrapping key has " + std::to_string(bk_words) +
                                        " words, expected " + std::to_string(want));
    const size_t npolys = ntrgsw * TFHEpp::BootstrappingTRGSWRows<P> * (P::k + 1);
    void* tmp = nullptr;
    CUFHE_HIP_TRY(hipMalloc(&tmp, want * sizeof(uint32_t)));
    hipError_t e = hipMemcpy(tmp, bk, want * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_HIP, std::string("bk upload: ") + hipGetErrorString(e));
    }
    e = hipMalloc(reinterpret_cast<void**>(&lk.bk_ntt), npolys * P::n * sizeof(uint32_t));
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_NOMEM, "bootstrapping key NTT allocation failed");
    }
    launch_to_ntt(lk.bk_ntt, static_cast<const uint32_t*>(tmp), static_cast<int>(npolys), ts.tb,
                  nullptr);
    e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(CUFHE_ERR_HIP, std::string("bk conversion: ") + hipGetErrorString(e));
    lk.nk.bk = lk.bk_ntt;
    lk.nk.tb = ts.tb;
    // gadget identity h_q = 2^(32 - (q+1) Bgbit) as torus_to_ntt_mod, Montgomery form
    for (int q = 0; q < static_cast<int>(P::l); q++) {
        const uint64_t h = 1ULL << (32 - (q + 1) * P::Bgbit);
        uint64_t v = (h * ntt31::P + (1ULL << 31)) >> 32;
        if (v >= ntt31::P) v -= ntt31::P;
        lk.nk.one_mont[q] = host_to_mont(static_cast<uint32_t>(v));
    }
    return CUFHE_OK;
}

// lvl2 key on the NTT backend: torus_to_ntt_mod (Goldilocks) + forward
// transform per polynomial (reference src/bootstrap_gpu.cu:213-340).
int load_bk_ntt64(LevelKeys<lvl2param>& lk, const Ntt64TableSet& ts, const uint64_t* bk,
                  size_t bk_words, bool kb)
{
    using P = lvl2param;
    const size_t ntrgsw = TFHEpp::kBKTRGSWs<TFHEpp::lvl02param>(kb);
    const size_t want = ntrgsw * TFHEpp::kTRGSWWords<P>;
    if (bk_words != want)
        return fail(CUFHE_ERR_SIZE, "bootstrapping key has " + std::to_string(bk_words) +
                                        " words, expected " + std::to_string(want));
    const size_t npolys = ntrgsw * TFHEpp::BootstrappingTRGSWRows<P> * (P::k + 1);
    void* tmp = nullptr;
    CUFHE_HIP_TRY(hipMalloc(&tmp, want * sizeof(uint64_t)));
    hipError_t e = hipMemcpy(tmp, bk, want * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_HIP, std::string("bk upload: ") + hipGetErrorString(e));
    }
    e = hipMalloc(reinterpret_cast<void**>(&lk.bk_ntt64), npolys * P::n * sizeof(uint64_t));
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(CUFHE_ERR_NOMEM, "bootstrapping key NTT allocation failed");
    }
    launch_to_ntt64(lk.bk_ntt64, static_cast<const uint64_t*>(tmp), static_cast<int>(npolys), ts.tb,
                    nullptr);
    e = hipDeviceSynchronize();
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(CUFHE_ERR_HIP, std::string("bk conversion: ") + hipGetErrorString(e));
    lk.n64.bk = lk.bk_ntt64;
    lk.n64.tb = ts.tb;
    // gadget identity h_q = 2^(64 - (q+1) Bgbit) as torus_to_ntt_mod
    for (int q = 0; q < static_cast<int>(P::l); q++)
        lk.n64.one[q] = torus64_to_gold(1ULL << (64 - (q + 1) * P::Bgbit));
    lk.kb = kb;
    return CUFHE_OK;
}

int load_ksk(uint32_t** dst, const uint32_t* ksk, size_t words, size_t want)
{
    if (words != want)
        return fail(CUFHE_ERR_SIZE, "key-switching key has " + std::to_string(words) +
                                        " words, expected " + std::to_string(want));
    CUFHE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), words * sizeof(uint32_t)));
    CUFHE_HIP_TRY(hipMemcpy(*dst, ksk, words * sizeof(uint32_t), hipMemcpyHostToDevice));
    return CUFHE_OK;
}

// One key of a level into the device state, the other key kept; the
// caller holds the state lock with the device idle. On failure the key is
// absent. The level's gates run once both keys are present (note_load).
int bk_load_lvl1(State& s, const uint32_t* bk, size_t words, int backend, int variant)
{
    free_bk(s.k1);
    const int rc = backend == CUFHE_BACKEND_NTT
                       ? load_bk_ntt(s.k1, s.tn, bk, words, variant)
                       : load_bk<TFHEpp::lvl01param>(s.k1, s.t1, bk, words,
                                                     variant == CUFHE_BR_KEY_BUNDLE);
    if (rc != CUFHE_OK) {
        free_bk(s.k1);
        return rc;
    }
    s.k1.backend = backend;
    s.k1.variant = variant;
    s.k1.kb = variant == CUFHE_BR_KEY_BUNDLE;
    s.k1.bk_loaded = true;
    return CUFHE_OK;
}
int bk_load_lvl2(State& s, const uint64_t* bk, size_t words, int backend, bool kb)
{
    free_bk(s.k2);
    const int rc = backend == CUFHE_BACKEND_NTT ? load_bk_ntt64(s.k2, s.t64, bk, words, kb)
                                                : load_bk<TFHEpp::lvl02param>(s.k2, s.t2, bk, words, kb);
    if (rc != CUFHE_OK) {
        free_bk(s.k2);
        return rc;
    }
    s.k2.backend = backend;
    s.k2.kb = kb;
    s.k2.bk_loaded = true;
    return CUFHE_OK;
}
template <class P>
int ksk_load(LevelKeys<P>& k, int level, const uint32_t* ksk, size_t words, bool subset)
{
    free_ksk(k);
    size_t want = TFHEpp::kKSKWords<TFHEpp::lvl20param>;
    if (level == 1) {
        using KS = TFHEpp::lvl10param;
        want = subset ? static_cast<size_t>(KS::domainP::n - KS::targetP::n) * KS::t *
                            (1u << (KS::basebit - 1)) * (KS::targetP::n + 1)
                      : TFHEpp::kKSKWords<KS>;
    }
    int rc = load_ksk(&k.ksk, ksk, words, want);
    if (rc == CUFHE_OK && ks_gemm_prepare(k.ksg, k.ksk, level, subset, nullptr) != 0)
        rc = fail(CUFHE_ERR_HIP, "key-switching key byte planes");
    if (rc != CUFHE_OK) {
        free_ksk(k);
        return rc;
    }
    k.subset = subset;
    k.ksk_loaded = true;
    return CUFHE_OK;
}
// A key of the level changed: plans and servers built before see a new
// generation.
template <class P>
void note_load(State& s, LevelKeys<P>& k)
{
    k.loaded = k.bk_loaded && k.ksk_loaded;
    k.generation = ++s.generations;
}

// Linear combination of a binary gate (reference src/bootstrap_gpu.cu
// :1335-1881 and :1905-1946): (ca, cb, offset in units of mu).
struct GateLin {
    int32_t ca, cb, off_mu;
};
bool gate_lin(int op, GateLin& g)
{
    switch (op) {
        case CUFHE_NAND: g = {-1, -1, 1}; return true;
        case CUFHE_OR: g = {1, 1, 1}; return true;
        case CUFHE_ORYN: g = {1, -1, 1}; return true;
        case CUFHE_ORNY: g = {-1, 1, 1}; return true;
        case CUFHE_AND: g = {1, 1, -1}; return true;
        case CUFHE_ANDYN: g = {1, -1, -1}; return true;
        case CUFHE_ANDNY: g = {-1, 1, -1}; return true;
        case CUFHE_NOR: g = {-1, -1, -1}; return true;
        case CUFHE_XOR: g = {2, 2, 2}; return true;
        case CUFHE_XNOR: g = {-2, -2, -2}; return true;
        default: return false;
    }
}

template <class T>
int alloc_upload(T** dst, const std::vector<T>& v)
{
    if (v.empty()) {
        *dst = nullptr;
        return CUFHE_OK;
    }
    CUFHE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T)));
    CUFHE_HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return CUFHE_OK;
}

// Batched lvl1 blind rotation + SampleExtract on the loaded backend.
void br_lvl1(const BrJob* jobs, int n, int c, uint32_t mu, hipStream_t st)
{
    const LevelKeys<lvl1param>& k = state().k1;
    if (n <= 0) return;
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_br_ntt(k.variant, jobs, n, c, mu, k.nk, st);
    else if (k.kb)
        launch_br_fft<lvl1param, true>(jobs, n, c, mu, k.fk, st);
    else
        launch_br_fft<lvl1param, false>(jobs, n, c, mu, k.fk, st);
}

// Batched lvl1 -> lvl0 identity key switch (full or subset key).
// ws: GEMM workspace for n jobs (ks_gemm_workspace_bytes) or null. Returns
// false only if the GEMM path failed to launch.
bool ks_lvl1(const KsJob* jobs, int n, void* ws, hipStream_t st)
{
    const LevelKeys<lvl1param>& k = state().k1;
    if (n <= 0) return true;
    if (ws && n >= kKsGemmMinJobs && k.ksg.planes) return ks_gemm_run(k.ksg, jobs, n, ws, st) == 0;
    launch_ks_batch(1, k.subset, jobs, n, k.ksk, st);
    return true;
}

// Jobs per workgroup of the lvl2 FFT rotation (1 or 2), cufhe_set_lvl2_jobs_per_group.
std::atomic<int> g_lvl2_c{1};

// Batched lvl2 blind rotation (+ SampleExtract or the TRLWE) on the loaded
// lvl2 backend.
void br_lvl2(const BrJob* jobs, int n, uint64_t mu, hipStream_t st)
{
    const LevelKeys<lvl2param>& k = state().k2;
    if (n <= 0) return;
    const int c = g_lvl2_c.load(std::memory_order_relaxed);
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_br_ntt64(k.kb, jobs, n, mu, k.n64, st);
    else if (k.kb)
        launch_br_fft<lvl2param, true>(jobs, n, c, mu, k.fk, st);
    else
        launch_br_fft<lvl2param, false>(jobs, n, c, mu, k.fk, st);
}

// One lvl02 gate (BR-IKS) as a single launch on the loaded lvl2 backend.
void gate_lvl02_one(bool mux, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                    const uint32_t* inc, int32_t ca, int32_t cb, uint32_t offset, int32_t mux_sign,
                    hipStream_t st)
{
    const LevelKeys<lvl2param>& k = state().k2;
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_gate_lvl02_ntt(k.kb, mux, out, in0, in1, inc, ca, cb, offset, mux_sign, k.ksk, k.n64, st);
    else
        launch_gate_lvl02_fft(k.kb, mux, out, in0, in1, inc, ca, cb, offset, mux_sign, k.ksk, k.fk, st);
}

bool ks_lvl2(const KsJob* jobs, int n, void* ws, hipStream_t st)
{
    const LevelKeys<lvl2param>& k = state().k2;
    if (n <= 0) return true;
    if (ws && n >= kKsGemmMinJobs && k.ksg.planes) return ks_gemm_run(k.ksg, jobs, n, ws, st) == 0;
    launch_ks_batch(2, false, jobs, n, k.ksk, st);
    return true;
}

// One lvl1 gate (IKS-BR) as a single launch on the loaded backend.
void gate_lvl1_one(bool mux, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                   const uint32_t* inc, int32_t ca, int32_t cb, uint32_t offset, int32_t mux_sign,
                   hipStream_t st)
{
    const LevelKeys<lvl1param>& k = state().k1;
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_gate_lvl1_ntt(k.variant, k.subset, mux, out, in0, in1, inc, ca, cb, offset, mux_sign,
                             k.ksk, k.nk, st);
    else
        launch_gate_lvl1_fft(k.kb, mux, out, in0, in1, inc, ca, cb, offset, mux_sign, k.ksk, k.fk,
                             st);
}

}  // namespace

int lvl1_count(const cufhe_gate_desc* gates, int n, int* nks, int* nmux)
{
    int k = 0, m = 0;
    for (int i = 0; i < n; i++) {
        const int op = gates[i].op;
        if (op < 0 || op >= CUFHE_GATE_COUNT)
            return set_error(CUFHE_ERR_ARG, "unknown gate op " + std::to_string(op));
        if (op == CUFHE_MUX || op == CUFHE_NMUX) {
            k += 2;
            m++;
        }
        else if (op != CUFHE_NOT && op != CUFHE_COPY) {
            k += 1;
        }
    }
    *nks = k;
    *nmux = m;
    return CUFHE_OK;
}

// Gate list -> jobs (reference src/bootstrap_gpu.cu:1223-1246 IKS-BR gates,
// :2069-2197 MUX, :2200-2222 NOT/COPY): a binary gate is one key switch of
// ca*in0 + cb*in1 + offset followed by one rotation; MUX/NMUX are two
// (c + t - mu, -c + f - mu) combined as +-(r1 + r0 + mu); NOT/COPY are
// coefficient-wise.
void lvl1_build(const cufhe_gate_desc* gates, int n, uint32_t* d_pool, uint32_t* ws, Lvl1Jobs& jobs)
{
    constexpr uint32_t W1 = kWireWords1, W0 = kLvl0Stride;
    constexpr uint32_t mu = static_cast<uint32_t>(lvl1param::μ);
    int nks = 0, nmux = 0;
    (void)lvl1_count(gates, n, &nks, &nmux);
    uint32_t* lvl0 = ws;
    uint32_t* res = ws + static_cast<size_t>(nks) * W0;
    int ks_i = 0, mux_i = 0;
    auto wire = [&](uint32_t w) { return d_pool + static_cast<size_t>(w) * W1; };
    for (int i = 0; i < n; i++) {
        const cufhe_gate_desc& gd = gates[i];
        if (gd.op == CUFHE_NOT || gd.op == CUFHE_COPY) {
            jobs.lin.push_back(LinJob{wire(gd.in0), nullptr, wire(gd.out),
                                      gd.op == CUFHE_NOT ? -1 : 1, 0, 0u, static_cast<int32_t>(W1)});
            continue;
        }
        if (gd.op == CUFHE_MUX || gd.op == CUFHE_NMUX) {
            uint32_t* r1 = res + static_cast<size_t>(2 * mux_i) * W1;
            uint32_t* r0 = r1 + W1;
            mux_i++;
            const int sgn = gd.op == CUFHE_MUX ? 1 : -1;
            for (int half = 0; half < 2; half++) {
                uint32_t* l0 = lvl0 + static_cast<size_t>(ks_i++) * W0;
                jobs.ks.push_back(KsJob{wire(gd.inc), wire(half == 0 ? gd.in1 : gd.in0), l0,
                                        half == 0 ? 1 : -1, 1, static_cast<uint64_t>(0u - mu)});
                jobs.br.push_back(BrJob{l0, nullptr, half == 0 ? r1 : r0, 1, 0, 0u, 0u});
            }
            jobs.lin.push_back(LinJob{r1, r0, wire(gd.out), sgn, sgn,
                                      static_cast<uint32_t>(sgn) * mu, static_cast<int32_t>(W1)});
            continue;
        }
        GateLin gl;
        gate_lin(gd.op, gl);
        uint32_t* l0 = lvl0 + static_cast<size_t>(ks_i++) * W0;
        jobs.ks.push_back(KsJob{wire(gd.in0), wire(gd.in1), l0, gl.ca, gl.cb,
                                static_cast<uint64_t>(static_cast<uint32_t>(gl.off_mu) * mu)});
        jobs.br.push_back(BrJob{l0, nullptr, wire(gd.out), 1, 0, 0u, 0u});
    }
}

bool lvl1_loaded() { return state().k1.loaded; }

uint64_t lvl1_key_generation()
{
    const State& s = state();
    return s.k1.loaded ? s.k1.generation : 0;
}

// Autotuner results: batch size (bootstraps) -> fastest jobs-per-workgroup
// choice, for the lvl1 key load `generation` (cufhe_autotune_lvl1).
struct TuneTable {
    std::mutex mu;
    uint64_t generation = 0;
    std::map<int, int> best;
};
TuneTable& tune_table()
{
    static TuneTable t;
    return t;
}

// Recorded choice for a batch of nbr bootstraps: the entry of the smallest
// tuned size >= nbr, else the largest tuned size; 0 if none.
int tuned_c(int nbr)
{
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    const LevelKeys<lvl1param>& k = state().k1;
    if (!k.loaded || t.generation != k.generation || t.best.empty() || nbr <= 0) return 0;
    auto it = t.best.lower_bound(nbr);
    if (it == t.best.end()) --it;
    return it->second;
}

int lvl1_pick_c(int c, int nbr)
{
    const LevelKeys<lvl1param>& k = state().k1;
    if (c == 0) c = tuned_c(nbr);
    if (c == 0) {
        if (k.backend == CUFHE_BACKEND_NTT) {
            // NTT block-binary sweep (4096 gates, BR ms): C=1 18.1, 2 17.3,
            // 4 20.3, 6 (shared key) 21.6 -- the 4-byte NTT key is not worth
            // the ring's barriers
            c = 2;
        }
        else if (k.kb) {
            // The one-wave-per-job kernel spreads the jobs over all CUs (at
            // most 8 per workgroup; waves without a job only stream the key)
            // and is the fastest at every batch size (MI355X, BR ms, C = 6 /
            // 8): 256 jobs 5.84 / 3.54, 1024: 5.84 / 4.27, 1536: 6.19 / 5.55,
            // 2048: 12.05 / 6.41, 3072: 11.73 / 11.11, 4096: 17.71 / 12.41.
            c = 8;
        }
        else {
            c = 1;
        }
    }
    // 6: the kernels that share key chunks among six jobs through LDS (NTT
    // backend, all modes; FFT backend, key bundle); 8 with the key bundle:
    // one wave per job, key shared by eight
    if (c == 6) return (k.backend == CUFHE_BACKEND_NTT || k.kb) ? 6 : -1;
    if (c != 1 && c != 2 && c != 4 && c != 8) return -1;
    return c;
}

size_t lvl1_ks_workspace_bytes(int nks)
{
    const LevelKeys<lvl1param>& k = state().k1;
    return (nks >= kKsGemmMinJobs && k.ksg.planes) ? ks_gemm_workspace_bytes(k.ksg, nks) : 0;
}

bool lvl1_launch(const KsJob* ks, int nks, void* ks_ws, const BrJob* br, int nbr,
                 const LinJob* lin, int nlin, int c, hipStream_t st)
{
    const bool ok = ks_lvl1(ks, nks, ks_ws, st);
    br_lvl1(br, nbr, c, static_cast<uint32_t>(lvl1param::μ), st);
    if (nlin) launch_linear(lin, nlin, st);
    return ok;
}

}  // namespace cufhe_amd

using namespace cufhe_amd;

struct cufhe_plan {
    int level = 1;  // 1: lvl1 gates (IKS-BR); 2: lvl02 gates (BR-IKS); 3: lvl0 gates (BR-IKS at lvl1)
    uint64_t key_generation = 0;  // the key load the plan was built for
    int njobs_ks = 0, njobs_br = 0, njobs_lin = 0;
    int c = 2;
    bool use_graph = false;
    KsJob* ks = nullptr;
    BrJob* br = nullptr;
    LinJob* lin = nullptr;
    uint32_t* ws = nullptr;  // lvl0 TLWEs + MUX partial results
    void* ks_ws = nullptr;   // GEMM key-switch workspace (large batches)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

namespace {

// GEMM key-switch workspace of a plan (batches of kKsGemmMinJobs or more).
int plan_alloc_ks_ws(cufhe_plan* p)
{
    State& s = state();
    p->key_generation = p->level == 2 ? s.k2.generation : s.k1.generation;
    const KsGemmKey& key = p->level == 2 ? s.k2.ksg : s.k1.ksg;
    if (p->njobs_ks < kKsGemmMinJobs || !key.planes) return CUFHE_OK;
    if (hipMalloc(&p->ks_ws, ks_gemm_workspace_bytes(key, p->njobs_ks)) != hipSuccess)
        return fail(CUFHE_ERR_NOMEM, "key-switch workspace allocation failed");
    return CUFHE_OK;
}

int plan_launch(cufhe_plan* p, hipStream_t st)
{
    State& s = state();
    if (p->level == 3) {
        br_lvl1(p->br, p->njobs_br, p->c, static_cast<uint32_t>(lvl1param::μ), st);
        if (!ks_lvl1(p->ks, p->njobs_ks, p->ks_ws, st))
            return fail(CUFHE_ERR_HIP, "GEMM key switch failed");
        if (p->njobs_lin) launch_linear(p->lin, p->njobs_lin, st);
        return check_launch("batched lvl0 gate launch");
    }
    if (p->level == 2) {
        // BR-IKS order: rotations (pre-added lvl0 inputs) -> lvl2 TLWEs ->
        // key switch (MUX halves combined in its pre-add)
        br_lvl2(p->br, p->njobs_br, static_cast<uint64_t>(lvl2param::μ), st);
        if (!ks_lvl2(p->ks, p->njobs_ks, p->ks_ws, st))
            return fail(CUFHE_ERR_HIP, "GEMM key switch failed");
        if (p->njobs_lin) launch_linear(p->lin, p->njobs_lin, st);
        return check_launch("batched lvl02 gate launch");
    }
    if (!lvl1_launch(p->ks, p->njobs_ks, p->ks_ws, p->br, p->njobs_br, p->lin, p->njobs_lin, p->c, st))
        return fail(CUFHE_ERR_HIP, "GEMM key switch failed");
    return check_launch("batched gate launch");
}

// The plan's kernels, job arrays and workspace were sized for the keys loaded
// when it was created; a reload (even of the same sizes) invalidates it.
int plan_check_keys(const cufhe_plan* p)
{
    const State& s = state();
    const LevelKeys<lvl1param>& k1 = s.k1;
    if (p->level == 2) {
        if (!s.k2.loaded) return fail(CUFHE_ERR_STATE, "lvl02 keys not loaded");
        if (s.k2.generation != p->key_generation)
            return fail(CUFHE_ERR_STATE, "lvl02 keys were reloaded after the plan was created");
        return CUFHE_OK;
    }
    if (!k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (k1.generation != p->key_generation)
        return fail(CUFHE_ERR_STATE, "lvl1 keys were reloaded after the plan was created");
    return CUFHE_OK;
}

int pick_c(int c, bool kb, int nbr)
{
    (void)kb;
    return lvl1_pick_c(c, nbr);
}

}  // namespace

extern "C" {

const char* cufhe_version(void) { return "cufhe_amd 0.1 (gfx950)"; }
const char* cufhe_last_error(void) { return g_last_error.c_str(); }

int cufhe_set_gpu_num(int gpu_num)
{
    if (gpu_num != 1) return fail(CUFHE_ERR_UNSUPPORTED, "one process drives one GPU: gpu_num must be 1");
    return CUFHE_OK;
}

int cufhe_initialize(void)
{
    std::lock_guard<std::mutex> g(state().mu);
    return ensure_init();
}

int cufhe_initialize_lvl01(const uint32_t* bk, size_t bk_words, int key_bundle,
                           const uint32_t* ksk, size_t ksk_words)
{
    return cufhe_initialize_lvl01_ex(bk, bk_words, CUFHE_BACKEND_FFT,
                                     key_bundle ? CUFHE_BR_KEY_BUNDLE : CUFHE_BR_PLAIN, ksk,
                                     ksk_words, 0);
}

int cufhe_initialize_lvl01_ex(const uint32_t* bk, size_t bk_words, int backend, int variant,
                              const uint32_t* ksk, size_t ksk_words, int subset_ksk)
{
    if (!bk || !ksk) return fail(CUFHE_ERR_ARG, "null key pointer");
    if (backend != CUFHE_BACKEND_FFT && backend != CUFHE_BACKEND_NTT)
        return fail(CUFHE_ERR_ARG, "unknown backend " + std::to_string(backend));
    if (variant < CUFHE_BR_PLAIN || variant > CUFHE_BR_BLOCK_BINARY)
        return fail(CUFHE_ERR_ARG, "unknown blind-rotation variant " + std::to_string(variant));
    if (backend == CUFHE_BACKEND_FFT && variant == CUFHE_BR_BLOCK_BINARY)
        return fail(CUFHE_ERR_UNSUPPORTED, "block-binary blind rotation needs the NTT backend");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = ensure_init();
    if (rc != CUFHE_OK) return rc;
    (void)hipDeviceSynchronize();
    free_keys(s.k1);
    rc = bk_load_lvl1(s, bk, bk_words, backend, variant);
    if (rc == CUFHE_OK) rc = ksk_load(s.k1, 1, ksk, ksk_words, subset_ksk != 0);
    if (rc != CUFHE_OK) {
        free_keys(s.k1);
        return rc;
    }
    note_load(s, s.k1);
    return CUFHE_OK;
}

int cufhe_initialize_lvl02(const uint64_t* bk, size_t bk_words, int key_bundle,
                           const uint32_t* ksk, size_t ksk_words)
{
    return cufhe_initialize_lvl02_ex(bk, bk_words, CUFHE_BACKEND_FFT, key_bundle, ksk, ksk_words);
}

int cufhe_initialize_lvl02_ex(const uint64_t* bk, size_t bk_words, int backend, int key_bundle,
                              const uint32_t* ksk, size_t ksk_words)
{
    if (!bk || !ksk) return fail(CUFHE_ERR_ARG, "null key pointer");
    if (backend != CUFHE_BACKEND_FFT && backend != CUFHE_BACKEND_NTT)
        return fail(CUFHE_ERR_ARG, "unknown backend " + std::to_string(backend));
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = ensure_init();
    if (rc != CUFHE_OK) return rc;
    (void)hipDeviceSynchronize();
    free_keys(s.k2);
    rc = bk_load_lvl2(s, bk, bk_words, backend, key_bundle != 0);
    if (rc == CUFHE_OK) rc = ksk_load(s.k2, 2, ksk, ksk_words, false);
    if (rc != CUFHE_OK) {
        free_keys(s.k2);
        return rc;
    }
    note_load(s, s.k2);
    return CUFHE_OK;
}

int cufhe_bk_load(int level, const void* bk, size_t words, int backend, int variant)
{
    if (!bk) return fail(CUFHE_ERR_ARG, "null key pointer");
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "level must be 1 or 2");
    if (backend != CUFHE_BACKEND_FFT && backend != CUFHE_BACKEND_NTT)
        return fail(CUFHE_ERR_ARG, "unknown backend " + std::to_string(backend));
    if (variant < CUFHE_BR_PLAIN || variant > CUFHE_BR_BLOCK_BINARY)
        return fail(CUFHE_ERR_ARG, "unknown blind-rotation variant " + std::to_string(variant));
    if (variant == CUFHE_BR_BLOCK_BINARY && (level == 2 || backend == CUFHE_BACKEND_FFT))
        return fail(CUFHE_ERR_UNSUPPORTED, "block-binary blind rotation is lvl1 on the NTT backend");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = ensure_init();
    if (rc != CUFHE_OK) return rc;
    (void)hipDeviceSynchronize();
    if (level == 1) {
        rc = bk_load_lvl1(s, static_cast<const uint32_t*>(bk), words, backend, variant);
        note_load(s, s.k1);
    } else {
        rc = bk_load_lvl2(s, static_cast<const uint64_t*>(bk), words, backend,
                          variant == CUFHE_BR_KEY_BUNDLE);
        note_load(s, s.k2);
    }
    return rc;
}

int cufhe_bk_delete(int level)
{
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "level must be 1 or 2");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.initialized) return CUFHE_OK;
    (void)hipDeviceSynchronize();
    if (level == 1) {
        free_bk(s.k1);
        note_load(s, s.k1);
    } else {
        free_bk(s.k2);
        note_load(s, s.k2);
    }
    return CUFHE_OK;
}

int cufhe_ksk_load(int level, const uint32_t* ksk, size_t words, int subset)
{
    if (!ksk) return fail(CUFHE_ERR_ARG, "null key pointer");
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "level must be 1 or 2");
    if (level == 2 && subset) return fail(CUFHE_ERR_UNSUPPORTED, "the subset key is lvl10 only");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = ensure_init();
    if (rc != CUFHE_OK) return rc;
    (void)hipDeviceSynchronize();
    if (level == 1) {
        rc = ksk_load(s.k1, 1, ksk, words, subset != 0);
        note_load(s, s.k1);
    } else {
        rc = ksk_load(s.k2, 2, ksk, words, false);
        note_load(s, s.k2);
    }
    return rc;
}

int cufhe_ksk_delete(int level)
{
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "level must be 1 or 2");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.initialized) return CUFHE_OK;
    (void)hipDeviceSynchronize();
    if (level == 1) {
        free_ksk(s.k1);
        note_load(s, s.k1);
    } else {
        free_ksk(s.k2);
        note_load(s, s.k2);
    }
    return CUFHE_OK;
}

int cufhe_cleanup(void)
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.initialized) return CUFHE_OK;
    (void)hipDeviceSynchronize();
    free_keys(s.k1);
    free_keys(s.k2);
    (void)hipFree(s.ahk1);
    (void)hipFree(s.ahk2);
    (void)hipFree(s.ahk2_ntt);
    s.ahk1 = s.ahk2 = nullptr;
    s.ahk2_ntt = nullptr;
    free_tables(s.t1);
    free_tables(s.t2);
    free_ntt_tables(s.tn);
    free_ntt64_tables(s.t64);
    s.initialized = false;
    return CUFHE_OK;
}

int cufhe_cleanup_lvl02(void)
{
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    (void)hipDeviceSynchronize();
    free_keys(s.k2);
    return CUFHE_OK;
}

int cufhe_synchronize(void)
{
    CUFHE_HIP_TRY(hipDeviceSynchronize());
    return CUFHE_OK;
}

int cufhe_stream_create(void** stream)
{
    if (!stream) return fail(CUFHE_ERR_ARG, "null stream handle");
    hipStream_t st;
    CUFHE_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *stream = st;
    return CUFHE_OK;
}

int cufhe_stream_destroy(void* stream)
{
    if (!stream) return CUFHE_OK;
    CUFHE_HIP_TRY(hipStreamDestroy(as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_event_create(void** event)
{
    if (!event) return fail(CUFHE_ERR_ARG, "null event handle");
    hipEvent_t e;
    CUFHE_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return CUFHE_OK;
}

int cufhe_event_destroy(void* event)
{
    if (!event) return CUFHE_OK;
    CUFHE_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return CUFHE_OK;
}

int cufhe_event_record(void* event, void* stream)
{
    if (!event) return fail(CUFHE_ERR_ARG, "null event");
    CUFHE_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_event_elapsed_ms(void* start, void* end, float* ms)
{
    if (!start || !end || !ms) return fail(CUFHE_ERR_ARG, "null pointer");
    CUFHE_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(end)));
    CUFHE_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
    return CUFHE_OK;
}

int cufhe_stream_query(void* stream)
{
    const hipError_t e = hipStreamQuery(as_stream(stream));
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(CUFHE_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int cufhe_stream_synchronize(void* stream)
{
    CUFHE_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_device_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(CUFHE_ERR_ARG, "null pointer");
    if (hipMalloc(ptr, bytes) != hipSuccess) return fail(CUFHE_ERR_NOMEM, "hipMalloc failed");
    return CUFHE_OK;
}

int cufhe_device_free(void* ptr)
{
    CUFHE_HIP_TRY(hipFree(ptr));
    return CUFHE_OK;
}

int cufhe_host_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(CUFHE_ERR_ARG, "null pointer");
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess)
        return fail(CUFHE_ERR_NOMEM, "hipHostMalloc failed");
    return CUFHE_OK;
}

int cufhe_host_free(void* ptr)
{
    CUFHE_HIP_TRY(hipHostFree(ptr));
    return CUFHE_OK;
}

int cufhe_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    CUFHE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    CUFHE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream)
{
    CUFHE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_memset_d(void* dst, int value, size_t bytes, void* stream)
{
    CUFHE_HIP_TRY(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return CUFHE_OK;
}

int cufhe_gate_lvl1(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                    const uint32_t* inc, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    if (!out || !in0) return fail(CUFHE_ERR_ARG, "null ciphertext pointer");
    const hipStream_t st = as_stream(stream);
    constexpr uint32_t mu = static_cast<uint32_t>(lvl1param::μ);
    if (op == CUFHE_NOT || op == CUFHE_COPY) {
        launch_linear_one(LinJob{in0, nullptr, out, op == CUFHE_NOT ? -1 : 1, 0, 0u,
                                 static_cast<int32_t>(lvl1param::n + 1)},
                          st);
        return check_launch("not/copy");
    }
    if (op == CUFHE_MUX || op == CUFHE_NMUX) {
        if (!in1 || !inc) return fail(CUFHE_ERR_ARG, "null MUX operand");
        gate_lvl1_one(true, out, in0, in1, inc, 0, 0, 0u, op == CUFHE_MUX ? 1 : -1, st);
        return check_launch("mux gate");
    }
    GateLin gl;
    if (!gate_lin(op, gl)) return fail(CUFHE_ERR_ARG, "unknown gate op " + std::to_string(op));
    if (!in1) return fail(CUFHE_ERR_ARG, "null second operand");
    gate_lvl1_one(false, out, in0, in1, nullptr, gl.ca, gl.cb,
                  static_cast<uint32_t>(gl.off_mu) * mu, 0, st);
    return check_launch("gate");
}

int cufhe_gate_lvl02(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                     const uint32_t* inc, void* stream)
{
    State& s = state();
    if (!s.k2.loaded) return fail(CUFHE_ERR_STATE, "lvl02 keys not loaded (cufhe_initialize_lvl02)");
    if (!out || !in0) return fail(CUFHE_ERR_ARG, "null ciphertext pointer");
    const hipStream_t st = as_stream(stream);
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    if (op == CUFHE_NOT || op == CUFHE_COPY) {
        launch_linear_one(LinJob{in0, nullptr, out, op == CUFHE_NOT ? -1 : 1, 0, 0u,
                                 static_cast<int32_t>(lvl0param::n + 1)},
                          st);
        return check_launch("not/copy lvl0");
    }
    if (op == CUFHE_MUX || op == CUFHE_NMUX) {
        if (!in1 || !inc) return fail(CUFHE_ERR_ARG, "null MUX operand");
        gate_lvl02_one(true, out, in0, in1, inc, 0, 0, 0u, op == CUFHE_MUX ? 1 : -1, st);
        return check_launch("mux gate lvl02");
    }
    GateLin gl;
    if (!gate_lin(op, gl)) return fail(CUFHE_ERR_ARG, "unknown gate op " + std::to_string(op));
    if (!in1) return fail(CUFHE_ERR_ARG, "null second operand");
    gate_lvl02_one(false, out, in0, in1, nullptr, gl.ca, gl.cb, static_cast<uint32_t>(gl.off_mu) * mu0,
                   0, st);
    return check_launch("gate lvl02");
}

int cufhe_plan_create_lvl1(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                           int jobs_per_group, int use_graph, cufhe_plan** out)
{
    State& s = state();
    if (!out) return fail(CUFHE_ERR_ARG, "null plan handle");
    *out = nullptr;
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    if (n < 0 || (n > 0 && (!gates || !d_pool))) return fail(CUFHE_ERR_ARG, "bad gate list");
    int nks = 0, nmux = 0;
    int rc = lvl1_count(gates, n, &nks, &nmux);
    if (rc != CUFHE_OK) return rc;
    const int c = pick_c(jobs_per_group, s.k1.kb, nks);
    if (c < 0)
        return fail(CUFHE_ERR_UNSUPPORTED,
                    "jobs_per_group must be 0, 1, 2, 4 or 8 (6: FFT key-bundle keys only)");
    cufhe_plan* p = new cufhe_plan();
    p->c = c;
    p->use_graph = use_graph != 0;
    const size_t ws_words = lvl1_workspace_words(nks, nmux);
    if (ws_words) {
        if (hipMalloc(reinterpret_cast<void**>(&p->ws), ws_words * sizeof(uint32_t)) != hipSuccess) {
            delete p;
            return fail(CUFHE_ERR_NOMEM, "plan workspace allocation failed");
        }
    }
    Lvl1Jobs jobs;
    jobs.ks.reserve(nks);
    jobs.br.reserve(nks);
    lvl1_build(gates, n, d_pool, p->ws, jobs);
    rc = alloc_upload(&p->ks, jobs.ks);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->br, jobs.br);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->lin, jobs.lin);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    p->njobs_ks = static_cast<int>(jobs.ks.size());
    p->njobs_br = static_cast<int>(jobs.br.size());
    p->njobs_lin = static_cast<int>(jobs.lin.size());
    rc = plan_alloc_ks_ws(p);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    *out = p;
    return CUFHE_OK;
}

int cufhe_plan_create_lvl01(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                            int jobs_per_group, int use_graph, cufhe_plan** out)
{
    State& s = state();
    if (!out) return fail(CUFHE_ERR_ARG, "null plan handle");
    *out = nullptr;
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    if (n < 0 || (n > 0 && (!gates || !d_pool))) return fail(CUFHE_ERR_ARG, "bad gate list");
    constexpr uint32_t W0 = lvl0param::n + 1;  // lvl0 wire stride
    constexpr uint32_t W1 = lvl1param::n + 1;  // rotated lvl1 TLWE in the workspace
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    constexpr uint32_t mu1 = static_cast<uint32_t>(lvl1param::μ);
    int nbr = 0, nmux = 0;
    int rc = lvl1_count(gates, n, &nbr, &nmux);
    if (rc != CUFHE_OK) return rc;
    const int c = pick_c(jobs_per_group, s.k1.kb, nbr);
    if (c < 0)
        return fail(CUFHE_ERR_UNSUPPORTED,
                    "jobs_per_group must be 0, 1, 2, 4 or 8 (6: FFT key-bundle keys only)");
    cufhe_plan* p = new cufhe_plan();
    p->level = 3;
    p->c = c;
    p->use_graph = use_graph != 0;
    if (nbr) {
        if (hipMalloc(reinterpret_cast<void**>(&p->ws),
                      static_cast<size_t>(nbr) * W1 * sizeof(uint32_t)) != hipSuccess) {
            delete p;
            return fail(CUFHE_ERR_NOMEM, "plan workspace allocation failed");
        }
    }
    std::vector<KsJob> ks;
    std::vector<BrJob> br;
    std::vector<LinJob> lin;
    auto wire = [&](uint32_t w) { return d_pool + static_cast<size_t>(w) * W0; };
    for (int i = 0; i < n; i++) {
        const cufhe_gate_desc& gd = gates[i];
        if (gd.op == CUFHE_NOT || gd.op == CUFHE_COPY) {
            lin.push_back(LinJob{wire(gd.in0), nullptr, wire(gd.out), gd.op == CUFHE_NOT ? -1 : 1, 0,
                                 0u, static_cast<int32_t>(W0)});
            continue;
        }
        if (gd.op == CUFHE_MUX || gd.op == CUFHE_NMUX) {
            uint32_t* r1 = p->ws + br.size() * W1;
            br.push_back(BrJob{wire(gd.inc), wire(gd.in1), r1, 1, 1, 0u - mu0, 0u});
            uint32_t* r0 = p->ws + br.size() * W1;
            br.push_back(BrJob{wire(gd.inc), wire(gd.in0), r0, -1, 1, 0u - mu0, 0u});
            const int sg = gd.op == CUFHE_MUX ? 1 : -1;
            ks.push_back(KsJob{r1, r0, wire(gd.out), sg, sg,
                               static_cast<uint64_t>(sg > 0 ? mu1 : 0u - mu1)});
            continue;
        }
        GateLin gl;
        gate_lin(gd.op, gl);
        uint32_t* r = p->ws + br.size() * W1;
        br.push_back(BrJob{wire(gd.in0), wire(gd.in1), r, gl.ca, gl.cb,
                           static_cast<uint32_t>(gl.off_mu) * mu0, 0u});
        ks.push_back(KsJob{r, nullptr, wire(gd.out), 1, 0, 0});
    }
    rc = alloc_upload(&p->ks, ks);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->br, br);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->lin, lin);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    p->njobs_ks = static_cast<int>(ks.size());
    p->njobs_br = static_cast<int>(br.size());
    p->njobs_lin = static_cast<int>(lin.size());
    rc = plan_alloc_ks_ws(p);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    *out = p;
    return CUFHE_OK;
}

int cufhe_gate_lvl0(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                    const uint32_t* inc, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    if (!out || !in0) return fail(CUFHE_ERR_ARG, "null ciphertext pointer");
    const hipStream_t st = as_stream(stream);
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    const LevelKeys<lvl1param>& k = s.k1;
    if (op == CUFHE_NOT || op == CUFHE_COPY) {
        launch_linear_one(LinJob{in0, nullptr, out, op == CUFHE_NOT ? -1 : 1, 0, 0u,
                                 static_cast<int32_t>(lvl0param::n + 1)},
                          st);
        return check_launch("not/copy lvl0");
    }
    const bool mux = op == CUFHE_MUX || op == CUFHE_NMUX;
    GateLin gl{0, 0, 0};
    if (mux) {
        if (!in1 || !inc) return fail(CUFHE_ERR_ARG, "null MUX operand");
    }
    else {
        if (!gate_lin(op, gl)) return fail(CUFHE_ERR_ARG, "unknown gate op " + std::to_string(op));
        if (!in1) return fail(CUFHE_ERR_ARG, "null second operand");
    }
    const uint32_t off = static_cast<uint32_t>(gl.off_mu) * mu0;
    const int sign = op == CUFHE_NMUX ? -1 : 1;
    constexpr uint32_t mu1 = static_cast<uint32_t>(lvl1param::μ);
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_gate_lvl01_ntt(k.variant, k.subset, mux, out, in0, in1, inc, gl.ca, gl.cb, off, sign,
                              mu1, k.ksk, k.nk, st);
    else
        launch_gate_lvl01_fft(k.kb, mux, k.subset, out, in0, in1, inc, gl.ca, gl.cb, off, sign, mu1,
                              k.ksk, k.fk, st);
    return check_launch("lvl0 gate");
}

int cufhe_bootstrap(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null ciphertext pointer");
    const LevelKeys<lvl1param>& k = s.k1;
    const hipStream_t st = as_stream(stream);
    // the lvl0 gate kernel with (ca, cb, offset) = (1, 0, 0): rotation of the
    // input itself with test vector +-mu, SampleExtract, key switch to lvl0
    if (k.backend == CUFHE_BACKEND_NTT)
        launch_gate_lvl01_ntt(k.variant, k.subset, false, d_out, d_in, d_in, nullptr, 1, 0, 0u, 1, mu,
                              k.ksk, k.nk, st);
    else
        launch_gate_lvl01_fft(k.kb, false, k.subset, d_out, d_in, d_in, nullptr, 1, 0, 0u, 1, mu,
                              k.ksk, k.fk, st);
    return check_launch("bootstrap");
}

int cufhe_sample_extract_key_switch_lvl10(uint32_t* d_out, const uint32_t* d_trlwe, int n,
                                          void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_trlwe) return fail(CUFHE_ERR_ARG, "null pointer");
    launch_se_ks_lvl10(s.k1.subset, d_out, d_trlwe, n, s.k1.ksk, as_stream(stream));
    return check_launch("sample extract + key switch");
}

int cufhe_bootstrap_to_trlwe_lvl01(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, int n,
                                   void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const hipStream_t st = as_stream(stream);
    std::vector<BrJob> jobs(n);
    for (int i = 0; i < n; i++)
        jobs[i] = BrJob{d_in + static_cast<size_t>(i) * (lvl0param::n + 1), nullptr,
                        d_out + static_cast<size_t>(i) * 2 * lvl1param::n, 1, 0, 0u, kBrOutTrlwe};
    BrJob* d = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(BrJob), st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d, jobs.data(), n * sizeof(BrJob), hipMemcpyHostToDevice, st));
    // the default rotation kernel of the loaded keys for this batch size
    br_lvl1(d, n, lvl1_pick_c(0, n), mu, st);
    CUFHE_HIP_TRY(hipFreeAsync(d, st));
    CUFHE_HIP_TRY(hipStreamSynchronize(st));
    return check_launch("bootstrap to TRLWE");
}

int cufhe_refresh_lvl1(uint32_t* d_out, const uint32_t* d_in, int n, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const hipStream_t st = as_stream(stream);
    uint32_t* tl0 = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&tl0),
                                 static_cast<size_t>(n) * (lvl0param::n + 1) * sizeof(uint32_t), st));
    launch_se_ks_lvl10(s.k1.subset, tl0, d_in, n, s.k1.ksk, st);
    int rc = check_launch("refresh key switch");
    if (rc == CUFHE_OK)
        rc = cufhe_bootstrap_to_trlwe_lvl01(d_out, tl0, static_cast<uint32_t>(lvl1param::μ), n, stream);
    (void)hipFreeAsync(tl0, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

int cufhe_plan_create_lvl02(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                            int use_graph, cufhe_plan** out)
{
    State& s = state();
    if (!out) return fail(CUFHE_ERR_ARG, "null plan handle");
    *out = nullptr;
    if (!s.k2.loaded) return fail(CUFHE_ERR_STATE, "lvl02 keys not loaded (cufhe_initialize_lvl02)");
    if (n < 0 || (n > 0 && (!gates || !d_pool))) return fail(CUFHE_ERR_ARG, "bad gate list");
    constexpr uint32_t W0 = lvl0param::n + 1;   // lvl0 wire stride
    constexpr uint32_t W2 = lvl2param::n + 1;   // lvl2 TLWE (uint64 words)
    constexpr uint32_t mu0 = static_cast<uint32_t>(lvl0param::μ);
    constexpr uint64_t mu2 = static_cast<uint64_t>(lvl2param::μ);
    int nbr = 0;
    for (int i = 0; i < n; i++) {
        const int op = gates[i].op;
        if (op < 0 || op >= CUFHE_GATE_COUNT) return fail(CUFHE_ERR_ARG, "unknown gate op");
        if (op == CUFHE_MUX || op == CUFHE_NMUX)
            nbr += 2;
        else if (op != CUFHE_NOT && op != CUFHE_COPY)
            nbr += 1;
    }
    cufhe_plan* p = new cufhe_plan();
    p->level = 2;
    p->c = 1;
    p->use_graph = use_graph != 0;
    if (nbr) {
        if (hipMalloc(reinterpret_cast<void**>(&p->ws),
                      static_cast<size_t>(nbr) * W2 * sizeof(uint64_t)) != hipSuccess) {
            delete p;
            return fail(CUFHE_ERR_NOMEM, "plan workspace allocation failed");
        }
    }
    uint64_t* tl2 = reinterpret_cast<uint64_t*>(p->ws);
    std::vector<KsJob> ks;
    std::vector<BrJob> br;
    std::vector<LinJob> lin;
    auto wire = [&](uint32_t w) { return d_pool + static_cast<size_t>(w) * W0; };
    for (int i = 0; i < n; i++) {
        const cufhe_gate_desc& gd = gates[i];
        if (gd.op == CUFHE_NOT || gd.op == CUFHE_COPY) {
            lin.push_back(LinJob{wire(gd.in0), nullptr, wire(gd.out), gd.op == CUFHE_NOT ? -1 : 1, 0,
                                 0u, static_cast<int32_t>(W0)});
            continue;
        }
        if (gd.op == CUFHE_MUX || gd.op == CUFHE_NMUX) {
            uint64_t* r1 = tl2 + static_cast<size_t>(br.size()) * W2;
            br.push_back(BrJob{wire(gd.inc), wire(gd.in1), r1, 1, 1, 0u - mu0, 0u});
            uint64_t* r0 = tl2 + static_cast<size_t>(br.size()) * W2;
            br.push_back(BrJob{wire(gd.inc), wire(gd.in0), r0, -1, 1, 0u - mu0, 0u});
            const int sg = gd.op == CUFHE_MUX ? 1 : -1;
            ks.push_back(KsJob{r1, r0, wire(gd.out), sg, sg, sg > 0 ? mu2 : 0 - mu2});
            continue;
        }
        GateLin gl;
        gate_lin(gd.op, gl);
        uint64_t* r = tl2 + static_cast<size_t>(br.size()) * W2;
        br.push_back(BrJob{wire(gd.in0), wire(gd.in1), r, gl.ca, gl.cb,
                           static_cast<uint32_t>(gl.off_mu) * mu0, 0u});
        ks.push_back(KsJob{r, nullptr, wire(gd.out), 1, 0, 0});
    }
    int rc = alloc_upload(&p->ks, ks);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->br, br);
    if (rc == CUFHE_OK) rc = alloc_upload(&p->lin, lin);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    p->njobs_ks = static_cast<int>(ks.size());
    p->njobs_br = static_cast<int>(br.size());
    p->njobs_lin = static_cast<int>(lin.size());
    rc = plan_alloc_ks_ws(p);
    if (rc != CUFHE_OK) {
        cufhe_plan_destroy(p);
        return rc;
    }
    *out = p;
    return CUFHE_OK;
}

int cufhe_plan_run(cufhe_plan* p, void* stream)
{
    if (!p) return fail(CUFHE_ERR_ARG, "null plan");
    const int krc = plan_check_keys(p);
    if (krc != CUFHE_OK) return krc;
    const hipStream_t st = as_stream(stream);
    if (!p->use_graph) return plan_launch(p, st);
    if (!p->exec) {
        // Capture on a private stream; the graph is replayed on `stream`.
        hipStream_t cs;
        CUFHE_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        CUFHE_HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = plan_launch(p, cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        if (rc != CUFHE_OK) return rc;
        if (e != hipSuccess) return fail(CUFHE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        p->graph = g;
        CUFHE_HIP_TRY(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    }
    CUFHE_HIP_TRY(hipGraphLaunch(p->exec, st));
    return CUFHE_OK;
}

int cufhe_plan_destroy(cufhe_plan* p)
{
    if (!p) return CUFHE_OK;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->ks);
    (void)hipFree(p->br);
    (void)hipFree(p->lin);
    (void)hipFree(p->ws);
    (void)hipFree(p->ks_ws);
    delete p;
    return CUFHE_OK;
}

int cufhe_plan_bootstraps(const cufhe_plan* p) { return p ? p->njobs_br : 0; }

int cufhe_plan_jobs_per_group(const cufhe_plan* p) { return p ? p->c : 0; }

int cufhe_plan_run_timed(cufhe_plan* p, void* stream, float* kernel_ms)
{
    if (!p || !kernel_ms) return fail(CUFHE_ERR_ARG, "null argument");
    const int krc = plan_check_keys(p);
    if (krc != CUFHE_OK) return krc;
    State& s = state();
    const hipStream_t st = as_stream(stream);
    // events live until the function returns, on every path
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events()
        {
            for (auto& x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    for (auto& x : evs.e) CUFHE_HIP_TRY(hipEventCreate(&x));
    hipEvent_t* e = evs.e;
    if (p->level == 2) {
        CUFHE_HIP_TRY(hipEventRecord(e[0], st));
        br_lvl2(p->br, p->njobs_br, static_cast<uint64_t>(lvl2param::μ), st);
        CUFHE_HIP_TRY(hipEventRecord(e[1], st));
        if (!ks_lvl2(p->ks, p->njobs_ks, p->ks_ws, st))
            return fail(CUFHE_ERR_HIP, "key switch failed to launch");
        CUFHE_HIP_TRY(hipEventRecord(e[2], st));
        if (p->njobs_lin) launch_linear(p->lin, p->njobs_lin, st);
        CUFHE_HIP_TRY(hipEventRecord(e[3], st));
        int rc2 = check_launch("timed lvl02 plan");
        if (rc2 == CUFHE_OK) {
            const hipError_t er = hipEventSynchronize(e[3]);
            if (er != hipSuccess) rc2 = fail(CUFHE_ERR_HIP, std::string("timed plan: ") + hipGetErrorString(er));
        }
        if (rc2 == CUFHE_OK) {
            // report in the lvl1 order: [0] key switch, [1] blind rotation, [2] combine
            (void)hipEventElapsedTime(&kernel_ms[1], e[0], e[1]);
            (void)hipEventElapsedTime(&kernel_ms[0], e[1], e[2]);
            (void)hipEventElapsedTime(&kernel_ms[2], e[2], e[3]);
        }
        return rc2;
    }
    hipEvent_t* ev = e;
    CUFHE_HIP_TRY(hipEventRecord(ev[0], st));
    if (p->level == 3) {
        // BR-IKS order; reported as [0] key switch, [1] rotation, [2] combine
        br_lvl1(p->br, p->njobs_br, p->c, static_cast<uint32_t>(lvl1param::μ), st);
        CUFHE_HIP_TRY(hipEventRecord(ev[1], st));
        if (!ks_lvl1(p->ks, p->njobs_ks, p->ks_ws, st))
            return fail(CUFHE_ERR_HIP, "key switch failed to launch");
    }
    else {
        if (!ks_lvl1(p->ks, p->njobs_ks, p->ks_ws, st))
            return fail(CUFHE_ERR_HIP, "key switch failed to launch");
        CUFHE_HIP_TRY(hipEventRecord(ev[1], st));
        br_lvl1(p->br, p->njobs_br, p->c, static_cast<uint32_t>(lvl1param::μ), st);
    }
    CUFHE_HIP_TRY(hipEventRecord(ev[2], st));
    if (p->njobs_lin) launch_linear(p->lin, p->njobs_lin, st);
    CUFHE_HIP_TRY(hipEventRecord(ev[3], st));
    int rc = check_launch("timed plan");
    if (rc == CUFHE_OK) {
        const hipError_t e = hipEventSynchronize(ev[3]);
        if (e != hipSuccess) rc = fail(CUFHE_ERR_HIP, std::string("timed plan: ") + hipGetErrorString(e));
    }
    if (rc == CUFHE_OK) {
        for (int i = 0; i < 3; i++) (void)hipEventElapsedTime(&kernel_ms[i], ev[i], ev[i + 1]);
        if (p->level == 3) std::swap(kernel_ms[0], kernel_ms[1]);
    }
    return rc;
}

int cufhe_gate_batch_lvl1(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                          int jobs_per_group, void* stream)
{
    cufhe_plan* p = nullptr;
    int rc = cufhe_plan_create_lvl1(gates, n, d_pool, jobs_per_group, 0, &p);
    if (rc != CUFHE_OK) return rc;
    rc = cufhe_plan_run(p, stream);
    if (rc == CUFHE_OK) {
        const hipError_t e = hipStreamSynchronize(as_stream(stream));
        if (e != hipSuccess) rc = fail(CUFHE_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    }
    cufhe_plan_destroy(p);
    return rc;
}

int cufhe_autotune_lvl1(int nbootstraps, int* best, float* times_ms, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (nbootstraps <= 0 || nbootstraps > (1 << 20)) return fail(CUFHE_ERR_ARG, "nbootstraps out of range");
    const hipStream_t st = as_stream(stream);
    const int n = nbootstraps;
    constexpr size_t W0 = lvl0param::n + 1, W1 = lvl1param::n + 1;
    // synthetic jobs: pseudo-random lvl0 inputs (the rotation's work does not
    // depend on the values), extracted outputs
    std::vector<uint32_t> h_in(static_cast<size_t>(n) * W0);
    uint64_t x = 0x9e3779b97f4a7c15ull;
    for (auto& v : h_in) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        v = static_cast<uint32_t>(x >> 16);
    }
    uint32_t* d_in = nullptr;
    uint32_t* d_out = nullptr;
    BrJob* d_jobs = nullptr;
    std::vector<BrJob> jobs(n);
    struct Mem {
        void* p[3] = {nullptr, nullptr, nullptr};
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Mem()
        {
            for (void* q : p) (void)hipFree(q);
            for (hipEvent_t v : e)
                if (v) (void)hipEventDestroy(v);
        }
    } mem;
    CUFHE_HIP_TRY(hipMalloc(&mem.p[0], h_in.size() * 4));
    CUFHE_HIP_TRY(hipMalloc(&mem.p[1], static_cast<size_t>(n) * W1 * 4));
    CUFHE_HIP_TRY(hipMalloc(&mem.p[2], static_cast<size_t>(n) * sizeof(BrJob)));
    d_in = static_cast<uint32_t*>(mem.p[0]);
    d_out = static_cast<uint32_t*>(mem.p[1]);
    d_jobs = static_cast<BrJob*>(mem.p[2]);
    for (int i = 0; i < n; i++)
        jobs[i] = BrJob{d_in + static_cast<size_t>(i) * W0, nullptr, d_out + static_cast<size_t>(i) * W1,
                        1, 0, 0, 0};
    CUFHE_HIP_TRY(hipMemcpyAsync(d_in, h_in.data(), h_in.size() * 4, hipMemcpyHostToDevice, st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(BrJob), hipMemcpyHostToDevice, st));
    CUFHE_HIP_TRY(hipEventCreate(&mem.e[0]));
    CUFHE_HIP_TRY(hipEventCreate(&mem.e[1]));
    if (times_ms)
        for (int c = 0; c < 9; c++) times_ms[c] = 0.f;
    int bestc = 0;
    float bestt = 0.f;
    for (int c : {1, 2, 4, 6, 8}) {
        if (lvl1_pick_c(c, n) != c) continue;  // not valid for these keys
        if (c == 8 && !s.k1.kb) continue;
        float t = 0.f;
        for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
            CUFHE_HIP_TRY(hipEventRecord(mem.e[0], st));
            br_lvl1(d_jobs, n, c, static_cast<uint32_t>(lvl1param::μ), st);
            CUFHE_HIP_TRY(hipEventRecord(mem.e[1], st));
            const int rc = check_launch("autotune rotation");
            if (rc != CUFHE_OK) return rc;
            CUFHE_HIP_TRY(hipEventSynchronize(mem.e[1]));
            CUFHE_HIP_TRY(hipEventElapsedTime(&t, mem.e[0], mem.e[1]));
        }
        if (times_ms) times_ms[c] = t;
        if (bestc == 0 || t < bestt) {
            bestc = c;
            bestt = t;
        }
    }
    if (bestc == 0) return fail(CUFHE_ERR_STATE, "no rotation kernel for these keys");
    {
        TuneTable& t = tune_table();
        std::lock_guard<std::mutex> g(t.mu);
        if (t.generation != s.k1.generation) {
            t.best.clear();
            t.generation = s.k1.generation;
        }
        t.best[n] = bestc;
    }
    if (best) *best = bestc;
    return CUFHE_OK;
}

int cufhe_set_lvl2_jobs_per_group(int c)
{
    if (c != 1 && c != 2) return fail(CUFHE_ERR_ARG, "lvl2 jobs per workgroup must be 1 or 2");
    g_lvl2_c.store(c, std::memory_order_relaxed);
    return CUFHE_OK;
}

int cufhe_set_key_switch_split(int split)
{
    if (split < 0 || split > 64) return fail(CUFHE_ERR_ARG, "key switch split out of range 0-64");
    ks_gemm_set_split(split);
    return CUFHE_OK;
}

int cufhe_identity_key_switch_lvl10(uint32_t* d_out, const uint32_t* d_in, int n, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const hipStream_t st = as_stream(stream);
    std::vector<KsJob> jobs(n);
    for (int i = 0; i < n; i++)
        jobs[i] = KsJob{d_in + static_cast<size_t>(i) * (lvl1param::n + 1), nullptr,
                        d_out + static_cast<size_t>(i) * (lvl0param::n + 1), 1, 0, 0};
    KsJob* d = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(KsJob), st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d, jobs.data(), n * sizeof(KsJob), hipMemcpyHostToDevice, st));
    (void)ks_lvl1(d, n, nullptr, st);
    CUFHE_HIP_TRY(hipFreeAsync(d, st));
    CUFHE_HIP_TRY(hipStreamSynchronize(st));
    return check_launch("identity key switch");
}

int cufhe_identity_key_switch_lvl10_noalloc(uint32_t* d_out, const uint32_t* d_in, int n,
                                            void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n < 0) return fail(CUFHE_ERR_ARG, "negative count");
    if (n > kNoAllocMaxJobs)
        return fail(CUFHE_ERR_ARG, "identity key switch without allocation takes at most " +
                                       std::to_string(kNoAllocMaxJobs) + " TLWEs per call");
    if (n == 0) return CUFHE_OK;
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    if (!s.k1.ksg.planes) return fail(CUFHE_ERR_STATE, "key switch fragments not prepared");
    void* base = nullptr;
    const int src = scratch_on_current_device(&base);
    if (src != CUFHE_OK) return src;
    KsJob* jobs = static_cast<KsJob*>(base);
    void* ws = static_cast<unsigned char*>(base) + kNoAllocJobBytes;
    if (ks_gemm_workspace_bytes(s.k1.ksg, n) > kNoAllocScratchBytes - kNoAllocJobBytes)
        return fail(CUFHE_ERR_STATE, "no-alloc scratch smaller than the key switch workspace");
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_fill_ks_jobs, dim3((n + 255) / 256), dim3(256), 0, st, jobs, d_in, d_out, n);
    if (ks_gemm_run(s.k1.ksg, jobs, n, ws, st) != 0) return fail(CUFHE_ERR_HIP, "key switch launch failed");
    return check_launch("identity key switch (no allocation)");
}

int cufhe_identity_key_switch_lvl20(uint32_t* d_out, const uint64_t* d_in, int n, void* stream)
{
    State& s = state();
    if (!s.k2.loaded) return fail(CUFHE_ERR_STATE, "lvl02 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const hipStream_t st = as_stream(stream);
    std::vector<KsJob> jobs(n);
    for (int i = 0; i < n; i++)
        jobs[i] = KsJob{d_in + static_cast<size_t>(i) * (lvl2param::n + 1), nullptr,
                        d_out + static_cast<size_t>(i) * (lvl0param::n + 1), 1, 0, 0};
    KsJob* d = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(KsJob), st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d, jobs.data(), n * sizeof(KsJob), hipMemcpyHostToDevice, st));
    (void)ks_lvl2(d, n, nullptr, st);
    CUFHE_HIP_TRY(hipFreeAsync(d, st));
    CUFHE_HIP_TRY(hipStreamSynchronize(st));
    return check_launch("identity key switch lvl20");
}

int cufhe_bootstrap_lvl01(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, int n,
                          int jobs_per_group, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const int c = pick_c(jobs_per_group, s.k1.kb, n);
    if (c < 0) return fail(CUFHE_ERR_UNSUPPORTED, "unsupported jobs_per_group");
    const hipStream_t st = as_stream(stream);
    std::vector<BrJob> jobs(n);
    for (int i = 0; i < n; i++)
        jobs[i] = BrJob{d_in + static_cast<size_t>(i) * (lvl0param::n + 1), nullptr,
                        d_out + static_cast<size_t>(i) * (lvl1param::n + 1), 1, 0, 0u, 0u};
    BrJob* d = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(BrJob), st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d, jobs.data(), n * sizeof(BrJob), hipMemcpyHostToDevice, st));
    br_lvl1(d, n, c, mu, st);
    CUFHE_HIP_TRY(hipFreeAsync(d, st));
    CUFHE_HIP_TRY(hipStreamSynchronize(st));
    return check_launch("bootstrap");
}

int cufhe_bootstrap_lvl02(uint64_t* d_out, const uint32_t* d_in, uint64_t mu, int n, void* stream)
{
    State& s = state();
    if (!s.k2.loaded) return fail(CUFHE_ERR_STATE, "lvl02 keys not loaded");
    if (n <= 0) return n == 0 ? CUFHE_OK : fail(CUFHE_ERR_ARG, "negative count");
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const hipStream_t st = as_stream(stream);
    std::vector<BrJob> jobs(n);
    for (int i = 0; i < n; i++)
        jobs[i] = BrJob{d_in + static_cast<size_t>(i) * (lvl0param::n + 1), nullptr,
                        d_out + static_cast<size_t>(i) * (lvl2param::n + 1), 1, 0, 0u, 0u};
    BrJob* d = nullptr;
    CUFHE_HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), n * sizeof(BrJob), st));
    CUFHE_HIP_TRY(hipMemcpyAsync(d, jobs.data(), n * sizeof(BrJob), hipMemcpyHostToDevice, st));
    br_lvl2(d, n, mu, st);
    CUFHE_HIP_TRY(hipFreeAsync(d, st));
    CUFHE_HIP_TRY(hipStreamSynchronize(st));
    return check_launch("bootstrap lvl02");
}

int cufhe_ntt_transform(const uint32_t* d_in, uint32_t* d_out, int count, int inverse,
                        void* stream)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
    }
    if (count < 0 || !d_in || !d_out) return fail(CUFHE_ERR_ARG, "bad arguments");
    if (count == 0) return CUFHE_OK;
    launch_ntt_natural(inverse != 0, d_out, d_in, count, s.tn.tb, as_stream(stream));
    return check_launch("ntt transform");
}

uint32_t cufhe_ntt_root(void) { return ntt31::PSI; }

int cufhe_ntt64_transform(const uint64_t* d_in, uint64_t* d_out, int count, int inverse,
                          void* stream)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
    }
    if (count < 0 || !d_in || !d_out) return fail(CUFHE_ERR_ARG, "bad arguments");
    if (count == 0) return CUFHE_OK;
    launch_ntt64_natural(inverse != 0, d_out, d_in, count, s.t64.tb, as_stream(stream));
    return check_launch("ntt64 transform");
}

uint64_t cufhe_ntt64_root(void) { return gold_psi(); }

int cufhe_ffnt_forward(int N, const void* d_poly, int integer, void* d_spectra, int count,
                       void* stream)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
    }
    if (count < 0 || !d_poly || !d_spectra) return fail(CUFHE_ERR_ARG, "bad arguments");
    if (count == 0) return CUFHE_OK;
    if (N == 1024)
        launch_to_fourier<lvl1param>(static_cast<cplx*>(d_spectra), d_poly, count, integer != 0,
                                     keys_view(s.t1, nullptr), as_stream(stream));
    else if (N == 2048)
        launch_to_fourier<lvl2param>(static_cast<cplx*>(d_spectra), d_poly, count, integer != 0,
                                     keys_view(s.t2, nullptr), as_stream(stream));
    else
        return fail(CUFHE_ERR_UNSUPPORTED, "N must be 1024 or 2048");
    return check_launch("ffnt forward");
}

int cufhe_ffnt_inverse(int N, const void* d_spectra, double* d_poly, int count, void* stream)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
    }
    if (count < 0 || !d_poly || !d_spectra) return fail(CUFHE_ERR_ARG, "bad arguments");
    if (count == 0) return CUFHE_OK;
    if (N == 1024)
        launch_from_fourier<lvl1param>(d_poly, static_cast<const cplx*>(d_spectra), count,
                                       keys_view(s.t1, nullptr), as_stream(stream));
    else if (N == 2048)
        launch_from_fourier<lvl2param>(d_poly, static_cast<const cplx*>(d_spectra), count,
                                       keys_view(s.t2, nullptr), as_stream(stream));
    else
        return fail(CUFHE_ERR_UNSUPPORTED, "N must be 1024 or 2048");
    return check_launch("ffnt inverse");
}

size_t cufhe_trgsw_transform_bytes_lvl1(void)
{
    const LevelKeys<lvl1param>& k = state().k1;
    if (!k.loaded) return 0;
    return k.backend == CUFHE_BACKEND_NTT ? kTrgswNtt * sizeof(uint32_t)
                                          : kTrgswFourier<lvl1param> * sizeof(cplx);
}

int cufhe_trgsw_to_transform_lvl1(const uint32_t* d_trgsw, void* d_out, int count, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (they select the backend)");
    if (count < 0) return fail(CUFHE_ERR_ARG, "negative count");
    if (count == 0) return CUFHE_OK;
    if (!d_trgsw || !d_out) return fail(CUFHE_ERR_ARG, "null pointer");
    const int npolys = count * TFHEpp::BootstrappingTRGSWRows<lvl1param> * 2;
    const hipStream_t st = as_stream(stream);
    if (s.k1.backend == CUFHE_BACKEND_NTT)
        launch_to_ntt(static_cast<uint32_t*>(d_out), d_trgsw, npolys, s.tn.tb, st);
    else
        launch_to_fourier<lvl1param>(static_cast<cplx*>(d_out), d_trgsw, npolys, false, s.k1.fk, st);
    return check_launch("TRGSW transform");
}

static int cmux_launch(const CmuxBatch& b, int n, void* stream)
{
    State& s = state();
    if (!s.k1.loaded) return fail(CUFHE_ERR_STATE, "lvl1 keys not loaded (they select the backend)");
    if (n < 0) return fail(CUFHE_ERR_ARG, "negative count");
    if (n == 0) return CUFHE_OK;
    const hipStream_t st = as_stream(stream);
    if (s.k1.backend == CUFHE_BACKEND_NTT)
        launch_cmux_batch_ntt(b, n, s.k1.nk, st);
    else
        launch_cmux_batch_fft(b, n, s.k1.fk, st);
    return check_launch("CMUX");
}

int cufhe_cmux_lvl1(uint32_t* d_res, const void* d_cs, const uint32_t* d_c1, const uint32_t* d_c0,
                    int n, void* stream)
{
    if (n > 0 && (!d_res || !d_cs || !d_c1 || !d_c0)) return fail(CUFHE_ERR_ARG, "null pointer");
    CmuxBatch b{nullptr, d_res, d_cs, d_c1, d_c0, cufhe_trgsw_transform_bytes_lvl1()};
    return cmux_launch(b, n, stream);
}

int cufhe_cmux_batch_lvl1(const cufhe_cmux_desc* d_descs, int n, void* stream)
{
    static_assert(sizeof(cufhe_cmux_desc) == sizeof(CmuxDesc), "descriptor layout");
    if (n > 0 && !d_descs) return fail(CUFHE_ERR_ARG, "null pointer");
    CmuxBatch b{reinterpret_cast<const CmuxDesc*>(d_descs), nullptr, nullptr, nullptr, nullptr, 0};
    return cmux_launch(b, n, stream);
}

int cufhe_trgsw_to_fourier(int N, const void* d_trgsw, void* d_spectra, int count, void* stream)
{
    if (N == 1024)
        return cufhe_ffnt_forward(N, d_trgsw, 0, d_spectra,
                                  count * TFHEpp::BootstrappingTRGSWRows<lvl1param> * 2, stream);
    if (N == 2048)
        return cufhe_ffnt_forward(N, d_trgsw, 0, d_spectra,
                                  count * TFHEpp::BootstrappingTRGSWRows<lvl2param> * 2, stream);
    return fail(CUFHE_ERR_UNSUPPORTED, "N must be 1024 or 2048");
}

int cufhe_annihilate_key_load_ex(int level, const void* ahk, size_t words, int backend)
{
    if (!ahk) return fail(CUFHE_ERR_ARG, "null annihilation key");
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "annihilation level must be 1 or 2");
    if (backend != CUFHE_BACKEND_FFT && backend != CUFHE_BACKEND_NTT)
        return fail(CUFHE_ERR_ARG, "backend must be CUFHE_BACKEND_FFT or CUFHE_BACKEND_NTT");
    if (level == 1 && backend == CUFHE_BACKEND_NTT)
        return fail(CUFHE_ERR_UNSUPPORTED, "level-1 annihilation runs on the FFT backend only");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    const int rc = ensure_init();
    if (rc != CUFHE_OK) return rc;
    // key switches still running on any stream may read the key memory that
    // is rewritten or freed below
    CUFHE_HIP_TRY(hipDeviceSynchronize());
    if (level == 1) return load_annihilate_key(s.ahk1, s.t1, ahk, words);
    if (backend == CUFHE_BACKEND_FFT) {
        const int r2 = load_annihilate_key(s.ahk2, s.t2, ahk, words);
        if (r2 == CUFHE_OK && s.ahk2_ntt) {
            (void)hipFree(s.ahk2_ntt);
            s.ahk2_ntt = nullptr;
        }
        return r2;
    }
    const int r2 = load_annihilate_key_ntt64(s.ahk2_ntt, s.t64, ahk, words);
    if (r2 == CUFHE_OK && s.ahk2) {
        (void)hipFree(s.ahk2);
        s.ahk2 = nullptr;
    }
    return r2;
}

int cufhe_annihilate_key_load(int level, const void* ahk, size_t words)
{
    return cufhe_annihilate_key_load_ex(level, ahk, words, CUFHE_BACKEND_FFT);
}

int cufhe_annihilate_key_delete(int level)
{
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "annihilation level must be 1 or 2");
    State& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    (void)hipDeviceSynchronize();
    if (level == 1) {
        (void)hipFree(s.ahk1);
        s.ahk1 = nullptr;
    }
    else {
        (void)hipFree(s.ahk2);
        (void)hipFree(s.ahk2_ntt);
        s.ahk2 = nullptr;
        s.ahk2_ntt = nullptr;
    }
    return CUFHE_OK;
}

int cufhe_annihilate_key_switch_batch(int level, void* d_out, size_t out_stride, const void* d_in,
                                      size_t in_stride, size_t batch, void* stream)
{
    if (level != 1 && level != 2) return fail(CUFHE_ERR_ARG, "annihilation level must be 1 or 2");
    if (batch == 0) return CUFHE_OK;
    if (!d_out || !d_in) return fail(CUFHE_ERR_ARG, "null pointer");
    const size_t words = level == 1 ? 2 * lvl1param::n : 2 * lvl2param::n;
    if (out_stride < words || in_stride < words)
        return fail(CUFHE_ERR_ARG, "stride below the TRLWE size");
    if (batch > static_cast<size_t>(std::numeric_limits<int>::max()))
        return fail(CUFHE_ERR_ARG, "batch too large");
    State& s = state();
    const cplx* key;
    const uint64_t* key_ntt;
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
        key = level == 1 ? s.ahk1 : s.ahk2;
        key_ntt = level == 2 ? s.ahk2_ntt : nullptr;
    }
    if (!key && !key_ntt) return fail(CUFHE_ERR_STATE, "no annihilation key loaded for this level");
    if (key_ntt)
        launch_annihilate_ntt64(static_cast<uint64_t*>(d_out), out_stride,
                                static_cast<const uint64_t*>(d_in), in_stride,
                                static_cast<int>(batch), key_ntt, s.t64.tb, as_stream(stream));
    else if (level == 1)
        launch_annihilate_fft<lvl1param>(static_cast<uint32_t*>(d_out), out_stride,
                                         static_cast<const uint32_t*>(d_in), in_stride,
                                         static_cast<int>(batch), key, keys_view(s.t1, nullptr),
                                         as_stream(stream));
    else
        launch_annihilate_fft<lvl2param>(static_cast<uint64_t*>(d_out), out_stride,
                                         static_cast<const uint64_t*>(d_in), in_stride,
                                         static_cast<int>(batch), key, keys_view(s.t2, nullptr),
                                         as_stream(stream));
    return check_launch("annihilate");
}

int cufhe_annihilate_key_switch(int level, void* d_out, const void* d_in, void* stream)
{
    const size_t words = level == 2 ? 2 * lvl2param::n : 2 * lvl1param::n;
    return cufhe_annihilate_key_switch_batch(level, d_out, words, d_in, words, 1, stream);
}

int cufhe_cmux_accumulate(int N, void* d_trlwe, const void* d_trgsw_spectra, uint32_t abar,
                          void* stream)
{
    State& s = state();
    {
        std::lock_guard<std::mutex> g(s.mu);
        const int rc = ensure_init();
        if (rc != CUFHE_OK) return rc;
    }
    if (!d_trlwe || !d_trgsw_spectra) return fail(CUFHE_ERR_ARG, "null pointer");
    if (abar >= 2u * static_cast<uint32_t>(N)) return fail(CUFHE_ERR_ARG, "abar out of [0, 2N)");
    if (N == 1024)
        launch_cmux_fft<lvl1param>(static_cast<uint32_t*>(d_trlwe),
                                   static_cast<const cplx*>(d_trgsw_spectra), abar,
                                   keys_view(s.t1, nullptr), as_stream(stream));
    else if (N == 2048)
        launch_cmux_fft<lvl2param>(static_cast<uint64_t*>(d_trlwe),
                                   static_cast<const cplx*>(d_trgsw_spectra), abar,
                                   keys_view(s.t2, nullptr), as_stream(stream));
    else
        return fail(CUFHE_ERR_UNSUPPORTED, "N must be 1024 or 2048");
    return check_launch("cmux");
}

}  // extern "C"
