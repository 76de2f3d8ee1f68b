// Annihilation key switching through the C++ drop-in (include/annihilate_gpu.hpp),
// the cases of the reference's test (reference test/test_annihilate_gpu.cu:46-133)
// at lvl1param and lvl2param: a trivial body (mu at X^0), a mask of (i + 1)
// shifted to the upper half of the word, and an encryption of mu at X^0. The
// reference compares with TFHEpp's CPU annihilation; here the expectation is
// the identity the operation computes: the output phase keeps the input
// phase's constant coefficient and the others vanish, within the key's noise
// and decomposition error (2^26 of 2^32 at lvl1, 2^56 of 2^64 at lvl2: the
// reference's own tolerances). The batched form with a workspace must give
// the same words as the single form.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <type_traits>

#include "annihilate_gpu.hpp"
#include "cufhe_gpu.hpp"

using namespace TFHEpp;

namespace {

template <class P>
bool run_case(const char* name, int backend = CUFHE_BACKEND_FFT)
{
    using T = typename P::T;
    constexpr uint32_t N = P::n;
    std::printf("=== %s annihilate GPU test ===\n", name);
    auto sk = std::make_unique<SecretKey>(11);
    auto ahk = std::make_unique<cufhe::AnnihilateKeyPolynomial<P>>();
    cufhe::AnnihilateKeyPolynomialGen<P>(*ahk, *sk);
    cufhe::AnnihilateKeyPolynomialToDevice<P>(*ahk, 1, backend);

    constexpr size_t bytes = sizeof(TRLWE<P>);
    T *d_in = nullptr, *d_out = nullptr, *d_ws = nullptr;
    if (hipMalloc(&d_in, 3 * bytes) != hipSuccess || hipMalloc(&d_out, 3 * bytes) != hipSuccess ||
        hipMalloc(&d_ws, 3 * bytes) != hipSuccess)
        return false;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return false;

    const T mu = static_cast<T>(1) << (std::numeric_limits<T>::digits - 3);
    constexpr uint64_t tol = sizeof(T) == 4 ? (1ULL << 26) : (1ULL << 56);
    auto in = std::make_unique<std::array<TRLWE<P>, 3>>();
    auto out = std::make_unique<std::array<TRLWE<P>, 3>>();
    (*in)[0] = TRLWE<P>{};
    (*in)[0][1][0] = mu;
    (*in)[1] = TRLWE<P>{};
    for (uint32_t i = 0; i < N; i++)
        (*in)[1][0][i] = static_cast<T>(i + 1) << (std::numeric_limits<T>::digits / 2);
    Polynomial<P> m{};
    m[0] = mu;
    trlweSymEncrypt<P>((*in)[2], m, sk->key.get<P>());
    const char* kinds[3] = {"trivial body", "first mask", "encrypted"};

    bool ok = true;
    if (hipMemcpy(d_in, in->data(), 3 * bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
    for (int c = 0; c < 3; c++)
        cufhe::AnnihilateKeySwitching<P>(d_out + c * 2 * N, d_in + c * 2 * N, st, 0);
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(out->data(), d_out, 3 * bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    for (int c = 0; c < 3; c++) {
        const auto ph_in = trlwePhase<P>((*in)[c], sk->key.get<P>());
        const auto ph = trlwePhase<P>((*out)[c], sk->key.get<P>());
        uint64_t worst = 0;
        for (uint32_t i = 0; i < N; i++) {
            const T want = i == 0 ? ph_in[0] : 0;
            using S = std::make_signed_t<T>;
            const S d = static_cast<S>(ph[i] - want);
            const uint64_t a = d < 0 ? static_cast<uint64_t>(-(d + 1)) + 1 : static_cast<uint64_t>(d);
            if (a > worst) worst = a;
        }
        const bool pass = worst <= tol;
        std::printf("case %d (%s): phase max abs diff=2^%.1f %s\n", c, kinds[c],
                    worst ? std::log2(static_cast<double>(worst)) : 0.0, pass ? "pass" : "mismatch");
        ok = ok && pass;
    }
    // batched form with a (unused) workspace argument
    cufhe::AnnihilateKeySwitchingBatchWithWorkspace<P>(d_ws, 2 * N, d_in, 2 * N, nullptr, 0, 3, st, 0);
    auto batched = std::make_unique<std::array<TRLWE<P>, 3>>();
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(batched->data(), d_ws, 3 * bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    const bool same = std::memcmp(batched->data(), out->data(), 3 * bytes) == 0;
    std::printf("batched == single: %s\n", same ? "pass" : "mismatch");
    ok = ok && same;

    (void)hipStreamDestroy(st);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_ws);
    cufhe::DeleteAnnihilateKey<P>(1);
    return ok;
}

}  // namespace

int main()
{
    cufhe::SetGPUNum(1);
    cufhe::Initialize();
    bool ok = run_case<lvl1param>("lvl1param");
    ok = run_case<lvl2param>("lvl2param") && ok;
    ok = run_case<lvl2param>("lvl2param, NTT backend", CUFHE_BACKEND_NTT) && ok;
    cufhe::CleanUp();
    std::printf(ok ? "ALL ANNIHILATE TESTS PASSED\n" : "ANNIHILATE TESTS FAILED\n");
    return ok ? 0 : 1;
}
