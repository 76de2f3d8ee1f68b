// The lower launcher layer of the C++ drop-in, called the way the
// reference's gate API calls it (reference src/cufhe_gates_gpu.cu:198-537,
// include/bootstrap_gpu.cuh:40-217, include/keyswitch_gpu.cuh:258-275):
//   XxxBootstrap<lvl01param, mu, lvl10param>  (BR-IKS, lvl0 in/out)
//   XxxBootstrap<lvl10param, lvl01param, mu>  (IKS-BR, lvl1 in/out)
//   XxxBootstrap<lvl02param, mu, lvl20param>  (BR-IKS through lvl2)
//   Mux/NMux/Copy/Not launchers, Bootstrap(out, in, mu), SEIandKS,
//   BootstrapTLWE2TRLWE, IdentityKeySwitch<lvl10param / lvl20param>,
//   TRGSW2NTT and CMUXNTT.
// Raw device pointers and hipStream_t, as the reference's launchers take.
// Prints PASS/FAIL per group and ALL PASS, returns 0/1.
#include <cstdio>
#include <random>
#include <vector>

#include "cufhe_gpu.hpp"

using namespace TFHEpp;

namespace {

template <class P>
typename P::T* dev_alloc(size_t count)
{
    void* p = nullptr;
    cufhe::CuCheck(cufhe_device_alloc(&p, count * sizeof(typename P::T)), "alloc");
    return static_cast<typename P::T*>(p);
}

// Copies go on the stream the launchers use (it does not synchronise with
// the null stream), and the device is synchronised after each, so an input
// is on the device before the next launch and a result is read after the
// launches that write it.
hipStream_t g_st = nullptr;

template <class P>
typename P::T* upload(const TLWE<P>& c)
{
    auto* d = dev_alloc<P>(c.size());
    cufhe::CuCheck(cufhe_memcpy_h2d(d, c.data(), sizeof(c), g_st), "h2d");
    cufhe::Synchronize();
    return d;
}

template <class P>
TLWE<P> download(const typename P::T* d)
{
    TLWE<P> c;
    cufhe::Synchronize();
    cufhe::CuCheck(cufhe_memcpy_d2h(c.data(), d, sizeof(c), g_st), "d2h");
    cufhe::Synchronize();
    return c;
}

int report(const char* what, int errors, int total)
{
    std::printf("%-44s %3d / %3d  %s\n", what, total - errors, total, errors ? "FAIL" : "PASS");
    return errors;
}

using Fn2 = void (*)(uint32_t*, const uint32_t*, const uint32_t*, hipStream_t, int);

}  // namespace

int main()
{
    SecretKey sk(11);
    EvalKey ek;
    ek.emplacebk<lvl01param>(sk);
    ek.emplaceiksk<lvl10param>(sk);
    ek.emplacebk<lvl02param>(sk);
    ek.emplaceiksk<lvl20param>(sk);
    cufhe::SetGPUNum(1);
    cufhe::Initialize(ek);
    cufhe::Initialize_lvl02(ek, sk);
    cufhe::Stream s;
    s.Create();
    hipStream_t st = s.st();
    g_st = st;
    std::mt19937 rng(77);
    int bad = 0;
    constexpr uint32_t mu1 = lvl1param::μ;
    constexpr uint64_t mu2 = lvl2param::μ;

    struct G {
        const char* name;
        Fn2 br01, iks, br02;
        int (*plain)(int, int);
    };
    const G gates[] = {
        {"Nand", cufhe::NandBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::NandBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::NandBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return 1 - a * b; }},
        {"Or", cufhe::OrBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::OrBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::OrBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return a | b; }},
        {"OrYN", cufhe::OrYNBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::OrYNBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::OrYNBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return a | (1 - b); }},
        {"OrNY", cufhe::OrNYBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::OrNYBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::OrNYBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return (1 - a) | b; }},
        {"And", cufhe::AndBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::AndBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::AndBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return a & b; }},
        {"AndYN", cufhe::AndYNBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::AndYNBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::AndYNBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return a & (1 - b); }},
        {"AndNY", cufhe::AndNYBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::AndNYBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::AndNYBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return (1 - a) & b; }},
        {"Nor", cufhe::NorBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::NorBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::NorBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return 1 - (a | b); }},
        {"Xor", cufhe::XorBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::XorBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::XorBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return a ^ b; }},
        {"Xnor", cufhe::XnorBootstrap<lvl01param, mu1, lvl10param>,
         cufhe::XnorBootstrap<lvl10param, lvl01param, mu1>,
         cufhe::XnorBootstrap<lvl02param, mu2, lvl20param>, [](int a, int b) { return 1 - (a ^ b); }},
    };

    // binary gates: all four input pairs through each of the three forms
    {
        int e1 = 0, e0 = 0, e2 = 0, total = 0;
        for (const G& g : gates)
            for (int x = 0; x < 2; x++)
                for (int y = 0; y < 2; y++) {
                    const int want = g.plain(x, y);
                    auto c0x = tlweSymEncrypt<lvl0param>(x ? lvl0param::μ : -lvl0param::μ,
                                                         lvl0param::α, sk.key.lvl0);
                    auto c0y = tlweSymEncrypt<lvl0param>(y ? lvl0param::μ : -lvl0param::μ,
                                                         lvl0param::α, sk.key.lvl0);
                    auto c1x = tlweSymEncrypt<lvl1param>(x ? lvl1param::μ : -lvl1param::μ,
                                                         lvl1param::α, sk.key.lvl1);
                    auto c1y = tlweSymEncrypt<lvl1param>(y ? lvl1param::μ : -lvl1param::μ,
                                                         lvl1param::α, sk.key.lvl1);
                    uint32_t *dx0 = upload<lvl0param>(c0x), *dy0 = upload<lvl0param>(c0y);
                    uint32_t *dx1 = upload<lvl1param>(c1x), *dy1 = upload<lvl1param>(c1y);
                    uint32_t* o0 = dev_alloc<lvl0param>(lvl0param::n + 1);
                    uint32_t* o1 = dev_alloc<lvl1param>(lvl1param::n + 1);
                    uint32_t* o2 = dev_alloc<lvl0param>(lvl0param::n + 1);
                    g.br01(o0, dx0, dy0, st, 0);
                    g.iks(o1, dx1, dy1, st, 0);
                    g.br02(o2, dx0, dy0, st, 0);
                    e0 += tlweSymDecrypt<lvl0param>(download<lvl0param>(o0), sk.key.lvl0) != want;
                    e1 += tlweSymDecrypt<lvl1param>(download<lvl1param>(o1), sk.key.lvl1) != want;
                    e2 += tlweSymDecrypt<lvl0param>(download<lvl0param>(o2), sk.key.lvl0) != want;
                    total++;
                    for (uint32_t* p : {dx0, dy0, dx1, dy1, o0, o1, o2}) cufhe_device_free(p);
                }
        bad += report("XxxBootstrap<lvl01param, mu, lvl10param>", e0, total);
        bad += report("XxxBootstrap<lvl10param, lvl01param, mu>", e1, total);
        bad += report("XxxBootstrap<lvl02param, mu, lvl20param>", e2, total);
    }

    // the same lvl02 launchers with the lvl2 rotation on the NTT backend
    // (EvalKey::ntt, the reference's USE_FFT=OFF build), then back to FFT
    {
        ek.ntt = true;
        cufhe::Initialize_lvl02(ek, sk);
        int e = 0, total = 0;
        for (int gi : {0, 8})  // Nand, Xor
            for (int x = 0; x < 2; x++)
                for (int y = 0; y < 2; y++) {
                    const G& g = gates[gi];
                    auto c0x = tlweSymEncrypt<lvl0param>(x ? lvl0param::μ : -lvl0param::μ,
                                                         lvl0param::α, sk.key.lvl0);
                    auto c0y = tlweSymEncrypt<lvl0param>(y ? lvl0param::μ : -lvl0param::μ,
                                                         lvl0param::α, sk.key.lvl0);
                    uint32_t *dx0 = upload<lvl0param>(c0x), *dy0 = upload<lvl0param>(c0y);
                    uint32_t* o2 = dev_alloc<lvl0param>(lvl0param::n + 1);
                    g.br02(o2, dx0, dy0, st, 0);
                    e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o2), sk.key.lvl0) != g.plain(x, y);
                    total++;
                    for (uint32_t* p : {dx0, dy0, o2}) cufhe_device_free(p);
                }
        bad += report("XxxBootstrap<lvl02param, mu, lvl20param> on the NTT backend", e, total);
        ek.ntt = false;
        cufhe::Initialize_lvl02(ek, sk);
    }

    // Mux / NMux / Copy / Not launchers
    {
        int e = 0, total = 0;
        for (int i = 0; i < 8; i++) {
            const int c = i & 1, t = (i >> 1) & 1, f = (i >> 2) & 1;
            auto enc0 = [&](int b) {
                return upload<lvl0param>(tlweSymEncrypt<lvl0param>(
                    b ? lvl0param::μ : -lvl0param::μ, lvl0param::α, sk.key.lvl0));
            };
            auto enc1 = [&](int b) {
                return upload<lvl1param>(tlweSymEncrypt<lvl1param>(
                    b ? lvl1param::μ : -lvl1param::μ, lvl1param::α, sk.key.lvl1));
            };
            uint32_t *c0 = enc0(c), *t0 = enc0(t), *f0 = enc0(f);
            uint32_t *c1 = enc1(c), *t1 = enc1(t), *f1 = enc1(f);
            uint32_t* m0 = dev_alloc<lvl0param>(lvl0param::n + 1);
            uint32_t* n0 = dev_alloc<lvl0param>(lvl0param::n + 1);
            uint32_t* m1 = dev_alloc<lvl1param>(lvl1param::n + 1);
            uint32_t* n1 = dev_alloc<lvl1param>(lvl1param::n + 1);
            uint32_t* m2 = dev_alloc<lvl0param>(lvl0param::n + 1);
            uint32_t* cp = dev_alloc<lvl1param>(lvl1param::n + 1);
            uint32_t* nt = dev_alloc<lvl0param>(lvl0param::n + 1);
            cufhe::MuxBootstrap<lvl01param, mu1, lvl10param>(m0, c0, t0, f0, st, 0);
            cufhe::NMuxBootstrap<lvl01param, mu1, lvl10param>(n0, c0, t0, f0, st, 0);
            cufhe::MuxBootstrap<lvl10param, lvl01param, mu1>(m1, c1, t1, f1, st, 0);
            cufhe::NMuxBootstrap<lvl10param, lvl01param, mu1>(n1, c1, t1, f1, st, 0);
            cufhe::MuxBootstrap<lvl02param, mu2, lvl20param>(m2, c0, t0, f0, st, 0);
            cufhe::CopyBootstrap<lvl1param>(cp, t1, st, 0);
            cufhe::NotBootstrap<lvl0param>(nt, c0, st, 0);
            const int mux = c ? t : f;
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(m0), sk.key.lvl0) != mux;
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(n0), sk.key.lvl0) != !mux;
            e += tlweSymDecrypt<lvl1param>(download<lvl1param>(m1), sk.key.lvl1) != mux;
            e += tlweSymDecrypt<lvl1param>(download<lvl1param>(n1), sk.key.lvl1) != !mux;
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(m2), sk.key.lvl0) != mux;
            e += tlweSymDecrypt<lvl1param>(download<lvl1param>(cp), sk.key.lvl1) != t;
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(nt), sk.key.lvl0) != !c;
            total += 7;
            for (uint32_t* p : {c0, t0, f0, c1, t1, f1, m0, n0, m1, n1, m2, cp, nt}) cufhe_device_free(p);
        }
        bad += report("Mux/NMux/Copy/Not launchers", e, total);
    }

    // Bootstrap(out, in, mu): sign-preserving refresh of a lvl0 TLWE (also in
    // place), SEIandKS of BootstrapTLWE2TRLWE, IdentityKeySwitch lvl10 / lvl20
    {
        int e = 0, total = 0;
        for (int i = 0; i < 8; i++) {
            const int b = rng() & 1;
            auto c0 = tlweSymEncrypt<lvl0param>(b ? lvl0param::μ : -lvl0param::μ, lvl0param::α,
                                                sk.key.lvl0);
            uint32_t* d = upload<lvl0param>(c0);
            uint32_t* o = dev_alloc<lvl0param>(lvl0param::n + 1);
            cufhe::Bootstrap(o, d, lvl1param::μ, st, 0);
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != b;
            cufhe::Bootstrap(d, d, lvl1param::μ, st, 0);  // out aliases in
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(d), sk.key.lvl0) != b;
            uint32_t* trlwe = dev_alloc<lvl1param>(2 * lvl1param::n);
            cufhe::BootstrapTLWE2TRLWE(trlwe, d, lvl1param::μ, st, 0);
            cufhe::SEIandKS(o, trlwe, st, 0);
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != b;
            auto c1 = tlweSymEncrypt<lvl1param>(b ? lvl1param::μ : -lvl1param::μ, lvl1param::α,
                                                sk.key.lvl1);
            uint32_t* d1 = upload<lvl1param>(c1);
            cufhe::IdentityKeySwitch<lvl10param>(o, d1, st, 0);
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != b;
            auto c2 = tlweSymEncrypt<lvl2param>(b ? lvl2param::μ : -lvl2param::μ, lvl2param::α,
                                                sk.key.lvl2);
            uint64_t* d2 = dev_alloc<lvl2param>(lvl2param::n + 1);
            cufhe::CuCheck(cufhe_memcpy_h2d(d2, c2.data(), sizeof(c2), st), "h2d");
            cufhe::Synchronize();
            cufhe::IdentityKeySwitch<lvl20param>(o, d2, st, 0);
            e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != b;
            total += 5;
            cufhe_device_free(d);
            cufhe_device_free(o);
            cufhe_device_free(trlwe);
            cufhe_device_free(d1);
            cufhe_device_free(d2);
        }
        bad += report("Bootstrap / SEIandKS / IdentityKeySwitch", e, total);
    }
    // TRGSW2NTT / CMUXNTT (reference src/cufhe_gates_gpu.cu:116-133,
    // src/bootstrap_gpu.cu:165-186): the result decrypts to c1's message when
    // the TRGSW encrypts 1 and to c0's when it encrypts 0; CMUXNTTkernel on
    // the same device operands gives the same TRLWE
    {
        int e = 0, total = 0;
        for (int i = 0; i < 6; i++) {
            const int b = i & 1;
            TRGSW<lvl1param> g;
            trgswSymEncrypt(g, b, sk.key.lvl1);
            cufhe::cuFHETRGSWNTTlvl1 cs;
            cufhe::TRGSW2NTT(cs, g, s);
            cufhe::cuFHETRLWElvl1 c1, c0, res;
            Polynomial<lvl1param> p1, p0;
            for (uint32_t k = 0; k < lvl1param::n; k++) {
                p1[k] = (rng() & 1) ? mu1 : 0u - mu1;
                p0[k] = (rng() & 1) ? mu1 : 0u - mu1;
            }
            trlweSymEncrypt(c1.trlwehost, p1, sk.key.lvl1);
            trlweSymEncrypt(c0.trlwehost, p0, sk.key.lvl1);
            cufhe::CMUXNTT(res, cs, c1, c0, s);
            cufhe::Synchronize();
            // the raw-pointer form on the same device operands gives the
            // same TRLWE
            cufhe::cuFHETRLWElvl1 raw;
            cufhe::CMUXNTTkernel(raw.trlwedevices[0], cs.trgswdevices[0], c1.trlwedevices[0],
                                 c0.trlwedevices[0], st, 0);
            cufhe::CuCheck(cufhe_memcpy_d2h(raw.trlwehost.data(), raw.trlwedevices[0],
                                            sizeof(raw.trlwehost), st), "d2h");
            cufhe::Synchronize();
            e += raw.trlwehost != res.trlwehost;
            // SampleExtract of coefficient 0 on the host
            TLWE<lvl1param> t;
            t[0] = res.trlwehost[0][0];
            for (uint32_t k = 1; k < lvl1param::n; k++) t[k] = 0u - res.trlwehost[0][lvl1param::n - k];
            t[lvl1param::n] = res.trlwehost[1][0];
            const uint32_t want = b ? p1[0] : p0[0];
            e += tlweSymDecrypt<lvl1param>(t, sk.key.lvl1) != (want == mu1);
            total++;
        }
        bad += report("TRGSW2NTT / CMUXNTT", e, total);
    }
    {
        // BootstrappingKeyToNTT (and the flat lvl02 form) /
        // KeySwitchingKeyToDevice and the Delete* functions: one key at a
        // time; gates are refused until a level has both keys, then run
        // correctly
        int e = 0, total = 0;
        auto c0 = tlweSymEncrypt<lvl0param>(lvl0param::μ, lvl0param::α, sk.key.lvl0);
        auto c1 = tlweSymEncrypt<lvl0param>(-lvl0param::μ, lvl0param::α, sk.key.lvl0);
        uint32_t *d0 = upload<lvl0param>(c0), *d1 = upload<lvl0param>(c1);
        uint32_t* o = dev_alloc<lvl0param>(lvl0param::n + 1);
        cufhe::DeleteKeySwitchingKey(1);
        e += cufhe_gate_lvl0(CUFHE_NAND, o, d0, d1, nullptr, st) != CUFHE_ERR_STATE;
        cufhe::DeleteBootstrappingKeyNTT(1);
        cufhe::BootstrappingKeyToNTT<lvl01param>(ek.getbk<lvl01param>(), 1);
        e += cufhe_gate_lvl0(CUFHE_NAND, o, d0, d1, nullptr, st) != CUFHE_ERR_STATE;
        cufhe::KeySwitchingKeyToDevice(ek.getiksk<lvl10param>(), 1);
        total += 2;
        cufhe::DeleteKeySwitchingKey_lvl20(1);
        cufhe::DeleteBootstrappingKeyNTT_lvl02(1);
        cufhe::KeySwitchingKeyToDevice_lvl20(ek.getiksk<lvl20param>(), 1);
        cufhe::BootstrappingKeyToNTT<lvl02param>(ek.getbk<lvl02param>(), 1);
        // the same key as a flat TRGSW array
        static_assert(sizeof(TRGSW<lvl2param>) == kTRGSWWords<lvl2param> * sizeof(uint64_t));
        const auto& bk2 = ek.getbk<lvl02param>();
        cufhe::BootstrappingKeyFlatToNTT_lvl02(reinterpret_cast<const TRGSW<lvl2param>*>(bk2.data()),
                                               static_cast<uint32_t>(bk2.size() / kTRGSWWords<lvl2param>), 1);
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++) {
                uint32_t* a = x ? d0 : d1;
                uint32_t* b = y ? d0 : d1;
                cufhe::NandBootstrap<lvl01param, mu1, lvl10param>(o, a, b, st, 0);
                e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != (1 - x * y);
                cufhe::NandBootstrap<lvl02param, mu2, lvl20param>(o, a, b, st, 0);
                e += tlweSymDecrypt<lvl0param>(download<lvl0param>(o), sk.key.lvl0) != (1 - x * y);
                total += 2;
            }
        for (uint32_t* p : {d0, d1, o}) cufhe_device_free(p);
        bad += report("BootstrappingKeyToNTT / KeySwitchingKeyToDevice", e, total);
    }
    {
        // ctxtInitialize / ctxtDelete: device copies of a host TLWE, a
        // round trip through them is exact
        TLWE<lvl1param> h, back{};
        for (uint32_t k = 0; k <= lvl1param::n; k++) h[k] = 0x9E3779B9u * (k + 1);
        std::vector<uint32_t*> dev;
        cufhe::ctxtInitialize(h, dev);
        int e = dev.size() != 1 || dev[0] == nullptr;
        if (!e) {
            cufhe::CuCheck(cufhe_memcpy_h2d(dev[0], h.data(), sizeof(h), s.st()), "h2d");
            cufhe::CuCheck(cufhe_memcpy_d2h(back.data(), dev[0], sizeof(h), s.st()), "d2h");
            cufhe::Synchronize();
            e = back != h;
        }
        cufhe::ctxtDelete(h, dev);
        e += dev[0] != nullptr;
        CuSafeCall(hipStreamSynchronize(st));  // the reference's error helpers
        CuCheckError();
        bad += report("ctxtInitialize / ctxtDelete", e, 1);
    }
    s.Destroy();
    cufhe::CleanUp();
    std::printf("%s\n", bad == 0 ? "ALL PASS" : "FAIL");
    return bad == 0 ? 0 : 1;
}
