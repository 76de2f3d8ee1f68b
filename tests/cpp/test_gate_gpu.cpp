// Drop-in check of the C++ gate API, written the way the reference's own
// gate test is (reference test/test_gate_gpu.cc:36-107, test/test_util.h:8-95):
// encrypt random bits with the TFHEpp-style host layer, run every gate through
// cufhe::Ctxt / cufhe::Stream on several streams, decrypt, compare with the
// truth tables (reference test/plain.h:10-69). Prints PASS/FAIL, returns 0/1.
#include <cstdio>
#include <functional>
#include <memory>
#include <random>
#include <vector>

#include "cufhe_gpu.hpp"

using namespace TFHEpp;
using P = lvl1param;

int main()
{
    constexpr int kStreams = 8, kPerGate = 24;
    SecretKey sk(0);
    EvalKey ek;
    ek.emplacebk<lvl01param>(sk);
    ek.emplaceiksk<lvl10param>(sk);
    cufhe::SetGPUNum(1);
    cufhe::Initialize(ek);
    std::vector<cufhe::Stream> st(kStreams);
    for (auto& s : st) s.Create();

    using Gate = std::function<void(cufhe::Ctxt<P>&, cufhe::Ctxt<P>&, cufhe::Ctxt<P>&, cufhe::Stream)>;
    struct Case {
        const char* name;
        Gate gate;
        std::function<int(int, int)> plain;
    };
    const std::vector<Case> cases = {
        {"Nand", cufhe::Nand<P>, [](int a, int b) { return 1 - a * b; }},
        {"Or", cufhe::Or<P>, [](int a, int b) { return (a + b) > 0; }},
        {"OrYN", cufhe::OrYN<P>, [](int a, int b) { return (a + (1 - b)) > 0; }},
        {"OrNY", cufhe::OrNY<P>, [](int a, int b) { return ((1 - a) + b) > 0; }},
        {"And", cufhe::And<P>, [](int a, int b) { return a * b; }},
        {"AndYN", cufhe::AndYN<P>, [](int a, int b) { return a * (1 - b); }},
        {"AndNY", cufhe::AndNY<P>, [](int a, int b) { return (1 - a) * b; }},
        {"Nor", cufhe::Nor<P>, [](int a, int b) { return 1 - ((a + b) > 0); }},
        {"Xor", cufhe::Xor<P>, [](int a, int b) { return (a + b) & 1; }},
        {"Xnor", cufhe::Xnor<P>, [](int a, int b) { return 1 - ((a + b) & 1); }},
    };
    std::mt19937 rng(2024);
    int errors = 0, total = 0;
    for (const Case& c : cases) {
        std::vector<std::unique_ptr<cufhe::Ctxt<P>>> a, b, o;
        std::vector<int> pa, pb;
        for (int i = 0; i < kPerGate; i++) {
            a.emplace_back(new cufhe::Ctxt<P>());
            b.emplace_back(new cufhe::Ctxt<P>());
            o.emplace_back(new cufhe::Ctxt<P>());
            pa.push_back(rng() & 1);
            pb.push_back(rng() & 1);
            a[i]->tlwehost = tlweSymEncrypt<P>(pa[i] ? P::μ : -P::μ, P::α, sk.key.lvl1);
            b[i]->tlwehost = tlweSymEncrypt<P>(pb[i] ? P::μ : -P::μ, P::α, sk.key.lvl1);
        }
        for (int i = 0; i < kPerGate; i++) c.gate(*o[i], *a[i], *b[i], st[i % kStreams]);
        cufhe::Synchronize();
        for (int i = 0; i < kPerGate; i++) {
            const int got = tlweSymDecrypt<P>(o[i]->tlwehost, sk.key.lvl1);
            errors += got != c.plain(pa[i], pb[i]);
            total++;
        }
    }
    // Mux(out, inc, in1, in0) = inc ? in1 : in0, and Not / Copy
    for (int i = 0; i < kPerGate; i++) {
        cufhe::Ctxt<P> c, t, f, o1, o2, o3, o4;
        const int pc = rng() & 1, pt = rng() & 1, pf = rng() & 1;
        c.tlwehost = tlweSymEncrypt<P>(pc ? P::μ : -P::μ, P::α, sk.key.lvl1);
        t.tlwehost = tlweSymEncrypt<P>(pt ? P::μ : -P::μ, P::α, sk.key.lvl1);
        f.tlwehost = tlweSymEncrypt<P>(pf ? P::μ : -P::μ, P::α, sk.key.lvl1);
        cufhe::Stream s = st[i % kStreams];
        cufhe::Mux<P>(o1, c, t, f, s);
        cufhe::NMux<P>(o2, c, t, f, s);
        cufhe::Not<P>(o3, c, s);
        cufhe::Copy<P>(o4, t, s);
        cufhe::Synchronize();
        errors += tlweSymDecrypt<P>(o1.tlwehost, sk.key.lvl1) != (pc ? pt : pf);
        errors += tlweSymDecrypt<P>(o2.tlwehost, sk.key.lvl1) != !(pc ? pt : pf);
        errors += tlweSymDecrypt<P>(o3.tlwehost, sk.key.lvl1) != !pc;
        errors += tlweSymDecrypt<P>(o4.tlwehost, sk.key.lvl1) != pt;
        total += 4;
    }
    for (auto& s : st) s.Destroy();
    cufhe::CleanUp();
    std::printf("%d / %d gates decrypted correctly\n", total - errors, total);
    std::printf("%s\n", errors == 0 ? "PASS" : "FAIL");
    return errors == 0 ? 0 : 1;
}
