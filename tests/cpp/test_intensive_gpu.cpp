// Many-stream stress test of the plain per-gate form, written the way the
// reference's intensive test is (reference test/test_intensive.cc:1-116):
// M streams each run a chain of N gates, the host polls StreamQuery over all
// streams and launches a stream's next gate as soon as its previous one is
// done; afterwards every output must decrypt to the expected bit. Nand and
// Mux on lvl0 ciphertexts (BR-IKS with the lvl1 keys). Prints ALL PASS or
// the failure counts, returns 0/1.
#include <cstdio>
#include <memory>
#include <vector>

#include "cufhe_gpu.hpp"

using namespace TFHEpp;
using P = lvl0param;

namespace {

constexpr size_t kGates = 10, kStreams = 1000;

template <class Launch, class Check>
int run_and_verify(const char* name, Launch&& launch, Check&& check)
{
    std::vector<cufhe::Stream> st(kStreams);
    for (auto& s : st) s.Create();
    std::vector<size_t> next(kStreams, 0);
    for (size_t j = 0; j < kStreams; j++) launch(0, j, st[j]);
    size_t polls = 0;
    while (true) {
        bool pending = false;
        for (size_t j = 0; j < kStreams; j++) {
            if (next[j] == kGates) continue;
            pending = true;
            if (cufhe::StreamQuery(st[j])) {
                const size_t i = ++next[j];
                if (i < kGates) launch(i, j, st[j]);
            }
        }
        polls++;
        if (!pending) break;
    }
    cufhe::Synchronize();
    size_t errors = 0;
    for (size_t i = 0; i < kGates; i++)
        for (size_t j = 0; j < kStreams; j++) errors += !check(i, j);
    std::printf("%-6s %zu streams x %zu gates: %zu errors (%zu polling sweeps)  %s\n", name,
                kStreams, kGates, errors, polls, errors ? "FAIL" : "PASS");
    for (auto& s : st) s.Destroy();
    return errors ? 1 : 0;
}

}  // namespace

int main()
{
    SecretKey sk(5);
    EvalKey ek;
    ek.emplacebk<lvl01param>(sk);
    ek.emplaceiksk<lvl10param>(sk);
    cufhe::SetGPUNum(1);
    cufhe::Initialize(ek);
    int bad = 0;

    // Nand(true, false) = true
    {
        cufhe::Ctxt<P> ca, cb;
        ca.tlwehost = tlweSymEncrypt<P>(P::μ, P::α, sk.key.lvl0);
        cb.tlwehost = tlweSymEncrypt<P>(-P::μ, P::α, sk.key.lvl0);
        std::vector<std::unique_ptr<cufhe::Ctxt<P>>> res(kGates * kStreams);
        for (auto& r : res) r.reset(new cufhe::Ctxt<P>());
        bad += run_and_verify(
            "nand",
            [&](size_t i, size_t j, cufhe::Stream s) { cufhe::Nand(*res[i * kStreams + j], ca, cb, s); },
            [&](size_t i, size_t j) { return tlweSymDecrypt<P>(res[i * kStreams + j]->tlwehost, sk.key.lvl0) == 1; });
    }
    // Mux(c = true, t = false, f = true) = false
    {
        cufhe::Ctxt<P> cc, ct, cf;
        cc.tlwehost = tlweSymEncrypt<P>(P::μ, P::α, sk.key.lvl0);
        ct.tlwehost = tlweSymEncrypt<P>(-P::μ, P::α, sk.key.lvl0);
        cf.tlwehost = tlweSymEncrypt<P>(P::μ, P::α, sk.key.lvl0);
        std::vector<std::unique_ptr<cufhe::Ctxt<P>>> res(kGates * kStreams);
        for (auto& r : res) r.reset(new cufhe::Ctxt<P>());
        bad += run_and_verify(
            "mux",
            [&](size_t i, size_t j, cufhe::Stream s) { cufhe::Mux(*res[i * kStreams + j], cc, ct, cf, s); },
            [&](size_t i, size_t j) { return tlweSymDecrypt<P>(res[i * kStreams + j]->tlwehost, sk.key.lvl0) == 0; });
    }
    cufhe::CleanUp();
    std::printf(bad ? "SOME FAILED\n" : "ALL PASS\n");
    return bad ? 1 : 0;
}
