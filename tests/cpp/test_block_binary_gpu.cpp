// Drop-in check of the block-binary configuration through the C++ gate API,
// covering what the reference's block-binary test covers (reference
// test/test_block_binary_gpu.cc:9-169): the key structure (blocks of ell with
// at most one bit, lvl1 key starting with the lvl0 key), subset
// SampleExtractAndKeySwitch of a TRLWE, and NAND / MUX truth tables on lvl0
// and lvl1 ciphertexts with a block-binary bootstrapping key and the subset
// key-switching key. Prints PASS/FAIL lines, returns 0 when all pass.
#include <cstdio>

#include "cufhe_gpu.hpp"

using namespace TFHEpp;

namespace {

bool key_structure_ok(const SecretKey& sk)
{
    constexpr uint32_t ell = lvl0bbparam::ell;
    for (uint32_t blk = 0; blk < lvl0param::n / ell; blk++) {
        uint32_t ones = 0;
        for (uint32_t o = 0; o < ell; o++) {
            const uint32_t i = blk * ell + o;
            ones += sk.key.lvl0[i];
            if (sk.key.lvl0[i] != sk.key.lvl1[i]) {
                std::printf("lvl1 key does not start with the lvl0 key (index %u)\n", i);
                return false;
            }
        }
        if (ones > 1) {
            std::printf("block %u holds %u ones\n", blk, ones);
            return false;
        }
    }
    return true;
}

template <class P>
typename P::T enc_bit(int b)
{
    return b ? P::μ : static_cast<typename P::T>(0 - P::μ);
}

template <class P>
bool nand_table(const char* label, const SecretKey& sk, cufhe::Stream st)
{
    cufhe::Ctxt<P> x, y, z;
    bool ok = true;
    for (int a = 0; a <= 1; a++)
        for (int b = 0; b <= 1; b++) {
            tlweSymEncrypt<P>(x.tlwehost, enc_bit<P>(a), sk.key.get<P>());
            tlweSymEncrypt<P>(y.tlwehost, enc_bit<P>(b), sk.key.get<P>());
            cufhe::Nand<P>(z, x, y, st);
            cufhe::Synchronize();
            const bool got = tlweSymDecrypt<P>(z.tlwehost, sk.key.get<P>());
            if (got != !(a && b)) {
                std::printf("%s NAND(%d, %d) = %d\n", label, a, b, got);
                ok = false;
            }
        }
    std::printf("%s NAND: %s\n", label, ok ? "PASS" : "FAIL");
    return ok;
}

template <class P>
bool mux_table(const char* label, const SecretKey& sk, cufhe::Stream st)
{
    cufhe::Ctxt<P> c, t, f, z;
    bool ok = true;
    for (int s = 0; s <= 1; s++)
        for (int a = 0; a <= 1; a++)
            for (int b = 0; b <= 1; b++) {
                tlweSymEncrypt<P>(c.tlwehost, enc_bit<P>(s), sk.key.get<P>());
                tlweSymEncrypt<P>(t.tlwehost, enc_bit<P>(a), sk.key.get<P>());
                tlweSymEncrypt<P>(f.tlwehost, enc_bit<P>(b), sk.key.get<P>());
                cufhe::Mux<P>(z, c, t, f, st);
                cufhe::Synchronize();
                const bool got = tlweSymDecrypt<P>(z.tlwehost, sk.key.get<P>());
                if (got != static_cast<bool>(s ? a : b)) {
                    std::printf("%s MUX(%d ? %d : %d) = %d\n", label, s, a, b, got);
                    ok = false;
                }
            }
    std::printf("%s MUX: %s\n", label, ok ? "PASS" : "FAIL");
    return ok;
}

bool subset_sample_extraction(const SecretKey& sk, cufhe::Stream st)
{
    cufhe::cuFHETRLWElvl1 in;
    cufhe::Ctxt<lvl0param> out;
    bool ok = true;
    for (int m = 0; m <= 1; m++) {
        Polynomial<lvl1param> p{};
        p[0] = enc_bit<lvl1param>(m);
        trlweSymEncrypt<lvl1param>(in.trlwehost, p, sk.key.lvl1);
        cufhe::SampleExtractAndKeySwitch(out, in, st);
        cufhe::Synchronize();
        const bool got = tlweSymDecrypt<lvl0param>(out.tlwehost, sk.key.lvl0);
        if (got != static_cast<bool>(m)) {
            std::printf("subset sample extraction of %d gave %d\n", m, got);
            ok = false;
        }
    }
    std::printf("subset SampleExtractAndKeySwitch: %s\n", ok ? "PASS" : "FAIL");
    return ok;
}

}  // namespace

int main()
{
    cufhe::SetGPUNum(1);
    SecretKey sk(7, /*block_binary=*/true);
    if (!key_structure_ok(sk)) {
        std::printf("block-binary key structure: FAIL\n");
        return 1;
    }
    std::printf("block-binary key structure: PASS\n");
    EvalKey ek(sk);
    ek.emplacebk<lvl01param>(sk);
    ek.emplacesubiksk<lvl10param>(sk);
    cufhe::Initialize(ek);
    cufhe::Stream st;
    st.Create();
    bool ok = subset_sample_extraction(sk, st);
    ok = nand_table<lvl0param>("lvl0", sk, st) && ok;
    ok = nand_table<lvl1param>("lvl1", sk, st) && ok;
    ok = mux_table<lvl0param>("lvl0", sk, st) && ok;
    ok = mux_table<lvl1param>("lvl1", sk, st) && ok;
    st.Destroy();
    cufhe::CleanUp();
    std::printf("%s\n", ok ? "ALL PASS" : "SOME FAILED");
    return ok ? 0 : 1;
}
