// Key files of the host layer (tfhe_keyfile_save/_load through the C++
// SecretKey/EvalKey save and load): round trips bit for bit, flags kept,
// a loaded secret key decrypts what the original encrypted, and malformed
// or truncated files and wrongly sized arrays are refused. CPU only.
// Usage: test_keyfile.bin <scratch directory>. Prints ALL PASS, returns 0/1.
#include <cstdio>
#include <string>
#include <vector>

#include "tfhe_host.hpp"

using namespace TFHEpp;

namespace {
int bad = 0;
void check(bool ok, const char* what)
{
    std::printf("%-56s %s\n", what, ok ? "PASS" : "FAIL");
    bad += !ok;
}
template <class V>
bool same(const V& a, const V& b)
{
    if (!a || !b) return !a && !b;
    return *a == *b;
}
}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    SecretKey sk(7);
    EvalKey ek;
    ek.emplacebk<lvl01param>(sk);
    ek.emplaceiksk<lvl10param>(sk);
    ek.emplaceiksk<lvl20param>(sk);
    ek.ntt = true;
    const std::string ekp = dir + "/ek.bin", skp = dir + "/sk.bin";
    check(ek.save(ekp.c_str()), "EvalKey::save");
    EvalKey ek2;
    check(ek2.load(ekp.c_str()), "EvalKey::load");
    check(same(ek.bklvl01, ek2.bklvl01) && same(ek.iksklvl10, ek2.iksklvl10) &&
              same(ek.iksklvl20, ek2.iksklvl20) && !ek2.bklvl02,
          "arrays identical, absent array stays absent");
    check(ek2.key_bundle == ek.key_bundle && ek2.ntt && !ek2.block_binary && !ek2.subset_iksk,
          "flags kept");

    check(sk.save(skp.c_str()), "SecretKey::save");
    SecretKey sk2(99, true);
    check(sk2.load(skp.c_str()) && sk2.key.lvl0 == sk.key.lvl0 && sk2.key.lvl1 == sk.key.lvl1 &&
              sk2.key.lvl2 == sk.key.lvl2 && !sk2.block_binary,
          "SecretKey round trip");
    int dec_bad = 0;
    for (int i = 0; i < 32; i++) {
        const bool m = (i * 7) & 1;
        auto c = tlweSymEncrypt<lvl1param>(m ? lvl1param::μ : -lvl1param::μ, lvl1param::α, sk.key.lvl1);
        dec_bad += tlweSymDecrypt<lvl1param>(c, sk2.key.lvl1) != m;
    }
    check(dec_bad == 0, "loaded secret key decrypts");

    // truncated file: every record's size is checked
    {
        FILE* in = std::fopen(ekp.c_str(), "rb");
        std::vector<char> buf(1 << 20);
        const size_t n = std::fread(buf.data(), 1, buf.size(), in);
        std::fclose(in);
        const std::string tp = dir + "/trunc.bin";
        FILE* out = std::fopen(tp.c_str(), "wb");
        std::fwrite(buf.data(), 1, n / 2, out);
        std::fclose(out);
        EvalKey e3;
        check(!e3.load(tp.c_str()), "truncated file refused");
        const std::string gp = dir + "/garbage.bin";
        out = std::fopen(gp.c_str(), "wb");
        std::fwrite("NOTAKEY!", 1, 8, out);
        std::fclose(out);
        check(!e3.load(gp.c_str()), "wrong magic refused");
    }
    // a wrongly sized array is refused at save time
    {
        EvalKey e4;
        e4.bklvl01 = std::make_unique<std::vector<uint32_t>>(1000);
        check(!e4.save((dir + "/bad.bin").c_str()), "wrongly sized array refused");
    }
    std::printf(bad ? "SOME FAILED\n" : "ALL PASS\n");
    return bad ? 1 : 0;
}
