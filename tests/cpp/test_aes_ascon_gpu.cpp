// Drop-in check of the C++ AES-128 decryption and ASCON-XOF APIs
// (include/aes_gpu.hpp, include/ascon_gpu.hpp), called the way the
// reference's AESDec / ASCONXOF are (include/aes_gpu.cuh:27-32,
// include/ascon_gpu.cuh:40-43):
//  * AESKeyExpansion gives FIPS-197 Appendix A.1's last round key;
//  * the FIPS-197 Appendix C.1 and Appendix B blocks, encrypted bit by bit
//    with their expanded keys, decrypt through AESDec called from two host
//    threads at once (their layers share serving rounds);
//  * ASCONXOF of the empty message gives the published Ascon v1.2 Ascon-Xof
//    digest, and of the bytes 00 .. 0f the oracle's value
//    (oracle/ascon_ref.py, pinned by that digest), and ASCONXOFInitialize /
//    Absorb / Squeeze over an encrypted state give the same outputs.
// Prints PASS/FAIL, returns 0/1.
#include <array>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "aes_gpu.hpp"
#include "ascon_gpu.hpp"

using namespace TFHEpp;
using P = lvl1param;

static std::array<uint8_t, 16> hex16(const char* h)
{
    std::array<uint8_t, 16> b{};
    for (int i = 0; i < 16; i++) b[i] = static_cast<uint8_t>(std::stoi(std::string(h + 2 * i, 2), nullptr, 16));
    return b;
}

static std::string hex(const std::vector<uint8_t>& b)
{
    std::string s;
    char buf[3];
    for (uint8_t v : b) {
        std::snprintf(buf, sizeof buf, "%02x", v);
        s += buf;
    }
    return s;
}

template <std::size_t N>
static void encrypt_bytes(TLWE<P>* out, const std::array<uint8_t, N>& bytes, const SecretKey& sk)
{
    for (std::size_t i = 0; i < 8 * N; i++) {
        const bool bit = (bytes[i / 8] >> (i % 8)) & 1;
        out[i] = tlweSymEncrypt<P>(bit ? P::μ : -P::μ, P::α, sk.key.lvl1);
    }
}

static std::vector<uint8_t> decrypt_bytes(const TLWE<P>* in, std::size_t bits, const SecretKey& sk)
{
    std::vector<uint8_t> out(bits / 8, 0);
    for (std::size_t i = 0; i < bits; i++)
        if (tlweSymDecrypt<P>(in[i], sk.key.lvl1)) out[i / 8] |= static_cast<uint8_t>(1u << (i % 8));
    return out;
}

int main()
{
    SecretKey sk(0);
    EvalKey ek;
    ek.emplacebk<lvl01param>(sk);
    ek.emplaceiksk<lvl10param>(sk);
    cufhe::SetGPUNum(1);
    cufhe::Initialize(ek);
    int failures = 0;

    const auto a1 = cufhe::AESKeyExpansion(hex16("2b7e151628aed2a6abf7158809cf4f3c"));
    const bool ks_ok = hex(std::vector<uint8_t>(a1[10].begin(), a1[10].end())) ==
                       "d014f9a8c9ee2589e13f0cc8b6630ca6";
    std::printf("AESKeyExpansion FIPS-197 A.1 round key 10: %s\n", ks_ok ? "ok" : "WRONG");
    failures += !ks_ok;

    struct Block {
        const char *key, *ct, *pt;
    };
    const Block blocks[2] = {
        {"000102030405060708090a0b0c0d0e0f", "69c4e0d86a7b0430d8cdb78070b4c55a",
         "00112233445566778899aabbccddeeff"},  // FIPS-197 Appendix C.1
        {"2b7e151628aed2a6abf7158809cf4f3c", "3925841d02dc09fbdc118597196a0b32",
         "3243f6a8885a308d313198a2e0370734"},  // FIPS-197 Appendix B
    };
    cufhe::InitializeAES(ek, sk);
    std::vector<std::array<TLWE<P>, 128>> cipher(2), plain(2);
    std::vector<std::array<std::array<TLWE<P>, 128>, 11>> rk(2);
    for (int b = 0; b < 2; b++) {
        encrypt_bytes(cipher[b].data(), hex16(blocks[b].ct), sk);
        const auto exp = cufhe::AESKeyExpansion(hex16(blocks[b].key));
        for (int r = 0; r < 11; r++) encrypt_bytes(rk[b][r].data(), exp[r], sk);
    }
    std::vector<cufhe::Stream> sts(2);
    std::vector<std::thread> th;
    for (int b = 0; b < 2; b++)
        th.emplace_back([&, b] { cufhe::AESDec(plain[b], cipher[b], rk[b], ek, sts[b]); });
    for (auto& t : th) t.join();
    for (int b = 0; b < 2; b++) {
        const std::string got = hex(decrypt_bytes(plain[b].data(), 128, sk));
        const bool ok = got == blocks[b].pt;
        std::printf("AESDec block %d: %s (%s)\n", b, got.c_str(), ok ? "ok" : "WRONG");
        failures += !ok;
    }
    cufhe::CleanUpAES();

    cufhe::InitializeASCON(ek, sk);
    std::vector<TLWE<P>> out(256), empty;
    cufhe::ASCONXOF(out, empty);
    const std::string d0 = hex(decrypt_bytes(out.data(), out.size(), sk));
    const bool e_ok = d0 == "5d4cbde6350ea4c174bd65b5b332f8408f99740b81aa02735eaefbcf0ba0339e";
    std::printf("ASCONXOF(empty, 32): %s (%s)\n", d0.c_str(), e_ok ? "ok" : "WRONG");
    failures += !e_ok;
    std::array<uint8_t, 16> msg{};
    for (int i = 0; i < 16; i++) msg[i] = static_cast<uint8_t>(i);
    std::vector<TLWE<P>> in(128);
    encrypt_bytes(in.data(), msg, sk);
    cufhe::Stream st;
    cufhe::ASCONXOF(out, in, ek, st);
    const std::string d1 = hex(decrypt_bytes(out.data(), out.size(), sk));
    const bool m_ok = d1 == "c861a89cfb1335f278c96cf7ffc9753c290cbe1a4e186d2923b496bb4ea5e519";
    std::printf("ASCONXOF(00..0f, 32): %s (%s)\n", d1.c_str(), m_ok ? "ok" : "WRONG");
    failures += !m_ok;
    // the incremental form over an encrypted state: the same digests, and a
    // one-block squeeze (no gates) gives their first 8 bytes
    cufhe::ASCONState state;
    cufhe::ASCONXOFInitialize(state, ek, st);
    cufhe::ASCONXOFAbsorb(state, in, ek, st);
    std::vector<TLWE<P>> head(64);
    cufhe::ASCONState copy = state;
    cufhe::ASCONXOFSqueeze(copy, head);
    cufhe::ASCONXOFSqueeze(state, out, ek, st);
    const std::string d2 = hex(decrypt_bytes(out.data(), out.size(), sk));
    const std::string h2 = hex(decrypt_bytes(head.data(), head.size(), sk));
    const bool i_ok = d2 == d1 && h2 == d1.substr(0, 16);
    std::printf("ASCONXOFInitialize/Absorb/Squeeze(00..0f, 32): %s, first block %s (%s)\n",
                d2.c_str(), h2.c_str(), i_ok ? "ok" : "WRONG");
    failures += !i_ok;
    cufhe::ASCONXOFInitialize(state);
    cufhe::ASCONXOFAbsorb(state, empty);
    cufhe::ASCONXOFSqueeze(state, out);
    const std::string d3 = hex(decrypt_bytes(out.data(), out.size(), sk));
    const bool z_ok = d3 == d0;
    std::printf("ASCONXOFInitialize/Absorb/Squeeze(empty, 32): %s (%s)\n", d3.c_str(), z_ok ? "ok" : "WRONG");
    failures += !z_ok;
    cufhe::CleanUpASCON();

    cufhe::CleanUp();
    std::printf("%s\n", failures ? "FAIL" : "PASS");
    return failures ? 1 : 0;
}
