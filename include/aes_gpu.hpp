// C++ drop-in for cuFHEpp's AES-128 decryption (reference
// include/aes_gpu.cuh:21-32, src/aes_gpu.cu:904-1118 AESDec): the same names
// and argument meaning over the C ABI of libcufhe_amd.so. The decryption is
// a circuit of lvl1 gates (cufhe_circuit_aes128_decrypt,
// include/cufhe_circuits.h) run by the library's serving loop.
//
// Differences a caller can see:
//  * lvl1 TLWEs throughout: the engine's gates take and give lvl1
//    ciphertexts, where the reference's expanded key and plaintext are TLWEs
//    of its circuit-bootstrapping target level; its template parameters
//    (iksP, brP, ahP) select those levels and are not needed here;
//  * the gate keys loaded by cufhe::Initialize(ek) are all the circuit uses:
//    InitializeAES and AESDec accept the reference's EvalKey / SecretKey
//    arguments and do not read them;
//  * AESDec runs on the serving loop's own stream and returns when the
//    plaintext is in `plain`; the Stream argument is accepted and not used.
//    Calls from several host threads are batched into the same rounds.
//  * AESKeyExpansion gives the FIPS-197 5.2 round keys on the key owner's
//    side (the reference takes them from TFHEpp), bit b of byte i being
//    TLWE 8 i + b of a round key.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "cufhe_amd.h"
#include "cufhe_circuits.h"
#include "cufhe_gpu.hpp"
#include "tfhe_host.hpp"

namespace cufhe {

namespace detail_circ {
inline void check(int rc, const char* what)
{
    if (rc != CUFHE_OK) {
        std::fprintf(stderr, "cufhe: %s failed (%d): %s\n", what, rc, cufhe_last_error());
        std::exit(-1);
    }
}

// A circuit from include/cufhe_circuits.h and its descriptor.
struct Circuit {
    std::vector<cufhe_gate_desc> gates;
    std::vector<int> layer_end;
    std::vector<uint32_t> outputs;
    cufhe_circuit desc{};
    template <class F>
    void build(F fill, const char* what)
    {
        cufhe_circuit_sizes sz{};
        check(fill(&sz, nullptr, nullptr, nullptr), what);
        gates.resize(sz.n_gates);
        layer_end.resize(sz.n_layers);
        outputs.resize(sz.n_outputs);
        check(fill(&sz, gates.data(), layer_end.data(), outputs.data()), what);
        desc = cufhe_circuit{sz.n_inputs, sz.n_gates, gates.data(), sz.n_layers,
                             layer_end.data(), sz.n_outputs, outputs.data()};
    }
    std::size_t wire_bytes() const
    {
        return static_cast<std::size_t>(desc.n_inputs + desc.n_gates) *
               (TFHEpp::lvl1param::n + 1) * sizeof(uint32_t);
    }
};

// Runs one request and waits for it.
inline void run(cufhe_server* server, const Circuit& c, const std::vector<uint32_t>& in,
                std::vector<uint32_t>& out, const char* what)
{
    uint64_t rid = 0;
    out.assign(static_cast<std::size_t>(c.desc.n_outputs) * (TFHEpp::lvl1param::n + 1), 0);
    check(cufhe_server_submit(server, &c.desc, in.data(), out.data(), &rid), what);
    const int rc = cufhe_server_wait(server, rid, -1);
    check(rc == 1 ? CUFHE_OK : rc, what);
}

// Noiseless encryptions of 0 and 1 (a = 0, b = -+mu): the circuits' constants.
inline void append_constants(std::vector<uint32_t>& in)
{
    using P = TFHEpp::lvl1param;
    for (int v = 0; v < 2; v++) {
        in.insert(in.end(), P::n, 0u);
        in.push_back(v ? static_cast<uint32_t>(P::μ) : static_cast<uint32_t>(0u - P::μ));
    }
}

template <class TLWE>
inline void append(std::vector<uint32_t>& in, const TLWE& c)
{
    in.insert(in.end(), c.begin(), c.end());
}
}  // namespace detail_circ

namespace detail_aes {
struct State {
    std::mutex mu;
    cufhe_server* server = nullptr;
    detail_circ::Circuit circuit;
};
inline State& state()
{
    static State s;
    return s;
}
constexpr int kRequestsInPool = 8;      // concurrent AESDec calls held on the GPU
constexpr int kBatchBootstraps = 16384;  // per serving round
}  // namespace detail_aes

// Builds the circuit and starts the serving loop on the loaded lvl1 gate
// keys (call cufhe::Initialize(ek) first).
inline void InitializeAES()
{
    auto& s = detail_aes::state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.server) return;
    s.circuit.build(cufhe_circuit_aes128_decrypt, "cufhe_circuit_aes128_decrypt");
    detail_circ::check(cufhe_server_create(detail_aes::kRequestsInPool * s.circuit.wire_bytes(),
                                           detail_aes::kBatchBootstraps, 0, &s.server),
                       "InitializeAES");
}
inline void InitializeAES(const TFHEpp::EvalKey&, const TFHEpp::SecretKey&) { InitializeAES(); }

inline void CleanUpAES()
{
    auto& s = detail_aes::state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.server) detail_circ::check(cufhe_server_destroy(s.server), "CleanUpAES");
    s.server = nullptr;
}

// plain = AES-128 decryption of cipher under the expanded key (FIPS-197 5.3).
inline void AESDec(std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>& plain,
                   const std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>& cipher,
                   const std::array<std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>, 11>& expandedkey,
                   Stream st = Stream(0))
{
    (void)st;
    InitializeAES();
    auto& s = detail_aes::state();
    std::vector<uint32_t> in, out;
    in.reserve(static_cast<std::size_t>(s.circuit.desc.n_inputs) * (TFHEpp::lvl1param::n + 1));
    for (const auto& c : cipher) detail_circ::append(in, c);
    for (const auto& rk : expandedkey)
        for (const auto& c : rk) detail_circ::append(in, c);
    detail_circ::append_constants(in);
    detail_circ::run(s.server, s.circuit, in, out, "AESDec");
    for (int i = 0; i < 128; i++)
        std::copy(out.begin() + static_cast<std::size_t>(i) * (TFHEpp::lvl1param::n + 1),
                  out.begin() + static_cast<std::size_t>(i + 1) * (TFHEpp::lvl1param::n + 1),
                  plain[i].begin());
}
inline void AESDec(std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>& plain,
                   const std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>& cipher,
                   const std::array<std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 128>, 11>& expandedkey,
                   const TFHEpp::EvalKey&, Stream st)
{
    AESDec(plain, cipher, expandedkey, st);
}

// FIPS-197 5.2 key expansion on the key owner's side: 11 round keys.
inline std::array<std::array<uint8_t, 16>, 11> AESKeyExpansion(const std::array<uint8_t, 16>& key)
{
    auto mul = [](int a, int b) {
        int r = 0;
        for (int i = 0; i < 8; i++) {
            if (b & 1) r ^= a;
            a = ((a << 1) ^ ((a & 0x80) ? 0x1B : 0)) & 0xFF;
            b >>= 1;
        }
        return r;
    };
    std::array<uint8_t, 256> sbox{};
    for (int x = 0; x < 256; x++) {
        int g = 0;
        for (int y = 1; y < 256 && x; y++)
            if (mul(x, y) == 1) g = y;
        int s = 0x63;
        for (int i = 0; i < 8; i++)
            s ^= (((g >> i) ^ (g >> ((i + 4) % 8)) ^ (g >> ((i + 5) % 8)) ^ (g >> ((i + 6) % 8)) ^
                   (g >> ((i + 7) % 8))) & 1) << i;
        sbox[x] = static_cast<uint8_t>(s);
    }
    std::array<std::array<uint8_t, 4>, 44> w{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) w[i][j] = key[4 * i + j];
    uint8_t rcon = 1;
    for (int i = 4; i < 44; i++) {
        std::array<uint8_t, 4> t = w[i - 1];
        if (i % 4 == 0) {
            t = {sbox[t[1]], sbox[t[2]], sbox[t[3]], sbox[t[0]]};
            t[0] ^= rcon;
            rcon = static_cast<uint8_t>(mul(rcon, 2));
        }
        for (int j = 0; j < 4; j++) w[i][j] = w[i - 4][j] ^ t[j];
    }
    std::array<std::array<uint8_t, 16>, 11> rk{};
    for (int r = 0; r < 11; r++)
        for (int i = 0; i < 16; i++) rk[r][i] = w[4 * r + i / 4][i % 4];
    return rk;
}

}  // namespace cufhe
