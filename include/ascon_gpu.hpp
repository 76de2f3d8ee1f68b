// C++ drop-in for cuFHEpp's ASCON-XOF (reference include/ascon_gpu.cuh:21-43,
// src/ascon_gpu.cu:1040-1282): ASCONXOF and its incremental form
// ASCONXOFInitialize / ASCONXOFAbsorb / ASCONXOFSqueeze over an encrypted
// state, with the same names and argument meaning over the C ABI of
// libcufhe_amd.so. Each operation of a given length is a circuit of lvl1
// gates (include/cufhe_circuits.h), built once and run by the library's
// serving loop.
//
// Differences a caller can see:
//  * lvl1 TLWEs (the engine's gates) where the reference uses TLWEs of its
//    circuit-bootstrapping target level; its template parameters and the
//    EvalKey / SecretKey arguments are accepted where it has them and not
//    read: the gate keys loaded by cufhe::Initialize(ek) are all it uses;
//  * ASCONState is 320 lvl1 TLWEs, bit k of word x_i at 64 i + k;
//  * input and output are any contiguous containers of
//    TFHEpp::TLWE<lvl1param> (std::span, std::vector, std::array), one TLWE
//    per bit, bit b of byte i at 8 i + b; sizes multiples of 8 as in the
//    reference (src/ascon_gpu.cu:1215-1216);
//  * the calls run on the serving loop's own stream and return when their
//    output is written; the Stream argument is not used. Calls are
//    serialised.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "aes_gpu.hpp"

namespace cufhe {

using ASCONState = std::array<TFHEpp::TLWE<TFHEpp::lvl1param>, 320>;

namespace detail_ascon {
using TLWE1 = TFHEpp::TLWE<TFHEpp::lvl1param>;
constexpr std::size_t kWords = TFHEpp::lvl1param::n + 1;

struct State {
    std::mutex mu;
    cufhe_server* server = nullptr;
    std::size_t pool_bytes = 0;
    std::map<std::tuple<int, int, int>, std::unique_ptr<detail_circ::Circuit>> circuits;
};
inline State& state()
{
    static State s;
    return s;
}

inline void check_bits(std::size_t bits, bool allow_zero, const char* what)
{
    if (bits % 8 || (!allow_zero && bits == 0) || bits / 8 > CUFHE_ASCON_MAX_BYTES) {
        std::fprintf(stderr, "cufhe: %s: sizes must be whole bytes (at most %d)\n", what,
                     CUFHE_ASCON_MAX_BYTES);
        std::exit(-1);
    }
}

// Runs the circuit (kind, a, b) on the TLWEs `in` (the constants appended
// here) and returns its outputs. 0: XOF(a -> b bytes), 1: absorb a bytes,
// 2: squeeze b bytes. Caller holds s.mu.
inline std::vector<uint32_t> run(int kind, int a, int b, std::vector<uint32_t> in)
{
    auto& s = state();
    auto& slot = s.circuits[{kind, a, b}];
    if (!slot) {
        slot = std::make_unique<detail_circ::Circuit>();
        slot->build(
            [&](cufhe_circuit_sizes* sz, cufhe_gate_desc* g, int* le, uint32_t* o) {
                return kind == 0   ? cufhe_circuit_ascon_xof(a, b, sz, g, le, o)
                       : kind == 1 ? cufhe_circuit_ascon_absorb(a, sz, g, le, o)
                                   : cufhe_circuit_ascon_squeeze(b, sz, g, le, o);
            },
            "ASCON circuit");
    }
    detail_circ::append_constants(in);
    std::vector<uint32_t> out;
    if (slot->desc.n_gates == 0) {
        // every output is an input wire (a squeeze of one block)
        for (uint32_t w : slot->outputs)
            out.insert(out.end(), in.begin() + w * kWords, in.begin() + (w + 1) * kWords);
        return out;
    }
    if (!s.server || s.pool_bytes < slot->wire_bytes()) {
        // a server whose pool holds this circuit (the previous one has no
        // request left: calls are serialised)
        if (s.server) detail_circ::check(cufhe_server_destroy(s.server), "ASCON");
        s.server = nullptr;
        s.pool_bytes = slot->wire_bytes();
        detail_circ::check(cufhe_server_create(s.pool_bytes, 16384, 0, &s.server), "ASCON");
    }
    detail_circ::run(s.server, *slot, in, out, "ASCON");
    return out;
}

inline void append(std::vector<uint32_t>& v, const TLWE1* c, std::size_t n)
{
    for (std::size_t i = 0; i < n; i++) detail_circ::append(v, c[i]);
}
inline void scatter(const std::vector<uint32_t>& v, std::size_t first, TLWE1* c, std::size_t n)
{
    for (std::size_t i = 0; i < n; i++)
        std::copy(v.begin() + (first + i) * kWords, v.begin() + (first + i + 1) * kWords, c[i].begin());
}
}  // namespace detail_ascon

inline void InitializeASCON() {}
inline void InitializeASCON(const TFHEpp::EvalKey&, const TFHEpp::SecretKey&) {}

inline void CleanUpASCON()
{
    auto& s = detail_ascon::state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.server) detail_circ::check(cufhe_server_destroy(s.server), "CleanUpASCON");
    s.server = nullptr;
    s.pool_bytes = 0;
    s.circuits.clear();
}

// output (8 out_bytes TLWEs) = Ascon-Xof of input (8 in_bytes TLWEs).
inline void ASCONXOF(TFHEpp::TLWE<TFHEpp::lvl1param>* output, std::size_t output_bits,
                     const TFHEpp::TLWE<TFHEpp::lvl1param>* input, std::size_t input_bits,
                     Stream st = Stream(0))
{
    (void)st;
    detail_ascon::check_bits(input_bits, true, "ASCONXOF");
    detail_ascon::check_bits(output_bits, false, "ASCONXOF");
    auto& s = detail_ascon::state();
    std::lock_guard<std::mutex> lk(s.mu);
    std::vector<uint32_t> in;
    detail_ascon::append(in, input, input_bits);
    const auto out = detail_ascon::run(0, static_cast<int>(input_bits / 8),
                                       static_cast<int>(output_bits / 8), std::move(in));
    detail_ascon::scatter(out, 0, output, output_bits);
}

// The public initial state p^12(IV, 0, 0, 0, 0) as noiseless encryptions
// (src/ascon_gpu.cu:1040-1048).
inline void ASCONXOFInitialize(ASCONState& state, Stream st = Stream(0))
{
    (void)st;
    uint64_t x[5];
    detail_circ::check(cufhe_ascon_xof_initial_state(x), "ASCONXOFInitialize");
    using P = TFHEpp::lvl1param;
    for (int i = 0; i < 320; i++) {
        state[i].fill(0);
        state[i][P::n] = ((x[i / 64] >> (i % 64)) & 1) ? static_cast<uint32_t>(P::μ)
                                                       : static_cast<uint32_t>(0u - P::μ);
    }
}

// Absorbs the whole message with its padding (src/ascon_gpu.cu:1051-1077).
inline void ASCONXOFAbsorb(ASCONState& state, const TFHEpp::TLWE<TFHEpp::lvl1param>* input,
                           std::size_t input_bits, Stream st = Stream(0))
{
    (void)st;
    detail_ascon::check_bits(input_bits, true, "ASCONXOFAbsorb");
    auto& s = detail_ascon::state();
    std::lock_guard<std::mutex> lk(s.mu);
    std::vector<uint32_t> in;
    detail_ascon::append(in, state.data(), 320);
    detail_ascon::append(in, input, input_bits);
    const auto out = detail_ascon::run(1, static_cast<int>(input_bits / 8), 0, std::move(in));
    detail_ascon::scatter(out, 0, state.data(), 320);
}

// Writes output_bits / 8 bytes; p^12 between 8-byte blocks, none after the
// last (src/ascon_gpu.cu:1080-1104).
inline void ASCONXOFSqueeze(ASCONState& state, TFHEpp::TLWE<TFHEpp::lvl1param>* output,
                            std::size_t output_bits, Stream st = Stream(0))
{
    (void)st;
    detail_ascon::check_bits(output_bits, false, "ASCONXOFSqueeze");
    auto& s = detail_ascon::state();
    std::lock_guard<std::mutex> lk(s.mu);
    std::vector<uint32_t> in;
    detail_ascon::append(in, state.data(), 320);
    const auto out = detail_ascon::run(2, 0, static_cast<int>(output_bits / 8), std::move(in));
    detail_ascon::scatter(out, 0, output, output_bits);
    detail_ascon::scatter(out, output_bits, state.data(), 320);
}

// Container forms (std::span, std::vector, std::array of lvl1 TLWEs), with
// and without the reference's EvalKey argument.
template <class Out, class In>
void ASCONXOF(Out& output, const In& input, Stream st = Stream(0))
{
    ASCONXOF(output.data(), output.size(), input.data(), input.size(), st);
}
template <class Out, class In>
void ASCONXOF(Out& output, const In& input, const TFHEpp::EvalKey&, Stream st)
{
    ASCONXOF(output.data(), output.size(), input.data(), input.size(), st);
}
inline void ASCONXOFInitialize(ASCONState& state, const TFHEpp::EvalKey&, Stream st)
{
    ASCONXOFInitialize(state, st);
}
template <class In>
void ASCONXOFAbsorb(ASCONState& state, const In& input, Stream st = Stream(0))
{
    ASCONXOFAbsorb(state, input.data(), input.size(), st);
}
template <class In>
void ASCONXOFAbsorb(ASCONState& state, const In& input, const TFHEpp::EvalKey&, Stream st)
{
    ASCONXOFAbsorb(state, input.data(), input.size(), st);
}
template <class Out>
void ASCONXOFSqueeze(ASCONState& state, Out& output, Stream st = Stream(0))
{
    ASCONXOFSqueeze(state, output.data(), output.size(), st);
}
template <class Out>
void ASCONXOFSqueeze(ASCONState& state, Out& output, const TFHEpp::EvalKey&, Stream st)
{
    ASCONXOFSqueeze(state, output.data(), output.size(), st);
}

}  // namespace cufhe
