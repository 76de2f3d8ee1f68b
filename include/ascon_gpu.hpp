// C++ drop-in for cuFHEpp's ASCON-XOF (reference include/ascon_gpu.cuh:21-43,
// src/ascon_gpu.cu:1206-1282 ASCONXOF): the same name and argument meaning
// over the C ABI of libcufhe_amd.so. The XOF of a message length and an
// output length is a circuit of lvl1 gates (cufhe_circuit_ascon_xof,
// include/cufhe_circuits.h), built once per length pair and run by the
// library's serving loop.
//
// Differences a caller can see:
//  * lvl1 TLWEs (the engine's gates) where the reference uses TLWEs of its
//    circuit-bootstrapping target level; its template parameters and the
//    EvalKey / SecretKey arguments are accepted where it has them and not
//    read: the gate keys loaded by cufhe::Initialize(ek) are all it uses;
//  * the whole XOF of a message is one call (ASCONXOF); the incremental
//    ASCONXOFInitialize / Absorb / Squeeze with an encrypted state are not
//    provided;
//  * input and output are any contiguous containers of
//    TFHEpp::TLWE<lvl1param> (std::span, std::vector, std::array), one TLWE
//    per bit, bit b of byte i at 8 i + b; sizes multiples of 8 as in the
//    reference (src/ascon_gpu.cu:1215-1216);
//  * ASCONXOF runs on the serving loop's own stream and returns when the
//    output is written; the Stream argument is not used. Calls are
//    serialised.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <utility>

#include "aes_gpu.hpp"

namespace cufhe {

namespace detail_ascon {
struct State {
    std::mutex mu;
    cufhe_server* server = nullptr;
    std::size_t pool_bytes = 0;
    std::map<std::pair<int, int>, std::unique_ptr<detail_circ::Circuit>> circuits;
};
inline State& state()
{
    static State s;
    return s;
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
    if (input_bits % 8 || output_bits % 8 || output_bits == 0 ||
        input_bits / 8 > CUFHE_ASCON_MAX_BYTES || output_bits / 8 > CUFHE_ASCON_MAX_BYTES) {
        std::fprintf(stderr, "cufhe: ASCONXOF: sizes must be whole bytes (1 .. %d out, 0 .. %d in)\n",
                     CUFHE_ASCON_MAX_BYTES, CUFHE_ASCON_MAX_BYTES);
        std::exit(-1);
    }
    auto& s = detail_ascon::state();
    std::lock_guard<std::mutex> lk(s.mu);
    const int in_bytes = static_cast<int>(input_bits / 8), out_bytes = static_cast<int>(output_bits / 8);
    auto& slot = s.circuits[{in_bytes, out_bytes}];
    if (!slot) {
        slot = std::make_unique<detail_circ::Circuit>();
        slot->build(
            [&](cufhe_circuit_sizes* sz, cufhe_gate_desc* g, int* le, uint32_t* o) {
                return cufhe_circuit_ascon_xof(in_bytes, out_bytes, sz, g, le, o);
            },
            "cufhe_circuit_ascon_xof");
    }
    if (!s.server || s.pool_bytes < slot->wire_bytes()) {
        // a server whose pool holds this circuit (the previous one has no
        // request left: calls are serialised)
        if (s.server) detail_circ::check(cufhe_server_destroy(s.server), "ASCONXOF");
        s.server = nullptr;
        s.pool_bytes = slot->wire_bytes();
        detail_circ::check(cufhe_server_create(s.pool_bytes, 16384, 0, &s.server), "ASCONXOF");
    }
    std::vector<uint32_t> in, out;
    for (std::size_t i = 0; i < input_bits; i++) detail_circ::append(in, input[i]);
    detail_circ::append_constants(in);
    detail_circ::run(s.server, *slot, in, out, "ASCONXOF");
    for (std::size_t i = 0; i < output_bits; i++)
        std::copy(out.begin() + i * (TFHEpp::lvl1param::n + 1),
                  out.begin() + (i + 1) * (TFHEpp::lvl1param::n + 1), output[i].begin());
}

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

}  // namespace cufhe
