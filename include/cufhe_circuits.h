/*
 * Built-in Boolean circuits of lvl1 gates for the serving loop
 * (cufhe_server_*, include/cufhe_amd.h): AES-128 decryption and ASCON-XOF,
 * the functions of the reference's AESDec (include/aes_gpu.cuh:27-32,
 * src/aes_gpu.cu:904-1118) and ASCONXOF (include/ascon_gpu.cuh:40-43,
 * src/ascon_gpu.cu:1206-1282), which evaluate them with circuit
 * bootstrapping instead. Host code only: no GPU is touched.
 *
 * A circuit comes back in caller buffers: call with gates, layer_end and
 * outputs all NULL to get the sizes, then with buffers at least that large
 * (sizes holds their capacities on entry, the sizes used on return). The
 * arrays fill a cufhe_circuit directly: gate g writes wire n_inputs + g,
 * layers in order.
 *
 * Wire layouts (bit b of byte i is wire 8 i + b, least significant bit
 * first; outputs in the same order):
 *   AES-128 decryption: [0, 128) the ciphertext block, [128 + 128 r,
 *     256 + 128 r) round key r (r = 0 .. 10, the FIPS-197 5.2 expansion),
 *     then an encryption of 0 and one of 1; 128 outputs, the plaintext.
 *   ASCON-XOF of an in_bytes message into out_bytes bytes (Ascon v1.2
 *     Ascon-Xof): [0, 8 in_bytes) the message, then an encryption of 0 and
 *     one of 1; 8 out_bytes outputs.
 */
#ifndef CUFHE_CIRCUITS_H
#define CUFHE_CIRCUITS_H

#include <stdint.h>

#include "cufhe_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CUFHE_AES128_INPUTS (128 + 11 * 128 + 2)
/* An ASCON circuit holds every intermediate wire: about 20,000 gates (82 MB
 * of lvl1 wires on the serving loop) per 8-byte block absorbed or squeezed
 * beyond the first, so lengths are capped (1024 bytes: ~10 GB). */
#define CUFHE_ASCON_MAX_BYTES 1024

typedef struct {
    int n_inputs;
    int n_gates;
    int n_layers;
    int n_outputs;
} cufhe_circuit_sizes;

int cufhe_circuit_aes128_decrypt(cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                                 int* layer_end, uint32_t* outputs);
/* 0 <= in_bytes <= CUFHE_ASCON_MAX_BYTES, 0 < out_bytes <= CUFHE_ASCON_MAX_BYTES. */
int cufhe_circuit_ascon_xof(int in_bytes, int out_bytes, cufhe_circuit_sizes* sizes,
                            cufhe_gate_desc* gates, int* layer_end, uint32_t* outputs);
/* The XOF's two phases over an encrypted state (reference
 * src/ascon_gpu.cu:1051-1104 DeviceASCONXOFAbsorb / Squeeze). State wire
 * 64 i + k is bit k of word x_i (bit k = 2^k).
 *   absorb:  inputs [0, 320) the state, [320, 320 + 8 in_bytes) the message,
 *            then encryptions of 0 and 1; outputs the 320 state bits after
 *            the absorb (padding and the final p^12 included).
 *   squeeze: inputs [0, 320) the state, then encryptions of 0 and 1; outputs
 *            the 8 out_bytes output bits, then the 320 state bits after the
 *            squeeze (p^12 between 8-byte blocks, none after the last; with
 *            out_bytes <= 8 there are no gates and the outputs are inputs). */
/* words[5] = the public XOF state after p^12 of (IV, 0, 0, 0, 0) (reference
 * src/ascon_gpu.cu:1040-1048 DeviceASCONXOFInitialize). */
int cufhe_ascon_xof_initial_state(uint64_t* words);
int cufhe_circuit_ascon_absorb(int in_bytes, cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                               int* layer_end, uint32_t* outputs);
int cufhe_circuit_ascon_squeeze(int out_bytes, cufhe_circuit_sizes* sizes, cufhe_gate_desc* gates,
                                int* layer_end, uint32_t* outputs);

#ifdef __cplusplus
}
#endif

#endif
