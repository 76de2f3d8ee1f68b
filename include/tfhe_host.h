/*
 * Host-side TFHE layer: seeded key generation, encryption and decryption in
 * the memory layouts the gate-bootstrapping path consumes.
 *
 * The reference takes all of this from TFHEpp (an un-vendored submodule,
 * reference .gitmodules:1-3): SecretKey, EvalKey::emplacebk/emplaceiksk,
 * tlweSymEncrypt/Decrypt (used at reference test/test_util.h:25-27, 77-78 and
 * src/cufhe_gates_gpu.cu:42-61). These entry points are this project's own
 * replacement, so keys and ciphertexts are reproducible from a seed.
 *
 * Levels: 0 -> lvl0 (n=636, uint32), 1 -> lvl1 (N=1024, uint32),
 *         2 -> lvl2 (N=2048, uint64).
 * All functions return 0 on success and a negative CUFHE_ERR_* code on bad
 * arguments. None of them touches a GPU.
 */
#ifndef TFHE_HOST_H
#define TFHE_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Secret keys. lvl0: 636 words, lvl1: 1024 words, lvl2: 2048 words (u64).
 * block_binary != 0 draws the lvl0 key in blocks of 4 with at most one bit
 * set per block and makes the lvl1 key start with the lvl0 key. */
int tfhe_secret_key_gen(uint64_t seed, int block_binary, uint32_t* key_lvl0,
                        uint32_t* key_lvl1, uint64_t* key_lvl2);

/* Bootstrapping key in the polynomial domain (TFHEpp BootstrappingKey
 * layout): target_level 1 -> uint32 words, 2 -> uint64 words.
 * key_bundle != 0 writes n/2 groups of 3 TRGSWs:
 *   [3i+0] = TRGSW(s_{2i} s_{2i+1}), [3i+1] = TRGSW(s_{2i}(1-s_{2i+1})),
 *   [3i+2] = TRGSW((1-s_{2i}) s_{2i+1}).
 * Otherwise n TRGSWs, [i] = TRGSW(s_i). Words written are returned through
 * *words_out when it is non-NULL. */
int tfhe_bk_gen(int target_level, int key_bundle, uint64_t seed,
                const uint32_t* key_lvl0, const void* key_target, void* bk_out,
                size_t bk_capacity_words, size_t* words_out);

/* Key-switching key, layout [j][digit][|v|-1][0..n0] of uint32.
 * domain_level 1 (lvl10) or 2 (lvl20). subset != 0 only for level 1 with a
 * subset key (lvl1 key starts with the lvl0 key): rows exist only for
 * j >= n0, stored from row 0 (tfhe_ksk_subset_words() words); the subset
 * key switch copies coefficients j < n0 instead. */
int tfhe_ksk_gen(int domain_level, uint64_t seed, const void* key_domain,
                 const uint32_t* key_lvl0, int subset, uint32_t* ksk_out,
                 size_t ksk_capacity_words);

/* Encrypt count bits as TLWE(+-mu) at the level's fresh noise. Ciphertext i
 * is drawn from the stream (seed, i) so any subset can be regenerated.
 * Output stride is the TLWE size of the level (n+1 words). */
int tfhe_tlwe_encrypt_bits(int level, uint64_t seed, const void* key,
                           const uint8_t* bits, size_t count, void* ct_out);

/* One TLWE encryption of the torus word m (uint32 for levels 0/1, uint64
 * for level 2) with noise alpha, drawn from the stream (seed, index). */
int tfhe_tlwe_encrypt(int level, uint64_t seed, uint64_t index, const void* key, uint64_t m,
                      double alpha, void* ct_out);

/* One lvl1 TRLWE encryption (a then b, 2 x 1024 words) of the Torus32
 * polynomial m under the lvl1 key (TFHEpp trlweSymEncrypt). */
int tfhe_trlwe_encrypt(uint64_t seed, uint64_t index, const uint32_t* key_lvl1, const uint32_t* m,
                       double alpha, uint32_t* ct_out);

/* One lvl1 TRGSW encryption of the integer m (TFHEpp trgswSymEncrypt: 4
 * TRLWE encryptions of zero, m times the gadget added to row r's component
 * r / l), 4 x 2 x 1024 words in the bootstrapping key's TRGSW layout. */
int tfhe_trgsw_encrypt(uint64_t seed, uint64_t index, const uint32_t* key_lvl1, int m,
                       uint32_t* ct_out);

/* One TRLWE encryption (a then b, 2 x N words of the level's torus type)
 * of the polynomial m under the level's key (level 1: uint32, N = 1024;
 * level 2: uint64, N = 2048), noise alpha, stream (seed, index). */
int tfhe_trlwe_encrypt_level(int level, uint64_t seed, uint64_t index, const void* key,
                             const void* m, double alpha, void* ct_out);

/* Phase b - a*s of count TRLWEs (N torus words each) at level 1 or 2. */
int tfhe_trlwe_phase(int level, const void* key, const void* ct, size_t count, void* phase_out);

/* Annihilation key of level 1 or 2 (reference src/annihilate_gpu.cu:605-628
 * AnnihilateKeyPolynomialGen): for bit < nbit and d = 2^(bit+1) + 1 the half
 * TRGSW (l TRLWE rows, TFHEpp halftrgswSymEncrypt) of the automorphism
 * tau_d(s) of the level's key s, noise alpha (the level's alpha for TFHEpp's
 * keys). Layout [bit][row][a, b][N] words of the level's torus type:
 * tfhe_annihilate_key_words(level) words. */
size_t tfhe_annihilate_key_words(int level);
int tfhe_annihilate_key_gen(int level, uint64_t seed, const void* key, double alpha,
                            void* ahk_out, size_t capacity_words);

/* Phase of count TLWEs as a torus fraction in [-1/2, 1/2). */
int tfhe_tlwe_phase(int level, const void* key, const void* ct, size_t count,
                    double* phase_out);

/* Decrypt count TLWEs: bit = phase > 0. */
int tfhe_tlwe_decrypt_bits(int level, const void* key, const void* ct,
                           size_t count, uint8_t* bits_out);

/* ---- key files (no reference counterpart: TFHEpp serialises its EvalKey
 * with cereal, reference src/CMakeLists.txt:34; the un-vendored TFHEpp's
 * archive layout is not available, so this is a documented container of the
 * same TFHEpp-layout arrays). File: "CUFHEKY1", uint32 flags, uint32 record
 * count, then per record uint32 tag, uint32 element bytes (4 or 8), uint64
 * element count and the elements, little-endian. Tags: 1 bklvl01 (u32),
 * 2 bklvl02 (u64), 3 iksklvl10 (u32, full or subset), 4 iksklvl20 (u32),
 * 5 lvl0 secret key (u32), 6 lvl1 secret key (u32), 7 lvl2 secret key (u64).
 * Flags: bit 0 key bundle, bit 1 block-binary, bit 2 subset iksklvl10,
 * bit 3 NTT backend. Arrays absent from a file are passed as null / 0. */
#define TFHE_KEYFILE_BKLVL01 1
#define TFHE_KEYFILE_BKLVL02 2
#define TFHE_KEYFILE_IKSKLVL10 3
#define TFHE_KEYFILE_IKSKLVL20 4
#define TFHE_KEYFILE_SKLVL0 5
#define TFHE_KEYFILE_SKLVL1 6
#define TFHE_KEYFILE_SKLVL2 7
#define TFHE_KEYFILE_RECORDS 7
/* Write the arrays ptrs[t-1] of counts[t-1] elements (t = tag; null or 0 =
 * absent). Sizes are checked against the layouts: bootstrapping keys
 * tfhe_bk_words(level, flags & 1), key-switching keys tfhe_ksk_words (or
 * tfhe_ksk_subset_words with flags bit 2), secret keys n of their level.
 * Returns 0, or -1 (bad arguments / sizes) or -2 (I/O). */
int tfhe_keyfile_save(const char* path, uint32_t flags, const void* const* ptrs,
                      const size_t* counts);
/* Read a key file's flags and per-tag element counts (0 = absent). */
int tfhe_keyfile_info(const char* path, uint32_t* flags, size_t* counts);
/* Read the arrays into ptrs[t-1] (null: skip), each sized as reported by
 * tfhe_keyfile_info. Returns 0, -1 (format) or -2 (I/O). */
int tfhe_keyfile_load(const char* path, void* const* ptrs);

/* Words of the structures above (for sizing buffers). */
size_t tfhe_tlwe_words(int level);
size_t tfhe_bk_words(int target_level, int key_bundle);
size_t tfhe_ksk_words(int domain_level);
size_t tfhe_ksk_subset_words(void);

#ifdef __cplusplus
}
#endif

#endif /* TFHE_HOST_H */
