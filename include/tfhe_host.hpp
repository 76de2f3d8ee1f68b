// C++ host layer with TFHEpp's names (SecretKey, EvalKey, tlweSymEncrypt,
// tlweSymDecrypt) over the C ABI in tfhe_host.h. It stands in for the
// un-vendored TFHEpp submodule the reference builds against (reference
// .gitmodules:1-3); keys and ciphertexts use TFHEpp's layouts, so code written
// against TFHEpp's EvalKey/TLWE types compiles against this header.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <type_traits>
#include <vector>

#include "tfhe_host.h"
#include "tfhe_params.hpp"

namespace TFHEpp {

namespace detail {
inline void check(int rc, const char* what)
{
    if (rc != 0) {
        std::fprintf(stderr, "TFHEpp host layer: %s failed (%d)\n", what, rc);
        std::exit(-1);
    }
}
template <class P>
constexpr int level()
{
    if constexpr (std::is_same_v<P, lvl0param>) return 0;
    if constexpr (std::is_same_v<P, lvl1param>) return 1;
    if constexpr (std::is_same_v<P, lvl2param>) return 2;
    return -1;
}
inline uint64_t& encrypt_counter()
{
    static uint64_t c = 0;
    return c;
}
inline uint64_t& encrypt_seed()
{
    static uint64_t s = std::random_device{}() | (static_cast<uint64_t>(std::random_device{}()) << 32);
    return s;
}
}  // namespace detail

struct SecretKey {
    struct {
        Key<lvl0param> lvl0;
        Key<lvl1param> lvl1;
        Key<lvl2param> lvl2;
        template <class P>
        const Key<P>& get() const
        {
            if constexpr (std::is_same_v<P, lvl0param>) return lvl0;
            else if constexpr (std::is_same_v<P, lvl1param>) return lvl1;
            else return lvl2;
        }
    } key;
    uint64_t seed;
    bool block_binary = false;  // lvl0 key in blocks of lvl0bbparam::ell, lvl1 key starts with it
    explicit SecretKey(uint64_t s, bool bb = false) : seed(s), block_binary(bb)
    {
        detail::check(tfhe_secret_key_gen(s, bb ? 1 : 0, key.lvl0.data(), key.lvl1.data(),
                                          key.lvl2.data()),
                      "tfhe_secret_key_gen");
    }
    SecretKey() : SecretKey(std::random_device{}() | (static_cast<uint64_t>(std::random_device{}()) << 32)) {}
    // Key files (tfhe_host.h): the three keys and the block-binary flag.
    bool save(const char* path) const
    {
        const void* ptrs[TFHE_KEYFILE_RECORDS] = {};
        size_t counts[TFHE_KEYFILE_RECORDS] = {};
        ptrs[TFHE_KEYFILE_SKLVL0 - 1] = key.lvl0.data();
        counts[TFHE_KEYFILE_SKLVL0 - 1] = key.lvl0.size();
        ptrs[TFHE_KEYFILE_SKLVL1 - 1] = key.lvl1.data();
        counts[TFHE_KEYFILE_SKLVL1 - 1] = key.lvl1.size();
        ptrs[TFHE_KEYFILE_SKLVL2 - 1] = key.lvl2.data();
        counts[TFHE_KEYFILE_SKLVL2 - 1] = key.lvl2.size();
        return tfhe_keyfile_save(path, block_binary ? 2u : 0u, ptrs, counts) == 0;
    }
    bool load(const char* path)
    {
        uint32_t fl = 0;
        size_t counts[TFHE_KEYFILE_RECORDS] = {};
        if (tfhe_keyfile_info(path, &fl, counts) != 0) return false;
        if (!counts[TFHE_KEYFILE_SKLVL0 - 1] || !counts[TFHE_KEYFILE_SKLVL1 - 1] ||
            !counts[TFHE_KEYFILE_SKLVL2 - 1])
            return false;
        void* ptrs[TFHE_KEYFILE_RECORDS] = {};
        ptrs[TFHE_KEYFILE_SKLVL0 - 1] = key.lvl0.data();
        ptrs[TFHE_KEYFILE_SKLVL1 - 1] = key.lvl1.data();
        ptrs[TFHE_KEYFILE_SKLVL2 - 1] = key.lvl2.data();
        if (tfhe_keyfile_load(path, ptrs) != 0) return false;
        block_binary = (fl & 2u) != 0;
        return true;
    }
};

// Bootstrapping key of brP in the polynomial domain (TRGSW after TRGSW:
// one per lvl0 key bit, or three per bit pair in the key-bundle layout) and
// key-switching key of ksP, as EvalKey holds them.
template <class brP>
using BootstrappingKey = std::vector<typename brP::targetP::T>;
template <class ksP>
using KeySwitchingKey = std::vector<uint32_t>;

// Bootstrapping keys (polynomial domain) and key-switching keys.
struct EvalKey {
    // Key files (tfhe_keyfile_save/_load, tfhe_host.h): the present arrays
    // and the flags (key bundle, block-binary, subset iksklvl10, NTT).
    // Returns false on an I/O or format error.
    bool save(const char* path) const
    {
        const void* ptrs[TFHE_KEYFILE_RECORDS] = {};
        size_t counts[TFHE_KEYFILE_RECORDS] = {};
        auto put = [&](int tag, const auto& v) {
            if (v) {
                ptrs[tag - 1] = v->data();
                counts[tag - 1] = v->size();
            }
        };
        put(TFHE_KEYFILE_BKLVL01, bklvl01);
        put(TFHE_KEYFILE_BKLVL02, bklvl02);
        put(TFHE_KEYFILE_IKSKLVL10, iksklvl10);
        put(TFHE_KEYFILE_IKSKLVL20, iksklvl20);
        return tfhe_keyfile_save(path, flags(), ptrs, counts) == 0;
    }
    bool load(const char* path)
    {
        uint32_t fl = 0;
        size_t counts[TFHE_KEYFILE_RECORDS] = {};
        if (tfhe_keyfile_info(path, &fl, counts) != 0) return false;
        void* ptrs[TFHE_KEYFILE_RECORDS] = {};
        auto take = [&](int tag, auto& v) {
            using V = typename std::remove_reference_t<decltype(v)>::element_type;
            if (counts[tag - 1]) {
                v = std::make_unique<V>(counts[tag - 1]);
                ptrs[tag - 1] = v->data();
            }
            else {
                v.reset();
            }
        };
        take(TFHE_KEYFILE_BKLVL01, bklvl01);
        take(TFHE_KEYFILE_BKLVL02, bklvl02);
        take(TFHE_KEYFILE_IKSKLVL10, iksklvl10);
        take(TFHE_KEYFILE_IKSKLVL20, iksklvl20);
        if (tfhe_keyfile_load(path, ptrs) != 0) return false;
        key_bundle = (fl & 1u) != 0;
        block_binary = (fl & 2u) != 0;
        subset_iksk = (fl & 4u) != 0;
        ntt = (fl & 8u) != 0;
        return true;
    }
    uint32_t flags() const
    {
        return (key_bundle ? 1u : 0u) | (block_binary ? 2u : 0u) | (subset_iksk ? 4u : 0u) |
               (ntt ? 8u : 0u);
    }
    // getbk<lvl01param / lvl02param>(), getiksk<lvl10param / lvl20param>()
    template <class brP>
    const BootstrappingKey<brP>& getbk() const
    {
        static_assert(std::is_same_v<brP, lvl01param> || std::is_same_v<brP, lvl02param>,
                      "bootstrapping keys: lvl01param, lvl02param");
        if constexpr (std::is_same_v<brP, lvl01param>)
            return *bklvl01;
        else
            return *bklvl02;
    }
    template <class ksP>
    const KeySwitchingKey<ksP>& getiksk() const
    {
        static_assert(std::is_same_v<ksP, lvl10param> || std::is_same_v<ksP, lvl20param>,
                      "key-switching keys: lvl10param, lvl20param");
        if constexpr (std::is_same_v<ksP, lvl10param>)
            return *iksklvl10;
        else
            return *iksklvl20;
    }

    bool key_bundle = true;    // 2-bit key bundle (reference default build)
    bool block_binary = false;  // block-binary rotation (NTT backend; reference USE_BLOCK_BINARY)
    bool subset_iksk = false;   // iksklvl10 holds the subset key (emplacesubiksk)
    bool ntt = false;           // NTT transform backend at both levels (reference USE_FFT=OFF)
    EvalKey() = default;
    // Keys for sk's configuration: a block-binary secret key selects the
    // block-binary rotation (one TRGSW per key bit, no key bundle).
    explicit EvalKey(const SecretKey& sk)
    {
        if (sk.block_binary) {
            key_bundle = false;
            block_binary = true;
        }
    }
    std::unique_ptr<std::vector<uint32_t>> bklvl01;
    std::unique_ptr<std::vector<uint64_t>> bklvl02;
    std::unique_ptr<std::vector<uint32_t>> iksklvl10;
    std::unique_ptr<std::vector<uint32_t>> iksklvl20;

    template <class P>
    void emplacebk(const SecretKey& sk)
    {
        if constexpr (std::is_same_v<P, lvl01param>) {
            bklvl01 = std::make_unique<std::vector<uint32_t>>(tfhe_bk_words(1, key_bundle));
            detail::check(tfhe_bk_gen(1, key_bundle, sk.seed ^ 0xB01, sk.key.lvl0.data(),
                                      sk.key.lvl1.data(), bklvl01->data(), bklvl01->size(), nullptr),
                          "tfhe_bk_gen");
        }
        else {
            static_assert(std::is_same_v<P, lvl02param>, "unsupported bootstrapping key");
            bklvl02 = std::make_unique<std::vector<uint64_t>>(tfhe_bk_words(2, key_bundle));
            detail::check(tfhe_bk_gen(2, key_bundle, sk.seed ^ 0xB02, sk.key.lvl0.data(),
                                      sk.key.lvl2.data(), bklvl02->data(), bklvl02->size(), nullptr),
                          "tfhe_bk_gen");
        }
    }
    // Subset key switching lvl1 -> lvl0 (reference emplacesubiksk,
    // test/test_block_binary_gpu.cc:155): rows only for lvl1 coefficients
    // beyond the lvl0 key the lvl1 key starts with.
    template <class P>
    void emplacesubiksk(const SecretKey& sk)
    {
        static_assert(std::is_same_v<P, lvl10param>, "subset key switching is lvl10 only");
        iksklvl10 = std::make_unique<std::vector<uint32_t>>(tfhe_ksk_subset_words());
        detail::check(tfhe_ksk_gen(1, sk.seed ^ 0x510, sk.key.lvl1.data(), sk.key.lvl0.data(), 1,
                                   iksklvl10->data(), iksklvl10->size()),
                      "tfhe_ksk_gen");
        subset_iksk = true;
    }
    template <class P>
    void emplaceiksk(const SecretKey& sk)
    {
        if constexpr (std::is_same_v<P, lvl10param>) {
            iksklvl10 = std::make_unique<std::vector<uint32_t>>(tfhe_ksk_words(1));
            detail::check(tfhe_ksk_gen(1, sk.seed ^ 0x510, sk.key.lvl1.data(), sk.key.lvl0.data(), 0,
                                       iksklvl10->data(), iksklvl10->size()),
                          "tfhe_ksk_gen");
        }
        else {
            static_assert(std::is_same_v<P, lvl20param>, "unsupported key-switching key");
            iksklvl20 = std::make_unique<std::vector<uint32_t>>(tfhe_ksk_words(2));
            detail::check(tfhe_ksk_gen(2, sk.seed ^ 0x520, sk.key.lvl2.data(), sk.key.lvl0.data(), 0,
                                       iksklvl20->data(), iksklvl20->size()),
                          "tfhe_ksk_gen");
        }
    }
};

template <class P>
TLWE<P> tlweSymEncrypt(const typename P::T p, const double alpha, const Key<P>& key)
{
    TLWE<P> c;
    detail::check(tfhe_tlwe_encrypt(detail::level<P>(), detail::encrypt_seed(),
                                    detail::encrypt_counter()++, key.data(), static_cast<uint64_t>(p),
                                    alpha, c.data()),
                  "tfhe_tlwe_encrypt");
    return c;
}

inline void trlweSymEncrypt(TRLWE<lvl1param>& c, const Polynomial<lvl1param>& p,
                            const Key<lvl1param>& key)
{
    std::array<uint32_t, 2 * lvl1param::n> flat;
    detail::check(tfhe_trlwe_encrypt(detail::encrypt_seed(), detail::encrypt_counter()++,
                                     key.data(), p.data(), lvl1param::α, flat.data()),
                  "tfhe_trlwe_encrypt");
    for (uint32_t i = 0; i < lvl1param::n; i++) {
        c[0][i] = flat[i];
        c[1][i] = flat[lvl1param::n + i];
    }
}
// TRLWE<P> is [a, b] of N words each, the flat layout of the C ABI.
template <class P>
void trlweSymEncrypt(TRLWE<P>& c, const Polynomial<P>& p, const Key<P>& key)
{
    static_assert(std::is_same_v<P, lvl1param> || std::is_same_v<P, lvl2param>,
                  "lvl1 or lvl2 TRLWE");
    detail::check(tfhe_trlwe_encrypt_level(detail::level<P>(), detail::encrypt_seed(),
                                           detail::encrypt_counter()++, key.data(), p.data(), P::α,
                                           c[0].data()),
                  "tfhe_trlwe_encrypt_level");
}

// TFHEpp trlwePhase<P>: b - a * s.
template <class P>
Polynomial<P> trlwePhase(const TRLWE<P>& c, const Key<P>& key)
{
    Polynomial<P> ph;
    detail::check(tfhe_trlwe_phase(detail::level<P>(), key.data(), c[0].data(), 1, ph.data()),
                  "tfhe_trlwe_phase");
    return ph;
}

// TFHEpp trgswSymEncrypt<lvl1param>(c, m, key): TRGSW of the integer m.
inline void trgswSymEncrypt(TRGSW<lvl1param>& c, int m, const Key<lvl1param>& key)
{
    std::array<uint32_t, kTRGSWWords<lvl1param>> flat;
    detail::check(tfhe_trgsw_encrypt(detail::encrypt_seed(), detail::encrypt_counter()++,
                                     key.data(), m, flat.data()),
                  "tfhe_trgsw_encrypt");
    for (uint32_t r = 0; r < BootstrappingTRGSWRows<lvl1param>; r++)
        for (uint32_t j = 0; j < 2; j++)
            for (uint32_t i = 0; i < lvl1param::n; i++)
                c[r][j][i] = flat[(static_cast<size_t>(r) * 2 + j) * lvl1param::n + i];
}

// TFHEpp's in-place form: tlweSymEncrypt<P>(c, p, key) with the level's noise.
template <class P>
void tlweSymEncrypt(TLWE<P>& c, const typename P::T p, const Key<P>& key)
{
    c = tlweSymEncrypt<P>(p, P::α, key);
}

template <class P>
bool tlweSymDecrypt(const TLWE<P>& c, const Key<P>& key)
{
    uint8_t bit = 0;
    detail::check(tfhe_tlwe_decrypt_bits(detail::level<P>(), key.data(), c.data(), 1, &bit),
                  "tfhe_tlwe_decrypt_bits");
    return bit != 0;
}

}  // namespace TFHEpp
