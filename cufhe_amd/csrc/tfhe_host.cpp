// Host-side TFHE layer: seeded keygen, encryption, decryption.
//
// Stands in for the TFHEpp calls the reference makes (SecretKey, EvalKey
// emplacebk/emplaceiksk, tlweSymEncrypt/Decrypt; reference
// src/cufhe_gates_gpu.cu:42-61, test/test_util.h:25-27). Layouts follow
// include/tfhe_params.hpp. The random source is ChaCha20 (RFC 8439 block
// function) keyed by the seed; every ciphertext / TRGSW has its own stream so
// results do not depend on thread count.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/cufhe_amd.h"
#include "../../include/tfhe_host.h"
#include "../../include/tfhe_params.hpp"

namespace {

using namespace TFHEpp;

inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

class ChaCha20 {
   public:
    ChaCha20(uint64_t seed, uint64_t stream)
    {
        // Expand the 64-bit seed into a 256-bit key with splitmix64.
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9E3779B97F4A7C15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
            x ^= x >> 31;
            key_[2 * i] = static_cast<uint32_t>(x);
            key_[2 * i + 1] = static_cast<uint32_t>(x >> 32);
        }
        nonce_[0] = static_cast<uint32_t>(stream);
        nonce_[1] = static_cast<uint32_t>(stream >> 32);
        nonce_[2] = 0x54464845u;  // "TFHE"
        counter_ = 0;
        pos_ = 16;
    }
    uint32_t next32()
    {
        if (pos_ == 16) refill();
        return buf_[pos_++];
    }
    uint64_t next64()
    {
        const uint64_t lo = next32();
        return lo | (static_cast<uint64_t>(next32()) << 32);
    }
    // Uniform in (0, 1] with 53 random bits.
    double uniform_open0()
    {
        return (static_cast<double>(next64() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    }
    double gaussian()
    {
        if (have_spare_) {
            have_spare_ = false;
            return spare_;
        }
        const double u1 = uniform_open0();
        const double u2 = uniform_open0();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = 2.0 * M_PI * u2;
        spare_ = r * std::sin(th);
        have_spare_ = true;
        return r * std::cos(th);
    }

   private:
    static void qr(uint32_t* s, int a, int b, int c, int d)
    {
        s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl32(s[d], 16);
        s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl32(s[b], 12);
        s[a] += s[b]; s[d] ^= s[a]; s[d] = rotl32(s[d], 8);
        s[c] += s[d]; s[b] ^= s[c]; s[b] = rotl32(s[b], 7);
    }
    void refill()
    {
        uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; i++) s[4 + i] = key_[i];
        s[12] = counter_++;
        s[13] = nonce_[0];
        s[14] = nonce_[1];
        s[15] = nonce_[2];
        uint32_t w[16];
        std::memcpy(w, s, sizeof(w));
        for (int i = 0; i < 10; i++) {
            qr(w, 0, 4, 8, 12); qr(w, 1, 5, 9, 13); qr(w, 2, 6, 10, 14); qr(w, 3, 7, 11, 15);
            qr(w, 0, 5, 10, 15); qr(w, 1, 6, 11, 12); qr(w, 2, 7, 8, 13); qr(w, 3, 4, 9, 14);
        }
        for (int i = 0; i < 16; i++) buf_[i] = w[i] + s[i];
        pos_ = 0;
    }
    uint32_t key_[8];
    uint32_t nonce_[3];
    uint32_t counter_;
    uint32_t buf_[16];
    int pos_;
    bool have_spare_ = false;
    double spare_ = 0.0;
};

// Stream tags keep the domains of the key material apart.
constexpr uint64_t kTagSecret = 1ULL << 56;
constexpr uint64_t kTagBK = 2ULL << 56;
constexpr uint64_t kTagKSK = 3ULL << 56;
constexpr uint64_t kTagEnc = 4ULL << 56;
constexpr uint64_t kTagAHK = 5ULL << 56;

template <class T>
T torus_uniform(ChaCha20& rng)
{
    if constexpr (sizeof(T) == 4)
        return rng.next32();
    else
        return rng.next64();
}

// Rounded Gaussian torus error of standard deviation alpha (torus units).
template <class T>
T torus_gaussian(ChaCha20& rng, double alpha)
{
    constexpr double scale = sizeof(T) == 4 ? 4294967296.0 : 18446744073709551616.0;
    const double e = rng.gaussian() * alpha * scale;
    return static_cast<T>(static_cast<int64_t>(std::llround(e)));
}

// b += a * s in Z_{2^w}[X]/(X^N+1) for a binary key s.
template <class T>
void poly_mul_binary_acc(T* b, const T* a, const T* s, uint32_t N)
{
    for (uint32_t i = 0; i < N; i++) {
        if (s[i] == 0) continue;
        for (uint32_t j = 0; j < N; j++) {
            const uint32_t idx = i + j;
            if (idx < N)
                b[idx] += a[j];
            else
                b[idx - N] -= a[j];
        }
    }
}

// One TRLWE encryption of zero: rows of (a_0 .. a_{k-1}, b).
template <class P>
void trlwe_encrypt_zero(typename P::T* out, const typename P::T* key, double alpha,
                        ChaCha20& rng)
{
    using T = typename P::T;
    constexpr uint32_t N = P::n;
    for (uint32_t i = 0; i < N; i++) out[i] = torus_uniform<T>(rng);
    T* b = out + N;
    for (uint32_t i = 0; i < N; i++) b[i] = torus_gaussian<T>(rng, alpha);
    poly_mul_binary_acc<T>(b, out, key, N);
}

template <class P>
void trgsw_encrypt(typename P::T* out, int m, const typename P::T* key, ChaCha20& rng)
{
    using T = typename P::T;
    constexpr uint32_t N = P::n;
    constexpr uint32_t rows = BootstrappingTRGSWRows<P>;
    for (uint32_t r = 0; r < rows; r++) trlwe_encrypt_zero<P>(out + r * 2 * N, key, P::α, rng);
    if (m == 0) return;
    for (uint32_t r = 0; r < rows; r++) {
        const uint32_t j = r / P::l;
        const uint32_t digit = r % P::l;
        const T h = static_cast<T>(1) << (std::numeric_limits<T>::digits - (digit + 1) * P::Bgbit);
        out[r * 2 * N + j * N] += static_cast<T>(m) * h;
    }
}

// p(X) -> p(X^d) in Z_{2^w}[X]/(X^N+1), d odd (TFHEpp Automorphism).
template <class T>
void automorphism(const T* in, uint32_t N, uint32_t d, T* out)
{
    for (uint32_t i = 0; i < N; i++) {
        const uint32_t idx = i * d;
        out[idx & (N - 1)] = (idx & N) ? static_cast<T>(0 - in[i]) : in[i];
    }
}

// Half TRGSW of the polynomial p (TFHEpp halftrgswSymEncrypt): l TRLWE
// encryptions of zero, row q's b component plus p * h_q (h_q = 2^(w-(q+1)Bgbit)).
template <class P>
void halftrgsw_encrypt(typename P::T* out, const typename P::T* p, const typename P::T* key,
                       double alpha, ChaCha20& rng)
{
    using T = typename P::T;
    constexpr uint32_t N = P::n;
    for (uint32_t q = 0; q < P::l; q++) {
        T* row = out + static_cast<size_t>(q) * 2 * N;
        trlwe_encrypt_zero<P>(row, key, alpha, rng);
        const T h = static_cast<T>(1) << (std::numeric_limits<T>::digits - (q + 1) * P::Bgbit);
        for (uint32_t j = 0; j < N; j++) row[N + j] += p[j] * h;
    }
}

// Annihilation key (reference src/annihilate_gpu.cu:605-628
// AnnihilateKeyPolynomialGen): for bit < nbit and d = 2^(bit+1) + 1, the half
// TRGSW of tau_d(s) under s. Layout: [bit][row q][component a, b][N].
template <class P>
void annihilate_key_gen(uint64_t seed, const typename P::T* key, double alpha, typename P::T* out)
{
    using T = typename P::T;
    constexpr uint32_t N = P::n;
    constexpr size_t per_bit = static_cast<size_t>(P::l) * 2 * N;
    std::vector<std::thread> pool;
    for (uint32_t bit = 0; bit < P::nbit; bit++) {
        pool.emplace_back([=]() {
            std::vector<T> autokey(N);
            automorphism<T>(key, N, (2u << bit) + 1, autokey.data());
            ChaCha20 rng(seed, kTagAHK | bit);
            halftrgsw_encrypt<P>(out + bit * per_bit, autokey.data(), key, alpha, rng);
        });
    }
    for (auto& th : pool) th.join();
}

template <class brP>
int bk_gen(bool kb, uint64_t seed, const uint32_t* key0, const typename brP::targetP::T* keyt,
           typename brP::targetP::T* out, size_t capacity, size_t* words_out)
{
    using P = typename brP::targetP;
    constexpr size_t tw = kTRGSWWords<P>;
    const size_t count = kBKTRGSWs<brP>(kb);
    if (capacity < count * tw) return CUFHE_ERR_SIZE;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < hw; t++) {
        pool.emplace_back([=]() {
            for (size_t i = t; i < count; i += hw) {
                int m;
                if (kb) {
                    const size_t g = i / 3, which = i % 3;
                    const int s0 = static_cast<int>(key0[2 * g]);
                    const int s1 = static_cast<int>(key0[2 * g + 1]);
                    m = which == 0 ? s0 * s1 : which == 1 ? s0 * (1 - s1) : (1 - s0) * s1;
                }
                else {
                    m = static_cast<int>(key0[i]);
                }
                ChaCha20 rng(seed, kTagBK | i);
                trgsw_encrypt<P>(out + i * tw, m, keyt, rng);
            }
        });
    }
    for (auto& th : pool) th.join();
    if (words_out) *words_out = count * tw;
    return CUFHE_OK;
}

template <class P>
void tlwe_encrypt(typename P::T* out, typename P::T m, const typename P::T* key, double alpha,
                  ChaCha20& rng)
{
    using T = typename P::T;
    constexpr uint32_t n = P::k * P::n;
    T b = m + torus_gaussian<T>(rng, alpha);
    for (uint32_t i = 0; i < n; i++) {
        out[i] = torus_uniform<T>(rng);
        b += out[i] * key[i];
    }
    out[n] = b;
}

template <class ksP>
int ksk_gen(uint64_t seed, const typename ksP::domainP::T* keyd, const uint32_t* key0, bool subset,
            uint32_t* out, size_t capacity)
{
    using TP = typename ksP::targetP;
    constexpr uint32_t dn = ksP::domainP::k * ksP::domainP::n;
    constexpr uint32_t numbase = 1u << (ksP::basebit - 1);
    constexpr uint32_t tl = TP::k * TP::n + 1;
    const size_t words = (subset ? (dn - TP::n) : dn) * static_cast<size_t>(ksP::t) * numbase * tl;
    if (capacity < words) return CUFHE_ERR_SIZE;
    std::fill(out, out + words, 0u);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < hw; t++) {
        pool.emplace_back([=]() {
            for (uint32_t j = t; j < dn; j += hw) {
                if (subset && j < TP::n) continue;
                ChaCha20 rng(seed, kTagKSK | j);
                const size_t row = subset ? j - TP::n : j;
                for (uint32_t d = 0; d < ksP::t; d++)
                    for (uint32_t h = 0; h < numbase; h++) {
                        const uint32_t m = static_cast<uint32_t>(keyd[j]) * (h + 1) *
                                           (1u << (32 - (d + 1) * ksP::basebit));
                        uint32_t* dst = out + ((row * ksP::t + d) * numbase + h) * tl;
                        tlwe_encrypt<TP>(dst, m, key0, ksP::α, rng);
                    }
            }
        });
    }
    for (auto& th : pool) th.join();
    return CUFHE_OK;
}

template <class P>
void encrypt_bits(uint64_t seed, const typename P::T* key, const uint8_t* bits, size_t count,
                  typename P::T* out)
{
    using T = typename P::T;
    constexpr size_t tl = P::k * P::n + 1;
    for (size_t i = 0; i < count; i++) {
        ChaCha20 rng(seed, kTagEnc | i);
        const T m = bits[i] ? static_cast<T>(P::μ) : static_cast<T>(-P::μ);
        tlwe_encrypt<P>(out + i * tl, m, key, P::α, rng);
    }
}

template <class P>
void phase(const typename P::T* key, const typename P::T* ct, size_t count, double* out)
{
    using T = typename P::T;
    using ST = std::make_signed_t<T>;
    constexpr size_t n = P::k * P::n;
    constexpr double scale = sizeof(T) == 4 ? 4294967296.0 : 18446744073709551616.0;
    for (size_t c = 0; c < count; c++) {
        const T* x = ct + c * (n + 1);
        T p = x[n];
        for (size_t i = 0; i < n; i++) p -= x[i] * key[i];
        out[c] = static_cast<double>(static_cast<ST>(p)) / scale;
    }
}

}  // namespace

extern "C" {

int tfhe_secret_key_gen(uint64_t seed, int block_binary, uint32_t* key0, uint32_t* key1,
                        uint64_t* key2)
{
    if (!key0 || !key1 || !key2) return CUFHE_ERR_ARG;
    ChaCha20 rng(seed, kTagSecret);
    if (block_binary) {
        constexpr uint32_t ell = lvl0bbparam::ell;
        for (uint32_t blk = 0; blk < lvl0param::n / ell; blk++) {
            // at most one bit per block: uniform over ell + 1 patterns
            const uint32_t pick = rng.next32() % (ell + 1);
            for (uint32_t o = 0; o < ell; o++) key0[blk * ell + o] = (pick == o + 1) ? 1u : 0u;
        }
    }
    else {
        for (uint32_t i = 0; i < lvl0param::n; i++) key0[i] = rng.next32() & 1u;
    }
    for (uint32_t i = 0; i < lvl1param::n; i++) key1[i] = rng.next32() & 1u;
    if (block_binary)
        for (uint32_t i = 0; i < lvl0param::n; i++) key1[i] = key0[i];
    for (uint32_t i = 0; i < lvl2param::n; i++) key2[i] = rng.next32() & 1u;
    return CUFHE_OK;
}

int tfhe_bk_gen(int target_level, int key_bundle, uint64_t seed, const uint32_t* key0,
                const void* key_target, void* bk_out, size_t capacity, size_t* words_out)
{
    if (!key0 || !key_target || !bk_out) return CUFHE_ERR_ARG;
    if (target_level == 1)
        return bk_gen<lvl01param>(key_bundle != 0, seed, key0,
                                  static_cast<const uint32_t*>(key_target),
                                  static_cast<uint32_t*>(bk_out), capacity, words_out);
    if (target_level == 2)
        return bk_gen<lvl02param>(key_bundle != 0, seed, key0,
                                  static_cast<const uint64_t*>(key_target),
                                  static_cast<uint64_t*>(bk_out), capacity, words_out);
    return CUFHE_ERR_ARG;
}

int tfhe_ksk_gen(int domain_level, uint64_t seed, const void* key_domain, const uint32_t* key0,
                 int subset, uint32_t* ksk_out, size_t capacity)
{
    if (!key_domain || !key0 || !ksk_out) return CUFHE_ERR_ARG;
    if (domain_level == 1)
        return ksk_gen<lvl10param>(seed, static_cast<const uint32_t*>(key_domain), key0,
                                   subset != 0, ksk_out, capacity);
    if (domain_level == 2 && !subset)
        return ksk_gen<lvl20param>(seed, static_cast<const uint64_t*>(key_domain), key0, false,
                                   ksk_out, capacity);
    return CUFHE_ERR_ARG;
}

int tfhe_tlwe_encrypt_bits(int level, uint64_t seed, const void* key, const uint8_t* bits,
                           size_t count, void* ct)
{
    if (!key || (!bits && count) || (!ct && count)) return CUFHE_ERR_ARG;
    switch (level) {
        case 0:
            encrypt_bits<lvl0param>(seed, static_cast<const uint32_t*>(key), bits, count,
                                    static_cast<uint32_t*>(ct));
            return CUFHE_OK;
        case 1:
            encrypt_bits<lvl1param>(seed, static_cast<const uint32_t*>(key), bits, count,
                                    static_cast<uint32_t*>(ct));
            return CUFHE_OK;
        case 2:
            encrypt_bits<lvl2param>(seed, static_cast<const uint64_t*>(key), bits, count,
                                    static_cast<uint64_t*>(ct));
            return CUFHE_OK;
        default:
            return CUFHE_ERR_ARG;
    }
}

int tfhe_tlwe_encrypt(int level, uint64_t seed, uint64_t index, const void* key, uint64_t m,
                      double alpha, void* ct)
{
    if (!key || !ct) return CUFHE_ERR_ARG;
    ChaCha20 rng(seed, kTagEnc | index);
    switch (level) {
        case 0:
            tlwe_encrypt<lvl0param>(static_cast<uint32_t*>(ct), static_cast<uint32_t>(m),
                                    static_cast<const uint32_t*>(key), alpha, rng);
            return CUFHE_OK;
        case 1:
            tlwe_encrypt<lvl1param>(static_cast<uint32_t*>(ct), static_cast<uint32_t>(m),
                                    static_cast<const uint32_t*>(key), alpha, rng);
            return CUFHE_OK;
        case 2:
            tlwe_encrypt<lvl2param>(static_cast<uint64_t*>(ct), m,
                                    static_cast<const uint64_t*>(key), alpha, rng);
            return CUFHE_OK;
        default:
            return CUFHE_ERR_ARG;
    }
}

int tfhe_trlwe_encrypt(uint64_t seed, uint64_t index, const uint32_t* key, const uint32_t* m,
                       double alpha, uint32_t* ct)
{
    if (!key || !m || !ct) return CUFHE_ERR_ARG;
    ChaCha20 rng(seed, kTagEnc | index);
    trlwe_encrypt_zero<lvl1param>(ct, key, alpha, rng);
    for (uint32_t i = 0; i < lvl1param::n; i++) ct[lvl1param::n + i] += m[i];
    return CUFHE_OK;
}

int tfhe_trlwe_encrypt_level(int level, uint64_t seed, uint64_t index, const void* key,
                             const void* m, double alpha, void* ct)
{
    if (!key || !m || !ct) return CUFHE_ERR_ARG;
    auto run = [&](auto tag) {
        using P = decltype(tag);
        using T = typename P::T;
        ChaCha20 rng(seed, kTagEnc | index);
        T* c = static_cast<T*>(ct);
        trlwe_encrypt_zero<P>(c, static_cast<const T*>(key), alpha, rng);
        for (uint32_t i = 0; i < P::n; i++) c[P::n + i] += static_cast<const T*>(m)[i];
    };
    if (level == 1)
        run(lvl1param{});
    else if (level == 2)
        run(lvl2param{});
    else
        return CUFHE_ERR_ARG;
    return CUFHE_OK;
}

int tfhe_trlwe_phase(int level, const void* key, const void* ct, size_t count, void* out)
{
    if (!key || (count && (!ct || !out))) return CUFHE_ERR_ARG;
    auto run = [&](auto tag) {
        using P = decltype(tag);
        using T = typename P::T;
        constexpr uint32_t N = P::n;
        std::vector<T> as(N);
        for (size_t c = 0; c < count; c++) {
            const T* x = static_cast<const T*>(ct) + c * 2 * N;
            std::fill(as.begin(), as.end(), T(0));
            poly_mul_binary_acc<T>(as.data(), x, static_cast<const T*>(key), N);
            T* o = static_cast<T*>(out) + c * N;
            for (uint32_t i = 0; i < N; i++) o[i] = static_cast<T>(x[N + i] - as[i]);
        }
    };
    if (level == 1)
        run(lvl1param{});
    else if (level == 2)
        run(lvl2param{});
    else
        return CUFHE_ERR_ARG;
    return CUFHE_OK;
}

size_t tfhe_annihilate_key_words(int level)
{
    if (level == 1) return static_cast<size_t>(lvl1param::nbit) * lvl1param::l * 2 * lvl1param::n;
    if (level == 2) return static_cast<size_t>(lvl2param::nbit) * lvl2param::l * 2 * lvl2param::n;
    return 0;
}

int tfhe_annihilate_key_gen(int level, uint64_t seed, const void* key, double alpha, void* out,
                            size_t capacity_words)
{
    if (!key || !out || !(alpha >= 0.0)) return CUFHE_ERR_ARG;
    const size_t words = tfhe_annihilate_key_words(level);
    if (words == 0) return CUFHE_ERR_ARG;
    if (capacity_words < words) return CUFHE_ERR_SIZE;
    if (level == 1)
        annihilate_key_gen<lvl1param>(seed, static_cast<const uint32_t*>(key), alpha,
                                      static_cast<uint32_t*>(out));
    else
        annihilate_key_gen<lvl2param>(seed, static_cast<const uint64_t*>(key), alpha,
                                      static_cast<uint64_t*>(out));
    return CUFHE_OK;
}

int tfhe_trgsw_encrypt(uint64_t seed, uint64_t index, const uint32_t* key, int m, uint32_t* ct)
{
    if (!key || !ct) return CUFHE_ERR_ARG;
    ChaCha20 rng(seed, kTagEnc | index);
    trgsw_encrypt<lvl1param>(ct, m, key, rng);
    return CUFHE_OK;
}

int tfhe_tlwe_phase(int level, const void* key, const void* ct, size_t count, double* out)
{
    if (!key || (!ct && count) || (!out && count)) return CUFHE_ERR_ARG;
    switch (level) {
        case 0:
            phase<lvl0param>(static_cast<const uint32_t*>(key), static_cast<const uint32_t*>(ct),
                             count, out);
            return CUFHE_OK;
        case 1:
            phase<lvl1param>(static_cast<const uint32_t*>(key), static_cast<const uint32_t*>(ct),
                             count, out);
            return CUFHE_OK;
        case 2:
            phase<lvl2param>(static_cast<const uint64_t*>(key), static_cast<const uint64_t*>(ct),
                             count, out);
            return CUFHE_OK;
        default:
            return CUFHE_ERR_ARG;
    }
}

int tfhe_tlwe_decrypt_bits(int level, const void* key, const void* ct, size_t count,
                           uint8_t* bits)
{
    if (!bits && count) return CUFHE_ERR_ARG;
    std::vector<double> ph(count);
    const int rc = tfhe_tlwe_phase(level, key, ct, count, ph.data());
    if (rc != CUFHE_OK) return rc;
    for (size_t i = 0; i < count; i++) bits[i] = ph[i] > 0.0 ? 1 : 0;
    return CUFHE_OK;
}

size_t tfhe_tlwe_words(int level)
{
    switch (level) {
        case 0: return lvl0param::k * lvl0param::n + 1;
        case 1: return lvl1param::k * lvl1param::n + 1;
        case 2: return lvl2param::k * lvl2param::n + 1;
        default: return 0;
    }
}

size_t tfhe_bk_words(int target_level, int key_bundle)
{
    if (target_level == 1) return kBKTRGSWs<lvl01param>(key_bundle != 0) * kTRGSWWords<lvl1param>;
    if (target_level == 2) return kBKTRGSWs<lvl02param>(key_bundle != 0) * kTRGSWWords<lvl2param>;
    return 0;
}

size_t tfhe_ksk_subset_words(void)
{
    return kKSKWords<lvl10param> / lvl1param::n * (lvl1param::n - lvl0param::n);
}

size_t tfhe_ksk_words(int domain_level)
{
    if (domain_level == 1) return kKSKWords<lvl10param>;
    if (domain_level == 2) return kKSKWords<lvl20param>;
    return 0;
}

namespace {
constexpr char kKeyMagic[8] = {'C', 'U', 'F', 'H', 'E', 'K', 'Y', '1'};

// Element size and expected element count of record tag t under `flags`.
bool keyfile_shape(int tag, uint32_t flags, uint32_t* esize, size_t* count)
{
    const int kb = (flags & 1u) ? 1 : 0;
    switch (tag) {
        case TFHE_KEYFILE_BKLVL01: *esize = 4; *count = tfhe_bk_words(1, kb); return true;
        case TFHE_KEYFILE_BKLVL02: *esize = 8; *count = tfhe_bk_words(2, kb); return true;
        case TFHE_KEYFILE_IKSKLVL10:
            *esize = 4;
            *count = (flags & 4u) ? tfhe_ksk_subset_words() : tfhe_ksk_words(1);
            return true;
        case TFHE_KEYFILE_IKSKLVL20: *esize = 4; *count = tfhe_ksk_words(2); return true;
        case TFHE_KEYFILE_SKLVL0: *esize = 4; *count = lvl0param::n; return true;
        case TFHE_KEYFILE_SKLVL1: *esize = 4; *count = lvl1param::n; return true;
        case TFHE_KEYFILE_SKLVL2: *esize = 8; *count = lvl2param::n; return true;
        default: return false;
    }
}

struct FileCloser {
    FILE* f;
    ~FileCloser()
    {
        if (f) std::fclose(f);
    }
};
}  // namespace

int tfhe_keyfile_save(const char* path, uint32_t flags, const void* const* ptrs,
                      const size_t* counts)
{
    if (!path || !ptrs || !counts || (flags >> 4) != 0) return -1;
    uint32_t nrec = 0;
    for (int t = 1; t <= TFHE_KEYFILE_RECORDS; t++) {
        if (!ptrs[t - 1] || !counts[t - 1]) continue;
        uint32_t es = 0;
        size_t want = 0;
        if (!keyfile_shape(t, flags, &es, &want) || counts[t - 1] != want) return -1;
        nrec++;
    }
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return -2;
    bool ok = std::fwrite(kKeyMagic, 1, 8, fc.f) == 8 && std::fwrite(&flags, 4, 1, fc.f) == 1 &&
              std::fwrite(&nrec, 4, 1, fc.f) == 1;
    for (int t = 1; ok && t <= TFHE_KEYFILE_RECORDS; t++) {
        if (!ptrs[t - 1] || !counts[t - 1]) continue;
        uint32_t es = 0;
        size_t want = 0;
        keyfile_shape(t, flags, &es, &want);
        const uint32_t tag = static_cast<uint32_t>(t);
        const uint64_t n = counts[t - 1];
        ok = std::fwrite(&tag, 4, 1, fc.f) == 1 && std::fwrite(&es, 4, 1, fc.f) == 1 &&
             std::fwrite(&n, 8, 1, fc.f) == 1 && std::fwrite(ptrs[t - 1], es, n, fc.f) == n;
    }
    if (!ok) return -2;
    return std::fflush(fc.f) == 0 ? 0 : -2;
}

// Walk the records; with ptrs, copy each present array into ptrs[tag-1].
static int keyfile_walk(const char* path, uint32_t* flags_out, size_t* counts, void* const* ptrs)
{
    if (!path) return -1;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return -2;
    char magic[8];
    uint32_t flags = 0, nrec = 0;
    if (std::fread(magic, 1, 8, fc.f) != 8 || std::memcmp(magic, kKeyMagic, 8) != 0) return -1;
    if (std::fread(&flags, 4, 1, fc.f) != 1 || std::fread(&nrec, 4, 1, fc.f) != 1) return -1;
    if ((flags >> 4) != 0 || nrec > TFHE_KEYFILE_RECORDS) return -1;
    if (counts)
        for (int t = 0; t < TFHE_KEYFILE_RECORDS; t++) counts[t] = 0;
    bool seen[TFHE_KEYFILE_RECORDS] = {};
    for (uint32_t i = 0; i < nrec; i++) {
        uint32_t tag = 0, es = 0;
        uint64_t n = 0;
        if (std::fread(&tag, 4, 1, fc.f) != 1 || std::fread(&es, 4, 1, fc.f) != 1 ||
            std::fread(&n, 8, 1, fc.f) != 1)
            return -1;
        uint32_t want_es = 0;
        size_t want_n = 0;
        if (tag < 1 || tag > TFHE_KEYFILE_RECORDS || seen[tag - 1] ||
            !keyfile_shape(static_cast<int>(tag), flags, &want_es, &want_n) || es != want_es ||
            n != want_n)
            return -1;
        seen[tag - 1] = true;
        if (counts) counts[tag - 1] = static_cast<size_t>(n);
        void* dst = ptrs ? ptrs[tag - 1] : nullptr;
        if (dst) {
            if (std::fread(dst, es, n, fc.f) != n) return -2;
        }
        else if (std::fseek(fc.f, static_cast<long>(es * n), SEEK_CUR) != 0) {
            return -2;
        }
    }
    if (flags_out) *flags_out = flags;
    return 0;
}

int tfhe_keyfile_info(const char* path, uint32_t* flags, size_t* counts)
{
    if (!flags || !counts) return -1;
    return keyfile_walk(path, flags, counts, nullptr);
}

int tfhe_keyfile_load(const char* path, void* const* ptrs)
{
    if (!ptrs) return -1;
    return keyfile_walk(path, nullptr, nullptr, ptrs);
}

}  // extern "C"
