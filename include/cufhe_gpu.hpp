// C++ drop-in for cuFHEpp's public gate API (reference
// include/cufhe_gpu.cuh:40-358, src/cufhe_gates_gpu.cu:34-977): the same
// names (SetGPUNum, Initialize, Ctxt, Stream, Nand ... NMux, g* variants,
// *_lvl02 gates) and the same argument meaning, implemented over the C ABI of
// libcufhe_amd.so (cufhe_amd.h). Errors print and exit(-1) as the reference's
// CuSafeCall does (include/error_gpu.cuh:30-66).
//
// Differences a caller can see: one GPU per process (SetGPUNum(n > 1) is
// refused); the backend (FFT / NTT, block-binary) follows the EvalKey instead
// of a build option.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "cufhe_amd.h"
#include "error_gpu.hpp"
#include "tfhe_host.hpp"

namespace cufhe {

inline int _gpuNum = 1;
inline int streamCount = 0;

inline void CuCheck(int rc, const char* what)
{
    if (rc != CUFHE_OK) {
        std::fprintf(stderr, "cufhe: %s failed (%d): %s\n", what, rc, cufhe_last_error());
        std::exit(-1);
    }
}

inline void SetGPUNum(int gpuNum)
{
    CuCheck(cufhe_set_gpu_num(gpuNum), "SetGPUNum");
    _gpuNum = gpuNum;
}

inline void Initialize() { CuCheck(cufhe_initialize(), "Initialize"); }

inline void Initialize(const TFHEpp::EvalKey& ek)
{
    if (!ek.bklvl01 || !ek.iksklvl10) {
        std::fprintf(stderr, "cufhe: Initialize(ek) needs bklvl01 and iksklvl10\n");
        std::exit(-1);
    }
    // reference build options as key properties: block-binary keys run on the
    // NTT backend (USE_BLOCK_BINARY implies USE_FFT=OFF), others on the FFT one
    // unless ek.ntt asks for the NTT backend (USE_FFT=OFF)
    const int backend = (ek.block_binary || ek.ntt) ? CUFHE_BACKEND_NTT : CUFHE_BACKEND_FFT;
    const int variant = ek.block_binary ? CUFHE_BR_BLOCK_BINARY
                                        : (ek.key_bundle ? CUFHE_BR_KEY_BUNDLE : CUFHE_BR_PLAIN);
    CuCheck(cufhe_initialize_lvl01_ex(ek.bklvl01->data(), ek.bklvl01->size(), backend, variant,
                                      ek.iksklvl10->data(), ek.iksklvl10->size(),
                                      ek.subset_iksk ? 1 : 0),
            "Initialize");
}

inline void Initialize_lvl02(const TFHEpp::EvalKey& ek, const TFHEpp::SecretKey&)
{
    if (!ek.bklvl02 || !ek.iksklvl20) {
        std::fprintf(stderr, "cufhe: Initialize_lvl02 needs bklvl02 and iksklvl20\n");
        std::exit(-1);
    }
    // ek.ntt: the lvl2 rotation on the Goldilocks NTT backend (USE_FFT=OFF)
    CuCheck(cufhe_initialize_lvl02_ex(ek.bklvl02->data(), ek.bklvl02->size(),
                                      ek.ntt ? CUFHE_BACKEND_NTT : CUFHE_BACKEND_FFT,
                                      ek.key_bundle, ek.iksklvl20->data(), ek.iksklvl20->size()),
            "Initialize_lvl02");
}

inline void CleanUp() { CuCheck(cufhe_cleanup(), "CleanUp"); }
inline void CleanUp_lvl02() { CuCheck(cufhe_cleanup_lvl02(), "CleanUp_lvl02"); }
inline void Synchronize() { CuCheck(cufhe_synchronize(), "Synchronize"); }

/** One key of a level at a time, the other kept (reference
 *  include/bootstrap_gpu.cuh:40-51, include/keyswitch_gpu.cuh:258-271); the
 *  level's gates run once both are loaded. The bootstrapping key's layout
 *  (one TRGSW per lvl0 key bit, or the 2-bit bundle) follows from its size;
 *  backend: CUFHE_BACKEND_FFT (the reference's default build) or
 *  CUFHE_BACKEND_NTT; block_binary: the block-binary rotation (lvl01param,
 *  NTT). The subset lvl10 key is recognised by its size. gpuNum is the
 *  reference's GPU count (one GPU per process here). */
inline void InitializeNTThandlers(const int) { CuCheck(cufhe_initialize(), "InitializeNTThandlers"); }
inline void InitializeNTThandlers_lvl02(const int)
{
    CuCheck(cufhe_initialize(), "InitializeNTThandlers_lvl02");
}
template <class brP>
void BootstrappingKeyToNTT(const TFHEpp::BootstrappingKey<brP>& bk, const int,
                           const int backend = CUFHE_BACKEND_FFT, const bool block_binary = false)
{
    static_assert(std::is_same_v<brP, TFHEpp::lvl01param> || std::is_same_v<brP, TFHEpp::lvl02param>,
                  "bootstrapping keys: lvl01param, lvl02param");
    constexpr int level = std::is_same_v<brP, TFHEpp::lvl01param> ? 1 : 2;
    const bool kb = bk.size() == tfhe_bk_words(level, 1);
    const int variant = block_binary ? CUFHE_BR_BLOCK_BINARY
                                     : (kb ? CUFHE_BR_KEY_BUNDLE : CUFHE_BR_PLAIN);
    CuCheck(cufhe_bk_load(level, bk.data(), bk.size(), block_binary ? CUFHE_BACKEND_NTT : backend,
                          variant),
            "BootstrappingKeyToNTT");
}
template <class brP>
void BootstrappingKeyBundleToNTT(const TFHEpp::BootstrappingKey<brP>& bk, const int gpuNum,
                                 const int backend = CUFHE_BACKEND_FFT)
{
    constexpr int level = std::is_same_v<brP, TFHEpp::lvl01param> ? 1 : 2;
    if (bk.size() != tfhe_bk_words(level, 1)) {
        std::fprintf(stderr, "cufhe: BootstrappingKeyBundleToNTT: not a key-bundle key\n");
        std::exit(-1);
    }
    BootstrappingKeyToNTT<brP>(bk, gpuNum, backend);
}
/** BootstrappingKeyFlatToNTT_lvl02 (reference src/bootstrap_gpu.cu:382-452):
 *  the lvl02 bootstrapping key as num_elements TRGSWs in a row (n of them,
 *  or 3 n / 2 in the key-bundle layout). */
inline void BootstrappingKeyFlatToNTT_lvl02(const TFHEpp::TRGSW<TFHEpp::lvl2param>* flat_bk,
                                            const uint32_t num_elements, const int,
                                            const int backend = CUFHE_BACKEND_FFT)
{
    const size_t words = static_cast<size_t>(num_elements) * TFHEpp::kTRGSWWords<TFHEpp::lvl2param>;
    const int variant = words == tfhe_bk_words(2, 1) ? CUFHE_BR_KEY_BUNDLE : CUFHE_BR_PLAIN;
    CuCheck(cufhe_bk_load(2, flat_bk, words, backend, variant), "BootstrappingKeyFlatToNTT_lvl02");
}
inline void DeleteBootstrappingKeyNTT(const int) { CuCheck(cufhe_bk_delete(1), "DeleteBootstrappingKeyNTT"); }
inline void DeleteBootstrappingKeyNTT_lvl02(const int)
{
    CuCheck(cufhe_bk_delete(2), "DeleteBootstrappingKeyNTT_lvl02");
}
inline void KeySwitchingKeyToDevice(const TFHEpp::KeySwitchingKey<TFHEpp::lvl10param>& ksk, const int)
{
    const int subset = ksk.size() == tfhe_ksk_subset_words() ? 1 : 0;
    CuCheck(cufhe_ksk_load(1, ksk.data(), ksk.size(), subset), "KeySwitchingKeyToDevice");
}
inline void DeleteKeySwitchingKey(const int) { CuCheck(cufhe_ksk_delete(1), "DeleteKeySwitchingKey"); }
inline void KeySwitchingKeyToDevice_lvl20(const TFHEpp::KeySwitchingKey<TFHEpp::lvl20param>& ksk,
                                          const int)
{
    CuCheck(cufhe_ksk_load(2, ksk.data(), ksk.size(), 0), "KeySwitchingKeyToDevice_lvl20");
}
inline void DeleteKeySwitchingKey_lvl20(const int)
{
    CuCheck(cufhe_ksk_delete(2), "DeleteKeySwitchingKey_lvl20");
}

/** Device copies of a host object, one per GPU (reference
 *  cufhe_gpu.cuh:78-98). The host object is not page-locked: copies go
 *  through the runtime's staging buffers (registering existing host memory
 *  is not used by this library). */
template <typename dT, typename hT>
void ctxtInitialize(hT& host, std::vector<dT*>& devices)
{
    devices.resize(_gpuNum);
    for (int i = 0; i < _gpuNum; i++) {
        void* p = nullptr;
        CuCheck(cufhe_device_alloc(&p, sizeof(host)), "ctxtInitialize");
        devices[i] = static_cast<dT*>(p);
    }
}

template <typename dT, typename hT>
void ctxtDelete(hT&, std::vector<dT*>& devices)
{
    for (auto*& p : devices) {
        cufhe_device_free(p);
        p = nullptr;
    }
}

/** Ciphertext: host TLWE plus one device copy (reference cufhe_gpu.cuh:100-121). */
template <class P>
struct Ctxt {
    Ctxt()
    {
        tlwedevices.resize(_gpuNum);
        void* p = nullptr;
        CuCheck(cufhe_device_alloc(&p, sizeof(tlwehost)), "Ctxt alloc");
        tlwedevices[0] = static_cast<typename P::T*>(p);
    }
    ~Ctxt()
    {
        for (auto* p : tlwedevices) cufhe_device_free(p);
    }
    Ctxt(const Ctxt&) = delete;
    Ctxt& operator=(const Ctxt&) = delete;

    alignas(64) TFHEpp::TLWE<P> tlwehost;
    std::vector<typename P::T*> tlwedevices;
};

/** Stream handle (reference cufhe_gpu.cuh:162-199). */
class Stream {
   public:
    Stream() : st_(nullptr), device_id_(streamCount % _gpuNum) { streamCount++; }
    explicit Stream(int device_id) : st_(nullptr), device_id_(device_id) { streamCount++; }
    void Create()
    {
        void* s = nullptr;
        CuCheck(cufhe_stream_create(&s), "Stream::Create");
        st_ = static_cast<hipStream_t>(s);
    }
    void Destroy()
    {
        CuCheck(cufhe_stream_destroy(st_), "Stream::Destroy");
        st_ = nullptr;
    }
    hipStream_t st() { return st_; }
    int device_id() { return device_id_; }

   private:
    hipStream_t st_;
    int device_id_;
};

inline bool StreamQuery(Stream st)
{
    const int r = cufhe_stream_query(st.st());
    CuCheck(r < 0 ? r : CUFHE_OK, "StreamQuery");
    return r == 1;
}

/** lvl1 TRLWE with one device copy (reference cufhe_gpu.cuh:123-160). */
struct cuFHETRLWElvl1 {
    cuFHETRLWElvl1()
    {
        void* p = nullptr;
        CuCheck(cufhe_device_alloc(&p, sizeof(trlwehost)), "TRLWE alloc");
        trlwedevices.assign(_gpuNum, static_cast<uint32_t*>(p));
    }
    ~cuFHETRLWElvl1() { cufhe_device_free(trlwedevices[0]); }
    cuFHETRLWElvl1(const cuFHETRLWElvl1&) = delete;
    cuFHETRLWElvl1& operator=(const cuFHETRLWElvl1&) = delete;

    TFHEpp::TRLWE<TFHEpp::lvl1param> trlwehost;  // (a, b), 2 x 1024 words
    std::vector<uint32_t*> trlwedevices;
};

/** lvl1 TRGSW in the transform domain of the loaded lvl1 backend (reference
 *  cufhe_gpu.cuh:136-156 cuFHETRGSWNTTlvl1): FFT spectra or NTT residues,
 *  cufhe_trgsw_transform_bytes_lvl1() bytes; construct after the keys are
 *  loaded (they select the backend). */
struct cuFHETRGSWNTTlvl1 {
    cuFHETRGSWNTTlvl1() : bytes(cufhe_trgsw_transform_bytes_lvl1())
    {
        if (bytes == 0) {
            std::fprintf(stderr, "cuFHETRGSWNTTlvl1: load the lvl1 keys first\n");
            std::exit(-1);
        }
        void* p = nullptr;
        CuCheck(cufhe_device_alloc(&p, bytes), "TRGSW alloc");
        trgswhost.resize(bytes);
        trgswdevices.assign(_gpuNum, p);
    }
    ~cuFHETRGSWNTTlvl1() { cufhe_device_free(trgswdevices[0]); }
    cuFHETRGSWNTTlvl1(const cuFHETRGSWNTTlvl1&) = delete;
    cuFHETRGSWNTTlvl1& operator=(const cuFHETRGSWNTTlvl1&) = delete;

    size_t bytes;
    std::vector<uint8_t> trgswhost;  // host copy of the transformed TRGSW
    std::vector<void*> trgswdevices;
};

template <class P>
inline void CtxtCopyH2D(Ctxt<P>& c, Stream st)
{
    CuCheck(cufhe_memcpy_h2d(c.tlwedevices[0], c.tlwehost.data(), sizeof(c.tlwehost), st.st()),
            "CtxtCopyH2D");
}

template <class P>
inline void CtxtCopyD2H(Ctxt<P>& c, Stream st)
{
    CuCheck(cufhe_memcpy_d2h(c.tlwehost.data(), c.tlwedevices[0], sizeof(c.tlwehost), st.st()),
            "CtxtCopyD2H");
}

template <class P>
void CopyOnHost(Ctxt<P>& out, Ctxt<P>& in)
{
    out.tlwehost = in.tlwehost;
}

namespace detail {
// lvl1 ciphertexts: IKS-BR gates; lvl0 ciphertexts: BR-IKS gates at lvl1
// (reference src/cufhe_gates_gpu.cu:198-511 and :514-825).
template <class P>
inline void gate(int op, Ctxt<P>& out, Ctxt<P>* in0, Ctxt<P>* in1, Ctxt<P>* inc, Stream st)
{
    static_assert(std::is_same_v<P, TFHEpp::lvl1param> || std::is_same_v<P, TFHEpp::lvl0param>,
                  "gates take Ctxt<lvl0param> or Ctxt<lvl1param>");
    auto* fn = std::is_same_v<P, TFHEpp::lvl1param> ? cufhe_gate_lvl1 : cufhe_gate_lvl0;
    CuCheck(fn(op, out.tlwedevices[0], in0 ? in0->tlwedevices[0] : nullptr,
               in1 ? in1->tlwedevices[0] : nullptr, inc ? inc->tlwedevices[0] : nullptr, st.st()),
            "gate");
}
template <class P>
inline void host_gate(int op, Ctxt<P>& out, Ctxt<P>* in0, Ctxt<P>* in1, Ctxt<P>* inc, Stream st)
{
    for (Ctxt<P>* c : {in0, in1, inc})
        if (c) CtxtCopyH2D<P>(*c, st);
    gate<P>(op, out, in0, in1, inc, st);
    CtxtCopyD2H<P>(out, st);
}
inline void gate02(int op, Ctxt<TFHEpp::lvl0param>& out, Ctxt<TFHEpp::lvl0param>* in0,
                   Ctxt<TFHEpp::lvl0param>* in1, Ctxt<TFHEpp::lvl0param>* inc, Stream st)
{
    using C0 = Ctxt<TFHEpp::lvl0param>;
    for (C0* c : {in0, in1, inc})
        if (c) CtxtCopyH2D<TFHEpp::lvl0param>(*c, st);
    CuCheck(cufhe_gate_lvl02(op, out.tlwedevices[0], in0 ? in0->tlwedevices[0] : nullptr,
                             in1 ? in1->tlwedevices[0] : nullptr,
                             inc ? inc->tlwedevices[0] : nullptr, st.st()),
            "gate_lvl02");
    CtxtCopyD2H<TFHEpp::lvl0param>(out, st);
}
}  // namespace detail

// Host-buffer gates: copy inputs H2D, bootstrap, copy the result D2H (all
// asynchronous on st; results valid after Synchronize()).
#define CUFHE_BINARY_GATE(Name, Code)                                          \
    template <class P>                                                         \
    void Name(Ctxt<P>& out, Ctxt<P>& in0, Ctxt<P>& in1, Stream st)            \
    {                                                                          \
        detail::host_gate<P>(Code, out, &in0, &in1, nullptr, st);              \
    }                                                                          \
    template <class P>                                                         \
    void g##Name(Ctxt<P>& out, Ctxt<P>& in0, Ctxt<P>& in1, Stream st)         \
    {                                                                          \
        detail::gate<P>(Code, out, &in0, &in1, nullptr, st);                   \
    }                                                                          \
    inline void Name##_lvl02(Ctxt<TFHEpp::lvl0param>& out, Ctxt<TFHEpp::lvl0param>& in0, \
                             Ctxt<TFHEpp::lvl0param>& in1, Stream st)          \
    {                                                                          \
        detail::gate02(Code, out, &in0, &in1, nullptr, st);                    \
    }
CUFHE_BINARY_GATE(Nand, CUFHE_NAND)
CUFHE_BINARY_GATE(Or, CUFHE_OR)
CUFHE_BINARY_GATE(OrYN, CUFHE_ORYN)
CUFHE_BINARY_GATE(OrNY, CUFHE_ORNY)
CUFHE_BINARY_GATE(And, CUFHE_AND)
CUFHE_BINARY_GATE(AndYN, CUFHE_ANDYN)
CUFHE_BINARY_GATE(AndNY, CUFHE_ANDNY)
CUFHE_BINARY_GATE(Nor, CUFHE_NOR)
CUFHE_BINARY_GATE(Xor, CUFHE_XOR)
CUFHE_BINARY_GATE(Xnor, CUFHE_XNOR)
#undef CUFHE_BINARY_GATE

// Mux(out, inc, in1, in0): out = inc ? in1 : in0 (reference argument order).
template <class P>
void Mux(Ctxt<P>& out, Ctxt<P>& inc, Ctxt<P>& in1, Ctxt<P>& in0, Stream st)
{
    detail::host_gate<P>(CUFHE_MUX, out, &in0, &in1, &inc, st);
}
template <class P>
void gMux(Ctxt<P>& out, Ctxt<P>& inc, Ctxt<P>& in1, Ctxt<P>& in0, Stream st)
{
    detail::gate<P>(CUFHE_MUX, out, &in0, &in1, &inc, st);
}
template <class P>
void NMux(Ctxt<P>& out, Ctxt<P>& inc, Ctxt<P>& in1, Ctxt<P>& in0, Stream st)
{
    detail::host_gate<P>(CUFHE_NMUX, out, &in0, &in1, &inc, st);
}
template <class P>
void gNMux(Ctxt<P>& out, Ctxt<P>& inc, Ctxt<P>& in1, Ctxt<P>& in0, Stream st)
{
    detail::gate<P>(CUFHE_NMUX, out, &in0, &in1, &inc, st);
}
inline void Mux_lvl02(Ctxt<TFHEpp::lvl0param>& out, Ctxt<TFHEpp::lvl0param>& inc,
                      Ctxt<TFHEpp::lvl0param>& in1, Ctxt<TFHEpp::lvl0param>& in0, Stream st)
{
    detail::gate02(CUFHE_MUX, out, &in0, &in1, &inc, st);
}
inline void NMux_lvl02(Ctxt<TFHEpp::lvl0param>& out, Ctxt<TFHEpp::lvl0param>& inc,
                       Ctxt<TFHEpp::lvl0param>& in1, Ctxt<TFHEpp::lvl0param>& in0, Stream st)
{
    detail::gate02(CUFHE_NMUX, out, &in0, &in1, &inc, st);
}

template <class P>
void Not(Ctxt<P>& out, Ctxt<P>& in, Stream st)
{
    detail::host_gate<P>(CUFHE_NOT, out, &in, nullptr, nullptr, st);
}
template <class P>
void gNot(Ctxt<P>& out, Ctxt<P>& in, Stream st)
{
    detail::gate<P>(CUFHE_NOT, out, &in, nullptr, nullptr, st);
}
template <class P>
void Copy(Ctxt<P>& out, Ctxt<P>& in, Stream st)
{
    detail::host_gate<P>(CUFHE_COPY, out, &in, nullptr, nullptr, st);
}
template <class P>
void gCopy(Ctxt<P>& out, Ctxt<P>& in, Stream st)
{
    detail::gate<P>(CUFHE_COPY, out, &in, nullptr, nullptr, st);
}

// ---- lower launcher layer (reference include/bootstrap_gpu.cuh:40-217,
// include/keyswitch_gpu.cuh:258-275): raw device pointers, a stream and the
// GPU index (0: one GPU per process). Asynchronous on st like the reference.
namespace detail {
inline void check_gpu(int gpuNum)
{
    if (gpuNum != 0) {
        std::fprintf(stderr, "cufhe: gpuNum %d: one GPU per process (index 0)\n", gpuNum);
        std::exit(-1);
    }
}
// BR-IKS order: lvl0 in and out, rotation at brP's target level.
template <class brP, class iksP>
inline void br_iks(int op, typename iksP::targetP::T* out, const typename brP::domainP::T* in0,
                   const typename brP::domainP::T* in1, const typename brP::domainP::T* inc,
                   hipStream_t st, int gpuNum)
{
    check_gpu(gpuNum);
    static_assert(std::is_same_v<typename brP::domainP, TFHEpp::lvl0param> &&
                      std::is_same_v<typename iksP::targetP, TFHEpp::lvl0param> &&
                      std::is_same_v<typename brP::targetP, typename iksP::domainP>,
                  "BR-IKS gates: lvl0 -> lvl1 -> lvl0 or lvl0 -> lvl2 -> lvl0");
    auto* fn = std::is_same_v<brP, TFHEpp::lvl01param> ? cufhe_gate_lvl0 : cufhe_gate_lvl02;
    CuCheck(fn(op, out, in0, in1, inc, st), "XxxBootstrap (BR-IKS)");
}
// IKS-BR order: lvl1 in and out.
template <class iksP, class brP>
inline void iks_br(int op, typename brP::targetP::T* out, const typename iksP::domainP::T* in0,
                   const typename iksP::domainP::T* in1, const typename iksP::domainP::T* inc,
                   hipStream_t st, int gpuNum)
{
    check_gpu(gpuNum);
    static_assert(std::is_same_v<iksP, TFHEpp::lvl10param> &&
                      std::is_same_v<brP, TFHEpp::lvl01param>,
                  "IKS-BR gates: lvl1 -> lvl0 -> lvl1");
    CuCheck(cufhe_gate_lvl1(op, out, in0, in1, inc, st), "XxxBootstrap (IKS-BR)");
}
}  // namespace detail

// XxxBootstrap<brP, mu, iksP>(out, in0, in1, st, gpu): BR-IKS (lvl0 in/out);
// XxxBootstrap<iksP, brP, mu>(out, in0, in1, st, gpu): IKS-BR (lvl1 in/out).
// mu must be the target level's mu, the only value the reference's gate API
// passes (src/cufhe_gates_gpu.cu:206-977).
#define CUFHE_BOOTSTRAP_LAUNCHER(Name, Code)                                                     \
    template <class brP, typename brP::targetP::T μ, class iksP>                                 \
    void Name##Bootstrap(typename iksP::targetP::T* const out,                                  \
                         const typename brP::domainP::T* const in0,                             \
                         const typename brP::domainP::T* const in1, const hipStream_t st,       \
                         const int gpuNum)                                                      \
    {                                                                                           \
        static_assert(μ == static_cast<typename brP::targetP::T>(brP::targetP::μ),             \
                      "test vector mu must be the target level's mu");                          \
        detail::br_iks<brP, iksP>(Code, out, in0, in1, nullptr, st, gpuNum);                     \
    }                                                                                           \
    template <class iksP, class brP, typename brP::targetP::T μ>                                 \
    void Name##Bootstrap(typename brP::targetP::T* const out,                                   \
                         const typename iksP::domainP::T* const in0,                            \
                         const typename iksP::domainP::T* const in1, const hipStream_t st,      \
                         const int gpuNum)                                                      \
    {                                                                                           \
        static_assert(μ == static_cast<typename brP::targetP::T>(brP::targetP::μ),             \
                      "test vector mu must be the target level's mu");                          \
        detail::iks_br<iksP, brP>(Code, out, in0, in1, nullptr, st, gpuNum);                     \
    }
CUFHE_BOOTSTRAP_LAUNCHER(Nand, CUFHE_NAND)
CUFHE_BOOTSTRAP_LAUNCHER(Or, CUFHE_OR)
CUFHE_BOOTSTRAP_LAUNCHER(OrYN, CUFHE_ORYN)
CUFHE_BOOTSTRAP_LAUNCHER(OrNY, CUFHE_ORNY)
CUFHE_BOOTSTRAP_LAUNCHER(And, CUFHE_AND)
CUFHE_BOOTSTRAP_LAUNCHER(AndYN, CUFHE_ANDYN)
CUFHE_BOOTSTRAP_LAUNCHER(AndNY, CUFHE_ANDNY)
CUFHE_BOOTSTRAP_LAUNCHER(Nor, CUFHE_NOR)
CUFHE_BOOTSTRAP_LAUNCHER(Xor, CUFHE_XOR)
CUFHE_BOOTSTRAP_LAUNCHER(Xnor, CUFHE_XNOR)
#undef CUFHE_BOOTSTRAP_LAUNCHER

// MuxBootstrap / NMuxBootstrap: out = inc ? in1 : in0 (negated for NMux).
#define CUFHE_MUX_LAUNCHER(Name, Code)                                                           \
    template <class brP, typename brP::targetP::T μ, class iksP>                                 \
    void Name##Bootstrap(typename iksP::targetP::T* const out,                                  \
                         const typename brP::domainP::T* const inc,                             \
                         const typename brP::domainP::T* const in1,                             \
                         const typename brP::domainP::T* const in0, const hipStream_t st,       \
                         const int gpuNum)                                                      \
    {                                                                                           \
        static_assert(μ == static_cast<typename brP::targetP::T>(brP::targetP::μ),             \
                      "test vector mu must be the target level's mu");                          \
        detail::br_iks<brP, iksP>(Code, out, in0, in1, inc, st, gpuNum);                         \
    }                                                                                           \
    template <class iksP, class brP, typename brP::targetP::T μ>                                 \
    void Name##Bootstrap(typename brP::targetP::T* const out,                                   \
                         const typename iksP::domainP::T* const inc,                            \
                         const typename iksP::domainP::T* const in1,                            \
                         const typename iksP::domainP::T* const in0, const hipStream_t st,      \
                         const int gpuNum)                                                      \
    {                                                                                           \
        static_assert(μ == static_cast<typename brP::targetP::T>(brP::targetP::μ),             \
                      "test vector mu must be the target level's mu");                          \
        detail::iks_br<iksP, brP>(Code, out, in0, in1, inc, st, gpuNum);                         \
    }
CUFHE_MUX_LAUNCHER(Mux, CUFHE_MUX)
CUFHE_MUX_LAUNCHER(NMux, CUFHE_NMUX)
#undef CUFHE_MUX_LAUNCHER

// CopyBootstrap<P> / NotBootstrap<P>: coefficient-wise (reference
// src/bootstrap_gpu.cu:2200-2222, :3607-3640).
template <class P>
void CopyBootstrap(typename P::T* const out, const typename P::T* const in, const hipStream_t st,
                   const int gpuNum)
{
    detail::check_gpu(gpuNum);
    static_assert(std::is_same_v<P, TFHEpp::lvl0param> || std::is_same_v<P, TFHEpp::lvl1param>,
                  "CopyBootstrap: lvl0 or lvl1 TLWEs");
    auto* fn = std::is_same_v<P, TFHEpp::lvl1param> ? cufhe_gate_lvl1 : cufhe_gate_lvl0;
    CuCheck(fn(CUFHE_COPY, out, in, nullptr, nullptr, st), "CopyBootstrap");
}
template <class P>
void NotBootstrap(typename P::T* const out, const typename P::T* const in, const hipStream_t st,
                  const int gpuNum)
{
    detail::check_gpu(gpuNum);
    static_assert(std::is_same_v<P, TFHEpp::lvl0param> || std::is_same_v<P, TFHEpp::lvl1param>,
                  "NotBootstrap: lvl0 or lvl1 TLWEs");
    auto* fn = std::is_same_v<P, TFHEpp::lvl1param> ? cufhe_gate_lvl1 : cufhe_gate_lvl0;
    CuCheck(fn(CUFHE_NOT, out, in, nullptr, nullptr, st), "NotBootstrap");
}

// Bootstrap(out, in, mu, st, gpu) (reference include/bootstrap_gpu.cuh:61-64,
// src/bootstrap_gpu.cu:1045-1058): lvl0 -> blind rotation with test vector
// +-mu at lvl1 -> SampleExtract -> key switch -> lvl0.
inline void Bootstrap(TFHEpp::lvl0param::T* const out, const TFHEpp::lvl0param::T* const in,
                      const TFHEpp::lvl1param::T mu, const hipStream_t st, const int gpuNum)
{
    detail::check_gpu(gpuNum);
    CuCheck(cufhe_bootstrap(out, in, mu, st), "Bootstrap");
}

// BootstrapTLWE2TRLWE(out, in, mu, st, gpu) (reference
// src/bootstrap_gpu.cu:2371-2388): the rotated test-vector TRLWE, no extraction.
inline void BootstrapTLWE2TRLWE(TFHEpp::lvl1param::T* const out,
                                const TFHEpp::lvl0param::T* const in,
                                const TFHEpp::lvl1param::T mu, const hipStream_t st,
                                const int gpuNum)
{
    detail::check_gpu(gpuNum);
    CuCheck(cufhe_bootstrap_to_trlwe_lvl01(out, in, mu, 1, st), "BootstrapTLWE2TRLWE");
}

// SEIandKS(out, in, st, gpu) (reference include/keyswitch_gpu.cuh:273-275,
// src/keyswitch_gpu.cu:61-68): SampleExtract index 0 of a lvl1 TRLWE and key
// switch to lvl0.
inline void SEIandKS(TFHEpp::lvl0param::T* const out, const TFHEpp::lvl1param::T* const in,
                     const hipStream_t& st, const int gpuNum)
{
    detail::check_gpu(gpuNum);
    CuCheck(cufhe_sample_extract_key_switch_lvl10(out, in, 1, st), "SEIandKS");
}

// IdentityKeySwitch<iksP>(out, in, st, gpu): KeySwitchFromTLWE<iksP> of one
// extracted TLWE (reference include/keyswitch_gpu.cuh:128-187), lvl10 or lvl20.
template <class iksP>
void IdentityKeySwitch(TFHEpp::lvl0param::T* const out, const typename iksP::domainP::T* const in,
                       const hipStream_t st, const int gpuNum)
{
    detail::check_gpu(gpuNum);
    if constexpr (std::is_same_v<iksP, TFHEpp::lvl10param>)
        CuCheck(cufhe_identity_key_switch_lvl10(out, in, 1, st), "IdentityKeySwitch<lvl10param>");
    else {
        static_assert(std::is_same_v<iksP, TFHEpp::lvl20param>, "lvl10param or lvl20param");
        CuCheck(cufhe_identity_key_switch_lvl20(out, in, 1, st), "IdentityKeySwitch<lvl20param>");
    }
}

// TRGSW2NTT(out, trgsw, st) (reference src/bootstrap_gpu.cu:165-186): the
// TRGSW to the transform domain on the device, and back into out's host copy
// (asynchronous; valid after Synchronize()).
inline void TRGSW2NTT(cuFHETRGSWNTTlvl1& out, const TFHEpp::TRGSW<TFHEpp::lvl1param>& trgsw,
                      Stream& st)
{
    void* d = nullptr;
    CuCheck(cufhe_device_alloc(&d, sizeof(trgsw)), "TRGSW2NTT alloc");
    CuCheck(cufhe_memcpy_h2d(d, trgsw.data(), sizeof(trgsw), st.st()), "TRGSW2NTT H2D");
    CuCheck(cufhe_trgsw_to_transform_lvl1(static_cast<const uint32_t*>(d), out.trgswdevices[0], 1,
                                          st.st()),
            "TRGSW2NTT");
    CuCheck(cufhe_memcpy_d2h(out.trgswhost.data(), out.trgswdevices[0], out.bytes, st.st()),
            "TRGSW2NTT D2H");
    CuCheck(cufhe_stream_synchronize(st.st()), "TRGSW2NTT sync");
    cufhe_device_free(d);
}

// CMUXNTT(res, cs, c1, c0, st) (reference src/cufhe_gates_gpu.cu:116-133):
// res = c1 if cs encrypts 1, c0 if it encrypts 0. Host copies in and out as
// the reference does; gCMUXNTT works on the device copies only.
inline void gCMUXNTT(cuFHETRLWElvl1& res, cuFHETRGSWNTTlvl1& cs, cuFHETRLWElvl1& c1,
                     cuFHETRLWElvl1& c0, Stream st)
{
    CuCheck(cufhe_cmux_lvl1(res.trlwedevices[0], cs.trgswdevices[0], c1.trlwedevices[0],
                            c0.trlwedevices[0], 1, st.st()),
            "CMUXNTT");
}
/** CMUXNTTkernel (reference src/bootstrap_gpu.cu:2357-2369): res = c0 +
 *  ExtProd(c1 - c0, cs) on device pointers, cs a TRGSW in the loaded
 *  backend's transform domain (TRGSW2NTT's device copy). */
inline void CMUXNTTkernel(uint32_t* const res, const void* const cs, uint32_t* const c1,
                          uint32_t* const c0, hipStream_t st, const int)
{
    CuCheck(cufhe_cmux_lvl1(res, cs, c1, c0, 1, st), "CMUXNTTkernel");
}
inline void CMUXNTT(cuFHETRLWElvl1& res, cuFHETRGSWNTTlvl1& cs, cuFHETRLWElvl1& c1,
                    cuFHETRLWElvl1& c0, Stream st)
{
    CuCheck(cufhe_memcpy_h2d(cs.trgswdevices[0], cs.trgswhost.data(), cs.bytes, st.st()),
            "CMUXNTT H2D");
    CuCheck(cufhe_memcpy_h2d(c1.trlwedevices[0], c1.trlwehost.data(), sizeof(c1.trlwehost),
                             st.st()),
            "CMUXNTT H2D");
    CuCheck(cufhe_memcpy_h2d(c0.trlwedevices[0], c0.trlwehost.data(), sizeof(c0.trlwehost),
                             st.st()),
            "CMUXNTT H2D");
    gCMUXNTT(res, cs, c1, c0, st);
    CuCheck(cufhe_memcpy_d2h(res.trlwehost.data(), res.trlwedevices[0], sizeof(res.trlwehost),
                             st.st()),
            "CMUXNTT D2H");
}

// TRLWE extras (reference src/cufhe_gates_gpu.cu:134-196).
inline void gSampleExtractAndKeySwitch(Ctxt<TFHEpp::lvl0param>& out, const cuFHETRLWElvl1& in,
                                       Stream st)
{
    CuCheck(cufhe_sample_extract_key_switch_lvl10(out.tlwedevices[0], in.trlwedevices[0], 1,
                                                  st.st()),
            "SampleExtractAndKeySwitch");
}
inline void SampleExtractAndKeySwitch(Ctxt<TFHEpp::lvl0param>& out, const cuFHETRLWElvl1& in,
                                      Stream st)
{
    CuCheck(cufhe_memcpy_h2d(in.trlwedevices[0], in.trlwehost.data(), sizeof(in.trlwehost),
                             st.st()),
            "TRLWE H2D");
    gSampleExtractAndKeySwitch(out, in, st);
    CtxtCopyD2H<TFHEpp::lvl0param>(out, st);
}
inline void gGateBootstrappingTLWE2TRLWElvl01NTT(cuFHETRLWElvl1& out,
                                                 Ctxt<TFHEpp::lvl0param>& in, Stream st)
{
    CuCheck(cufhe_bootstrap_to_trlwe_lvl01(out.trlwedevices[0], in.tlwedevices[0],
                                           TFHEpp::lvl1param::μ, 1, st.st()),
            "GateBootstrappingTLWE2TRLWElvl01NTT");
}
inline void GateBootstrappingTLWE2TRLWElvl01NTT(cuFHETRLWElvl1& out,
                                                Ctxt<TFHEpp::lvl0param>& in, Stream st)
{
    CtxtCopyH2D<TFHEpp::lvl0param>(in, st);
    gGateBootstrappingTLWE2TRLWElvl01NTT(out, in, st);
    CuCheck(cufhe_memcpy_d2h(out.trlwehost.data(), out.trlwedevices[0], sizeof(out.trlwehost),
                             st.st()),
            "TRLWE D2H");
}
inline void gRefresh(cuFHETRLWElvl1& out, cuFHETRLWElvl1& in, Stream st)
{
    CuCheck(cufhe_refresh_lvl1(out.trlwedevices[0], in.trlwedevices[0], 1, st.st()), "Refresh");
}
inline void Refresh(cuFHETRLWElvl1& out, cuFHETRLWElvl1& in, Stream st)
{
    CuCheck(cufhe_memcpy_h2d(in.trlwedevices[0], in.trlwehost.data(), sizeof(in.trlwehost),
                             st.st()),
            "TRLWE H2D");
    gRefresh(out, in, st);
    CuCheck(cufhe_memcpy_d2h(out.trlwehost.data(), out.trlwedevices[0], sizeof(out.trlwehost),
                             st.st()),
            "TRLWE D2H");
}

}  // namespace cufhe
