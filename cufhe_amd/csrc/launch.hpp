// Host-side launch wrappers of the HIP kernels (defined in kernels_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

namespace cufhe_amd {

struct cplx;
template <class P>
struct FourierKeys;
struct NttKeys;
struct NttTables;
struct Ntt64Keys;
struct Ntt64Tables;

template <class P>
void launch_to_fourier(cplx* out, const void* in, int npolys, bool integer,
                       const FourierKeys<P>& K, hipStream_t st);
template <class P>
void launch_from_fourier(double* out, const cplx* in, int npolys, const FourierKeys<P>& K,
                         hipStream_t st);
template <class P>
void launch_xai_poly(int32_t* out, hipStream_t st);
template <class P, bool KB>
void launch_br_fft(const BrJob* jobs, int njobs, int c, typename P::T mu,
                   const FourierKeys<P>& K, hipStream_t st);
void launch_gate_lvl1_fft(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                          const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                          uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                          const FourierKeys<lvl1param>& K, hipStream_t st);
// lvl0 API gate (BR-IKS at lvl1). tv_mu: the test vector's amplitude (the
// gate's mu; the reference Bootstrap() passes its own); MUX/NMUX combine
// with lvl1param's mu.
void launch_gate_lvl01_fft(bool kb, bool mux, bool subset, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
                           const FourierKeys<lvl1param>& K, hipStream_t st);
void launch_se_ks_lvl10(bool subset, uint32_t* out, const uint32_t* trlwe, int n,
                        const uint32_t* ksk, hipStream_t st);
void launch_gate_lvl02_fft(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                           const FourierKeys<lvl2param>& K, hipStream_t st);
void launch_ks_batch(int domain_level, bool subset, const KsJob* jobs, int njobs,
                     const uint32_t* ksk, hipStream_t st);
void launch_linear(const LinJob* jobs, int njobs, hipStream_t st);
void launch_linear_one(const LinJob& job, hipStream_t st);
template <class P>
void launch_cmux_fft(typename P::T* trlwe, const cplx* trgsw, uint32_t abar,
                     const FourierKeys<P>& K, hipStream_t st);

// Annihilation key switching of `batch` TRLWEs (kernels_annihilate.hip);
// strides in torus words, ahk: the key's nbit x l x 2 spectra.
template <class P>
void launch_annihilate_fft(typename P::T* out, size_t out_stride, const typename P::T* in,
                           size_t in_stride, int batch, const cplx* ahk, const FourierKeys<P>& K,
                           hipStream_t st);
// lvl2 on the Goldilocks NTT backend; ahk: the key's nbit x l x 2 transforms
// (launch_to_ntt64 layout).
void launch_annihilate_ntt64(uint64_t* out, size_t out_stride, const uint64_t* in,
                             size_t in_stride, int batch, const uint64_t* ahk,
                             const Ntt64Tables& tb, hipStream_t st);

// NTT backend (lvl1): mode 0 plain CMUX, 1 key bundle, 2 block-binary.
void launch_cmux_batch_fft(const CmuxBatch& batch, int n, const FourierKeys<lvl1param>& K,
                           hipStream_t st);
void launch_cmux_batch_ntt(const CmuxBatch& batch, int n, const NttKeys& K, hipStream_t st);
void launch_to_ntt(uint32_t* out, const uint32_t* in, int npolys, const NttTables& tb,
                   hipStream_t st);
void launch_ntt_natural(bool inverse, uint32_t* out, const uint32_t* in, int npolys,
                        const NttTables& tb, hipStream_t st);
void launch_br_ntt(int mode, const BrJob* jobs, int njobs, int c, uint32_t mu, const NttKeys& K,
                   hipStream_t st);
void launch_gate_lvl01_ntt(int mode, bool subset, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, uint32_t tv_mu, const uint32_t* ksk,
                           const NttKeys& K, hipStream_t st);
void launch_gate_lvl1_ntt(int mode, bool subset, bool mux, uint32_t* out, const uint32_t* in0,
                          const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                          uint32_t offset, int32_t mux_sign, const uint32_t* ksk, const NttKeys& K,
                          hipStream_t st);

// NTT backend (lvl2, Goldilocks prime): plain CMUX or key bundle.
void launch_to_ntt64(uint64_t* out, const uint64_t* in, int npolys, const Ntt64Tables& tb,
                     hipStream_t st);
void launch_ntt64_natural(bool inverse, uint64_t* out, const uint64_t* in, int npolys,
                          const Ntt64Tables& tb, hipStream_t st);
void launch_br_ntt64(bool kb, const BrJob* jobs, int njobs, uint64_t mu, const Ntt64Keys& K,
                     hipStream_t st);
void launch_gate_lvl02_ntt(bool kb, bool mux, uint32_t* out, const uint32_t* in0,
                           const uint32_t* in1, const uint32_t* inc, int32_t ca, int32_t cb,
                           uint32_t offset, int32_t mux_sign, const uint32_t* ksk,
                           const Ntt64Keys& K, hipStream_t st);

}  // namespace cufhe_amd
