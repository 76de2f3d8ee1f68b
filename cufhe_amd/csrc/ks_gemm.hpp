// Batched identity key switching as an exact int8 MFMA product (ks_gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_common.hpp"

namespace cufhe_amd {

// Key-switching key pre-tiled for the int8 MFMA key switch: for each block
// of 32 output words, each 32-deep step of the digit dimension and each of
// the four balanced byte planes, one 1 KiB operand fragment in register
// order. rows = switched coefficients x t x 2^(basebit-1) (the digit rows
// of the TFHEpp layout), cols = 4 (n0 + 1).
struct KsGemmKey {
    int8_t* planes = nullptr;
    int rows = 0, cols = 0;
    int level = 1;  // domain level: 1 (lvl10) or 2 (lvl20)
    bool subset = false;
};

// Batches of at least this many jobs use the MFMA key switch when a
// workspace is given (it wins at every batch size); the gather kernel
// (k_ks_batch) serves the entry points that take no workspace.
constexpr int kKsGemmMinJobs = 1;

// Blocks of 32 jobs per wave in the MFMA kernel (MM x 4 accumulators of
// 32 x 32 int32) and waves per workgroup (sharing one LDS key chunk).
constexpr int kKsMfmaJobBlocks = 2;
constexpr int kKsMfmaWaves = 8;
// 32-deep steps of the digit dimension of the full lvl10 key, the largest
// lvl1 key switch (workspace: 4 bytes per step per job).
constexpr int kKsLvl10StepsMax = 448;
// Steps of the digit dimension per LDS chunk (one workgroup barrier each).
constexpr int kKsMfmaChunk = 4;

// Builds the pre-tiled key from a device KSK (TFHEpp layout); 0 on success.
int ks_gemm_prepare(KsGemmKey& key, const uint32_t* d_ksk, int domain_level, bool subset,
                    hipStream_t st);
void ks_gemm_release(KsGemmKey& key);
// Device workspace bytes for njobs jobs.
size_t ks_gemm_workspace_bytes(const KsGemmKey& key, int njobs);
// Key-switches njobs jobs (device KsJob array) on st; 0 on success.
int ks_gemm_run(const KsGemmKey& key, const KsJob* d_jobs, int njobs, void* ws, hipStream_t st);
// Number of waves the digit dimension is split over (their partial sums are
// added atomically); 0 restores the per-batch choice (about two waves per
// SIMD).
void ks_gemm_set_split(int split);

}  // namespace cufhe_amd
