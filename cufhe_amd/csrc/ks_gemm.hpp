// Batched identity key switching as an exact int8 GEMM (ks_gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_common.hpp"

namespace cufhe_amd {

// Key-switching key as four signed byte planes: [row][4 (n0 + 1)] int8,
// row = (switched coefficient, digit, |digit| - 1).
struct KsGemmKey {
    int8_t* planes = nullptr;
    int rows = 0, cols = 0;
    int level = 1;  // domain level: 1 (lvl10) or 2 (lvl20)
    bool subset = false;
};

// Batches below this many jobs use the gather kernel (k_ks_batch). The GEMM
// wins at every batch size measured (16 jobs: 0.08 ms against 0.55 ms for
// the eight-waves-per-job gather), so every plan with a workspace uses it;
// the gather form serves the entry points that take no workspace.
constexpr int kKsGemmMinJobs = 1;

// Builds the byte planes of a device KSK (TFHEpp layout); 0 on success.
int ks_gemm_prepare(KsGemmKey& key, const uint32_t* d_ksk, int domain_level, bool subset,
                    hipStream_t st);
void ks_gemm_release(KsGemmKey& key);
// Device workspace bytes for njobs jobs.
size_t ks_gemm_workspace_bytes(const KsGemmKey& key, int njobs);
// Key-switches njobs jobs (device KsJob array) on st; 0 on success.
int ks_gemm_run(const KsGemmKey& key, const KsJob* d_jobs, int njobs, void* ws, hipStream_t st);

}  // namespace cufhe_amd
