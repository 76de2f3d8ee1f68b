// Error helpers of cuFHEpp's public headers (reference
// include/error_gpu.cuh:28-66) for HIP: CuSafeCall(e) reports a failed HIP
// call with its source location and exits; CuCheckError() does the same for
// the last launch error, and with CUFHE_SAFER defined also waits for the
// device so that asynchronous faults surface at the call.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace cufhe {

inline void CuSafeCall__(hipError_t e, const char* file, const int line)
{
    if (e == hipSuccess) return;
    std::fprintf(stderr, "cufhe: CuSafeCall failed at %s:%d: %s\n", file, line, hipGetErrorString(e));
    std::exit(-1);
}

inline void CuCheckError__(const char* file, const int line)
{
    CuSafeCall__(hipGetLastError(), file, line);
#ifdef CUFHE_SAFER
    CuSafeCall__(hipDeviceSynchronize(), file, line);
#endif
}

}  // namespace cufhe

#define CuSafeCall(e) ::cufhe::CuSafeCall__((e), __FILE__, __LINE__)
#define CuCheckError() ::cufhe::CuCheckError__(__FILE__, __LINE__)
