// Internal engine interface shared by the C ABI's plans (runtime.hip) and the
// serving loop (server.hip): turning a lvl1 gate list into key-switch,
// blind-rotation and combine jobs over a device wire pool, and launching
// them on the loaded lvl1 keys.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cufhe_amd.h"
#include "device_common.hpp"

namespace cufhe_amd {

// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string& msg);

constexpr uint32_t kWireWords1 = lvl1param::n + 1;  // lvl1 TLWE words per wire
constexpr uint32_t kLvl0Stride = 640;                // lvl0 TLWE stride in workspaces

struct Lvl1Jobs {
    std::vector<KsJob> ks;
    std::vector<BrJob> br;
    std::vector<LinJob> lin;
};

// Key switches (= bootstraps) and MUX/NMUX gates of a gate list; fails on an
// unknown op.
int lvl1_count(const cufhe_gate_desc* gates, int n, int* nks, int* nmux);
// Workspace words: one lvl0 TLWE per bootstrap plus two lvl1 TLWEs per MUX.
inline size_t lvl1_workspace_words(int nks, int nmux)
{
    return static_cast<size_t>(nks) * kLvl0Stride + static_cast<size_t>(nmux) * 2 * kWireWords1;
}
// Appends the jobs of gates[0, n) (wire w at d_pool + w * kWireWords1) to
// jobs, using ws (lvl1_workspace_words of this list) for intermediates.
void lvl1_build(const cufhe_gate_desc* gates, int n, uint32_t* d_pool, uint32_t* ws, Lvl1Jobs& jobs);
bool lvl1_loaded();
// Changes whenever lvl1 keys are (re)loaded; 0 when none are loaded.
uint64_t lvl1_key_generation();
// Validated jobs-per-workgroup for nbr bootstraps (c = 0: the default for
// the loaded keys and that batch size; nbr < 0: size unknown), or -1.
int lvl1_pick_c(int c, int nbr);
// Workspace bytes of the GEMM key switch for nks jobs (0: gather path).
size_t lvl1_ks_workspace_bytes(int nks);
// Key switch -> blind rotation + SampleExtract -> combine, on stream st
// (ks_ws: lvl1_ks_workspace_bytes(nks) bytes or null). False if the key
// switch failed to launch.
bool lvl1_launch(const KsJob* ks, int nks, void* ks_ws, const BrJob* br, int nbr,
                 const LinJob* lin, int nlin, int c, hipStream_t st);

}  // namespace cufhe_amd
