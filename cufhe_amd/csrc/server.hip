// Serving loop behind cufhe_server_* (include/cufhe_amd.h): continuous
// batching of layered lvl1 Boolean circuits on one GPU.
//
// Data flow. The server owns one device wire pool (lvl1 TLWEs of 1025 words)
// carved by a first-fit range allocator. A submitted circuit waits in a FIFO
// until its wires fit; at admission its inputs are copied into its range and
// every later layer's outputs are written next to them, so ciphertexts stay
// in HBM for the circuit's whole life. Each round the worker thread takes
// the next unscheduled layer of every admitted circuit (in admission order,
// whole layers, up to the round's bootstrap and gate budget), rewrites their
// local wire ids to pool slots, and runs them as one batch: key switch ->
// blind rotation -> combine (engine.hpp). Gates never depend on their
// batch-mates, so every result is bit-identical to running the circuit
// alone.
//
// Pipelining. Rounds are issued on one stream, so layer l + 1 of a circuit
// may be issued while layer l still runs: stream order makes it read the
// finished outputs. Up to two rounds are in flight; each has its own pinned
// and device job arrays and a pinned output staging area, and two events.
// While round k runs, the worker admits circuits, builds round k + 1 and
// issues it; it blocks only when two rounds are in flight or nothing is left
// to issue. A round retires when its end event completes: the outputs of the
// circuits whose last layer it ran are copied from the staging area to the
// caller's buffers, their wire ranges are freed (admitting waiting circuits)
// and the GPU time between its events is added to the busy time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

using namespace cufhe_amd;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator of contiguous wire ranges with coalescing.
class WireArena {
   public:
    void reset(size_t wires)
    {
        free_.clear();
        if (wires) free_[0] = wires;
        used_ = 0;
    }
    bool alloc(size_t n, size_t* base)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *base = it->first;
            const size_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    void release(size_t base, size_t n)
    {
        auto it = free_.emplace(base, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        used_ -= n;
    }
    size_t used() const { return used_; }

   private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct Request {
    uint64_t id = 0;
    int n_inputs = 0, n_gates = 0, n_layers = 0;
    std::vector<cufhe_gate_desc> gates;
    std::vector<int> layer_end;
    std::vector<int> layer_boot;  // bootstraps per layer
    std::vector<uint32_t> outputs;
    uint32_t* inputs = nullptr;  // pinned copy: n_inputs x 1025 words
    uint32_t* host_out = nullptr;
    size_t wires = 0, base = 0;
    int next_layer = 0, rounds = 0;  // next_layer: the next layer not yet issued
    int status = 0;  // 0 running, 1 done, < 0 error code
    double t_submit = 0, t_admit = -1, t_done = -1;
    ~Request() { (void)hipHostFree(inputs); }
};
using ReqPtr = std::shared_ptr<Request>;

// Buffers of one in-flight round.
struct RoundSlot {
    KsJob* d_ks = nullptr;
    BrJob* d_br = nullptr;
    LinJob* d_lin = nullptr;
    KsJob* h_ks = nullptr;
    BrJob* h_br = nullptr;
    LinJob* h_lin = nullptr;
    uint32_t* h_out = nullptr;  // outputs of the circuits finishing in the round
    size_t h_out_wires = 0;
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
};

struct InFlight {
    int slot = 0;
    std::vector<ReqPtr> finishing;
    std::vector<size_t> out_offset;  // staging offset (wires) of each finishing circuit
};

}  // namespace

struct cufhe_server {
    uint32_t* pool = nullptr;
    size_t pool_wires = 0;
    WireArena arena;
    int max_boot = 0, c = 1;
    uint64_t key_generation = 0;  // the lvl1 key load the workspaces were sized for
    hipStream_t st = nullptr;
    // stream-ordered, so shared by the rounds in flight
    uint32_t* d_ws = nullptr;
    void* d_ksws = nullptr;  // MFMA key-switch workspace (max_boot jobs)
    RoundSlot slots[2];
    std::deque<InFlight> inflight;  // worker thread only
    int next_slot = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> waiting;
    std::vector<ReqPtr> active;
    std::unordered_map<uint64_t, ReqPtr> all;
    uint64_t next_id = 1;
    bool stop = false;
    int error = 0;
    std::string error_msg;
    Clock::time_point t0;
    double stats[5] = {0, 0, 0, 0, 0};
    std::thread worker;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void fail_all_locked(cufhe_server* S, int code, const std::string& msg)
{
    S->error = code;
    S->error_msg = msg;
    for (auto& r : S->active) r->status = code;
    for (auto& r : S->waiting) r->status = code;
    S->active.clear();
    S->waiting.clear();
    S->cv_done.notify_all();
}

// Retire the oldest in-flight round once its end event has completed.
// `block`: wait for it. Returns false if it is still running (block false).
bool retire_oldest(cufhe_server* S, bool block)
{
    InFlight& f = S->inflight.front();
    RoundSlot& sl = S->slots[f.slot];
    hipError_t e = block ? hipEventSynchronize(sl.ev_end) : hipEventQuery(sl.ev_end);
    if (e == hipErrorNotReady) return false;
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, sl.ev_start, sl.ev_end);
    if (e == hipSuccess) {
        for (size_t i = 0; i < f.finishing.size(); i++) {
            const ReqPtr& r = f.finishing[i];
            std::memcpy(r->host_out, sl.h_out + f.out_offset[i] * kWireWords1,
                        r->outputs.size() * kWireWords1 * sizeof(uint32_t));
        }
    }
    const double te = S->now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (e != hipSuccess) {
        S->inflight.clear();
        fail_all_locked(S, CUFHE_ERR_HIP, std::string("server round: ") + hipGetErrorString(e));
        return true;
    }
    S->stats[3] += ms * 1e-3;
    for (auto& r : f.finishing) {
        r->status = 1;
        r->t_done = te;
        S->arena.release(r->base, r->wires);
        S->active.erase(std::find(S->active.begin(), S->active.end(), r));
        // keep the timestamps, drop the circuit
        std::vector<cufhe_gate_desc>().swap(r->gates);
        (void)hipHostFree(r->inputs);
        r->inputs = nullptr;
    }
    if (!f.finishing.empty()) S->cv_done.notify_all();
    S->inflight.pop_front();
    return true;
}

// The worker: retire finished rounds, admit, gather the next unscheduled
// layer of every admitted circuit, issue them as one round; at most two
// rounds in flight.
void worker_loop(cufhe_server* S)
{
    constexpr uint32_t W1 = kWireWords1;
    for (;;) {
        while (!S->inflight.empty() && retire_oldest(S, false)) {
        }
        std::vector<ReqPtr> admitted_now, finishing;
        std::vector<cufhe_gate_desc> gates;
        bool block_on_round = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            auto issuable = [&] {
                if (!S->waiting.empty()) return true;
                for (auto& r : S->active)
                    if (r->next_layer < r->n_layers) return true;
                return false;
            };
            if (!issuable()) {
                if (!S->inflight.empty()) {
                    block_on_round = true;
                }
                else {
                    if (S->stop && S->waiting.empty() && S->active.empty()) return;
                    S->cv_work.wait(lk, [&] { return S->stop || !S->waiting.empty(); });
                    continue;
                }
            }
            if (!block_on_round) {
                while (!S->waiting.empty()) {
                    ReqPtr r = S->waiting.front();
                    size_t base = 0;
                    if (!S->arena.alloc(r->wires, &base)) break;
                    r->base = base;
                    r->t_admit = S->now();
                    S->active.push_back(r);
                    admitted_now.push_back(r);
                    S->waiting.pop_front();
                }
                S->stats[4] = std::max(S->stats[4], static_cast<double>(S->arena.used()) * W1 * 4);
                int boots = 0, ng = 0;
                for (auto& r : S->active) {
                    const int l = r->next_layer;
                    if (l >= r->n_layers) continue;  // every layer issued; retiring
                    const int g0 = l == 0 ? 0 : r->layer_end[l - 1], g1 = r->layer_end[l];
                    const int lb = r->layer_boot[l], lg = g1 - g0;
                    if (!gates.empty() && (boots + lb > S->max_boot || ng + lg > S->max_boot)) continue;
                    const uint32_t b = static_cast<uint32_t>(r->base);
                    for (int g = g0; g < g1; g++) {
                        cufhe_gate_desc d = r->gates[g];
                        d.out = b + static_cast<uint32_t>(r->n_inputs + g);
                        d.in0 += b;
                        d.in1 += b;
                        d.inc += b;
                        gates.push_back(d);
                    }
                    boots += lb;
                    ng += lg;
                    r->next_layer++;
                    r->rounds++;
                    if (r->next_layer == r->n_layers) finishing.push_back(r);
                }
                if (!gates.empty()) {
                    S->stats[0] += 1;
                    S->stats[1] += boots;
                    S->stats[2] += ng;
                }
                else if (!S->inflight.empty()) {
                    block_on_round = true;  // waiting circuits need the space of running ones
                }
            }
        }
        if (block_on_round) {
            retire_oldest(S, true);
            continue;
        }
        if (gates.empty()) continue;
        if (lvl1_key_generation() != S->key_generation) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->inflight.clear();
            fail_all_locked(S, CUFHE_ERR_STATE, "lvl1 keys were reloaded while the server ran");
            continue;
        }
        // a free slot: at most two rounds in flight
        if (S->inflight.size() == 2) retire_oldest(S, true);
        if (S->error) continue;
        const int si = S->next_slot;
        S->next_slot ^= 1;
        RoundSlot& sl = S->slots[si];
        InFlight f;
        f.slot = si;
        size_t out_wires = 0;
        for (auto& r : finishing) {
            f.out_offset.push_back(out_wires);
            out_wires += r->outputs.size();
        }
        f.finishing = finishing;
        hipError_t e = hipSuccess;
        if (out_wires > sl.h_out_wires) {  // the slot is idle: grow its staging
            (void)hipHostFree(sl.h_out);
            sl.h_out = nullptr;
            sl.h_out_wires = 0;
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_out), out_wires * W1 * sizeof(uint32_t),
                              hipHostMallocDefault);
            if (e == hipSuccess) sl.h_out_wires = out_wires;
        }
        for (auto& r : admitted_now) {
            if (e != hipSuccess) break;
            e = hipMemcpyAsync(S->pool + r->base * W1, r->inputs,
                               static_cast<size_t>(r->n_inputs) * W1 * sizeof(uint32_t),
                               hipMemcpyHostToDevice, S->st);
        }
        Lvl1Jobs jobs;
        jobs.ks.reserve(S->max_boot);
        jobs.br.reserve(S->max_boot);
        lvl1_build(gates.data(), static_cast<int>(gates.size()), S->pool, S->d_ws, jobs);
        const int nks = static_cast<int>(jobs.ks.size()), nbr = static_cast<int>(jobs.br.size());
        const int nlin = static_cast<int>(jobs.lin.size());
        std::memcpy(sl.h_ks, jobs.ks.data(), nks * sizeof(KsJob));
        std::memcpy(sl.h_br, jobs.br.data(), nbr * sizeof(BrJob));
        std::memcpy(sl.h_lin, jobs.lin.data(), nlin * sizeof(LinJob));
        if (e == hipSuccess) e = hipEventRecord(sl.ev_start, S->st);
        if (e == hipSuccess && nks)
            e = hipMemcpyAsync(sl.d_ks, sl.h_ks, nks * sizeof(KsJob), hipMemcpyHostToDevice, S->st);
        if (e == hipSuccess && nbr)
            e = hipMemcpyAsync(sl.d_br, sl.h_br, nbr * sizeof(BrJob), hipMemcpyHostToDevice, S->st);
        if (e == hipSuccess && nlin)
            e = hipMemcpyAsync(sl.d_lin, sl.h_lin, nlin * sizeof(LinJob), hipMemcpyHostToDevice, S->st);
        if (e == hipSuccess) {
            // c = 0: choose per round (small rounds use one job per workgroup)
            const int c = S->c ? S->c : lvl1_pick_c(0, nbr);
            if (!lvl1_launch(sl.d_ks, nks, S->d_ksws, sl.d_br, nbr, sl.d_lin, nlin, c, S->st))
                e = hipErrorLaunchFailure;
            if (e == hipSuccess) e = hipGetLastError();
        }
        for (size_t i = 0; i < finishing.size() && e == hipSuccess; i++) {
            const ReqPtr& r = finishing[i];
            for (size_t o = 0; o < r->outputs.size() && e == hipSuccess; o++)
                e = hipMemcpyAsync(sl.h_out + (f.out_offset[i] + o) * W1,
                                   S->pool + (r->base + r->outputs[o]) * W1, W1 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, S->st);
        }
        if (e == hipSuccess) e = hipEventRecord(sl.ev_end, S->st);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(S->st);
            std::lock_guard<std::mutex> lk(S->mu);
            S->inflight.clear();
            fail_all_locked(S, CUFHE_ERR_HIP, std::string("server round: ") + hipGetErrorString(e));
            continue;
        }
        S->inflight.push_back(std::move(f));
    }
}

void free_server(cufhe_server* S)
{
    if (S->st) (void)hipStreamSynchronize(S->st);
    (void)hipFree(S->pool);
    (void)hipFree(S->d_ws);
    (void)hipFree(S->d_ksws);
    for (RoundSlot& sl : S->slots) {
        (void)hipFree(sl.d_ks);
        (void)hipFree(sl.d_br);
        (void)hipFree(sl.d_lin);
        (void)hipHostFree(sl.h_ks);
        (void)hipHostFree(sl.h_br);
        (void)hipHostFree(sl.h_lin);
        (void)hipHostFree(sl.h_out);
        if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
        if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
    }
    if (S->st) (void)hipStreamDestroy(S->st);
    delete S;
}

}  // namespace

extern "C" {

int cufhe_server_create(size_t pool_bytes, int max_batch_bootstraps, int jobs_per_group,
                        cufhe_server** server)
{
    if (!server) return set_error(CUFHE_ERR_ARG, "null server handle");
    *server = nullptr;
    if (!lvl1_loaded()) return set_error(CUFHE_ERR_STATE, "lvl1 keys not loaded (cufhe_initialize_lvl01)");
    const int c = jobs_per_group == 0 ? 0 : lvl1_pick_c(jobs_per_group, -1);
    if (c < 0) return set_error(CUFHE_ERR_UNSUPPORTED, "jobs_per_group must be 0, 1, 2, 4, 6 or 8");
    if (max_batch_bootstraps <= 0) max_batch_bootstraps = 8192;
    const size_t wire_bytes = static_cast<size_t>(kWireWords1) * sizeof(uint32_t);
    if (pool_bytes < wire_bytes) return set_error(CUFHE_ERR_ARG, "pool smaller than one ciphertext");
    auto* S = new cufhe_server();
    S->pool_wires = pool_bytes / wire_bytes;
    S->max_boot = max_batch_bootstraps;
    S->c = c;
    const size_t mb = static_cast<size_t>(max_batch_bootstraps);
    hipError_t e = hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&S->pool), S->pool_wires * wire_bytes);
    for (RoundSlot& sl : S->slots) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.d_ks), mb * sizeof(KsJob));
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.d_br), mb * sizeof(BrJob));
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.d_lin), mb * sizeof(LinJob));
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_ks), mb * sizeof(KsJob), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_br), mb * sizeof(BrJob), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_lin), mb * sizeof(LinJob), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreate(&sl.ev_start);
        if (e == hipSuccess) e = hipEventCreate(&sl.ev_end);
    }
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&S->d_ws),
                      lvl1_workspace_words(max_batch_bootstraps, max_batch_bootstraps / 2) *
                          sizeof(uint32_t));
    if (e == hipSuccess && lvl1_ks_workspace_bytes(max_batch_bootstraps))
        e = hipMalloc(&S->d_ksws, lvl1_ks_workspace_bytes(max_batch_bootstraps));
    if (e != hipSuccess) {
        free_server(S);
        return set_error(CUFHE_ERR_NOMEM, std::string("server allocation: ") + hipGetErrorString(e));
    }
    S->arena.reset(S->pool_wires);
    S->key_generation = lvl1_key_generation();
    S->t0 = Clock::now();
    S->worker = std::thread(worker_loop, S);
    *server = S;
    return CUFHE_OK;
}

int cufhe_server_submit(cufhe_server* S, const cufhe_circuit* c, const uint32_t* host_inputs,
                        uint32_t* host_outputs, uint64_t* request_id)
{
    if (!S || !c || !request_id) return set_error(CUFHE_ERR_ARG, "null argument");
    if (c->n_inputs < 0 || c->n_gates <= 0 || c->n_layers <= 0 || c->n_outputs <= 0 || !c->gates ||
        !c->layer_end || !c->outputs || (c->n_inputs && !host_inputs) || !host_outputs)
        return set_error(CUFHE_ERR_ARG, "bad circuit");
    auto r = std::make_shared<Request>();
    r->n_inputs = c->n_inputs;
    r->n_gates = c->n_gates;
    r->n_layers = c->n_layers;
    r->gates.assign(c->gates, c->gates + c->n_gates);
    r->layer_end.assign(c->layer_end, c->layer_end + c->n_layers);
    r->outputs.assign(c->outputs, c->outputs + c->n_outputs);
    r->wires = static_cast<size_t>(c->n_inputs) + c->n_gates;
    r->host_out = host_outputs;
    // validate layers and operands: a gate reads only inputs and earlier layers
    int prev = 0;
    for (int l = 0; l < c->n_layers; l++) {
        const int end = r->layer_end[l];
        if (end <= prev || end > c->n_gates || (l == c->n_layers - 1 && end != c->n_gates))
            return set_error(CUFHE_ERR_ARG, "layer_end must increase to n_gates");
        const uint32_t limit = static_cast<uint32_t>(c->n_inputs + prev);
        int boots = 0;
        for (int g = prev; g < end; g++) {
            cufhe_gate_desc& d = r->gates[g];
            const int op = d.op;
            if (op < 0 || op >= CUFHE_GATE_COUNT) return set_error(CUFHE_ERR_ARG, "unknown gate op");
            const bool mux = op == CUFHE_MUX || op == CUFHE_NMUX;
            const bool unary = op == CUFHE_NOT || op == CUFHE_COPY;
            if (!unary && d.in1 >= limit) return set_error(CUFHE_ERR_ARG, "operand from the same or a later layer");
            if (d.in0 >= limit || (mux && d.inc >= limit))
                return set_error(CUFHE_ERR_ARG, "operand from the same or a later layer");
            if (unary) d.in1 = 0;
            if (!mux) d.inc = 0;
            boots += mux ? 2 : (unary ? 0 : 1);
        }
        if (boots > S->max_boot || end - prev > S->max_boot)
            return set_error(CUFHE_ERR_SIZE, "a layer exceeds the server's batch budget");
        r->layer_boot.push_back(boots);
        prev = end;
    }
    for (uint32_t o : r->outputs)
        if (o >= r->wires) return set_error(CUFHE_ERR_ARG, "output wire out of range");
    if (r->wires > S->pool_wires) return set_error(CUFHE_ERR_SIZE, "circuit larger than the wire pool");
    if (c->n_inputs) {
        const size_t bytes = static_cast<size_t>(c->n_inputs) * kWireWords1 * sizeof(uint32_t);
        if (hipHostMalloc(reinterpret_cast<void**>(&r->inputs), bytes, hipHostMallocDefault) != hipSuccess) {
            r->inputs = nullptr;
            return set_error(CUFHE_ERR_NOMEM, "pinned input staging allocation failed");
        }
        std::memcpy(r->inputs, host_inputs, bytes);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->error) return set_error(S->error, S->error_msg);
    if (S->stop) return set_error(CUFHE_ERR_STATE, "server is shutting down");
    r->id = S->next_id++;
    r->t_submit = S->now();
    S->all[r->id] = r;
    S->waiting.push_back(r);
    *request_id = r->id;
    S->cv_work.notify_one();
    return CUFHE_OK;
}

int cufhe_server_wait(cufhe_server* S, uint64_t id, int timeout_ms)
{
    if (!S) return set_error(CUFHE_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return set_error(CUFHE_ERR_ARG, "unknown request");
    ReqPtr r = it->second;
    auto done = [&] { return r->status != 0; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, done);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return 0;
    if (r->status < 0) return set_error(r->status, S->error_msg);
    return 1;
}

int cufhe_server_request_info(cufhe_server* S, uint64_t id, double* times, int* rounds)
{
    if (!S) return set_error(CUFHE_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return set_error(CUFHE_ERR_ARG, "unknown request");
    if (times) {
        times[0] = it->second->t_submit;
        times[1] = it->second->t_admit;
        times[2] = it->second->t_done;
    }
    if (rounds) *rounds = it->second->rounds;
    return CUFHE_OK;
}

int cufhe_server_stats(cufhe_server* S, double* stats)
{
    if (!S || !stats) return set_error(CUFHE_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    for (int i = 0; i < 5; i++) stats[i] = S->stats[i];
    return CUFHE_OK;
}

int cufhe_server_destroy(cufhe_server* S)
{
    if (!S) return CUFHE_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    const int rc = S->error;
    free_server(S);
    return rc;
}

}  // extern "C"
