/*
 * C ABI of the MI355X gate-bootstrapping engine: the drop-in boundary for
 * cuFHEpp's gate path. Plain pointers, sizes and opaque stream handles only.
 *
 * Ciphertext layouts (TFHEpp, see include/tfhe_params.hpp):
 *   lvl1 TLWE: 1025 x uint32 (a[0..1023], b), lvl0 TLWE: 637 x uint32,
 *   lvl2 TLWE: 2049 x uint64.
 * Device pointers come from cufhe_device_alloc (or any hipMalloc). A stream
 * handle is a hipStream_t passed as void* (NULL = default stream).
 * Every entry point returns CUFHE_OK (0) or a negative CUFHE_ERR_* code;
 * cufhe_last_error() describes the last failure of the calling thread. The
 * reference prints and calls exit(-1) instead (include/error_gpu.cuh:30-66);
 * the C++ layer (include/cufhe_gpu.hpp) restores that behaviour.
 *
 * Reference interfaces each entry point replaces are cited per function
 * (paths relative to the cuFHEpp source tree).
 */
#ifndef CUFHE_AMD_H
#define CUFHE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUFHE_OK 0
#define CUFHE_ERR_ARG (-1)         /* bad argument (null pointer, bad op) */
#define CUFHE_ERR_SIZE (-2)        /* buffer size does not match the layout */
#define CUFHE_ERR_STATE (-3)       /* not initialised / keys not loaded */
#define CUFHE_ERR_HIP (-4)         /* HIP runtime error */
#define CUFHE_ERR_UNSUPPORTED (-5) /* configuration not provided */
#define CUFHE_ERR_NOMEM (-6)       /* device allocation failed */
#define CUFHE_ERR_NO_DEVICE (-7)   /* no GPU visible */

/* Gate codes (reference include/cufhe_gpu.cuh:249-304 gate set). */
typedef enum {
    CUFHE_NAND = 0,
    CUFHE_OR = 1,
    CUFHE_ORYN = 2,
    CUFHE_ORNY = 3,
    CUFHE_AND = 4,
    CUFHE_ANDYN = 5,
    CUFHE_ANDNY = 6,
    CUFHE_NOR = 7,
    CUFHE_XOR = 8,
    CUFHE_XNOR = 9,
    CUFHE_MUX = 10,  /* out = inc ? in1 : in0 */
    CUFHE_NMUX = 11, /* out = !(inc ? in1 : in0) */
    CUFHE_NOT = 12,
    CUFHE_COPY = 13,
    CUFHE_GATE_COUNT = 14
} cufhe_gate_t;

/* One gate of a batch: wire indices into a device pool of TLWEs. */
typedef struct {
    int32_t op;
    uint32_t out;
    uint32_t in0;
    uint32_t in1;
    uint32_t inc; /* MUX/NMUX selector; ignored otherwise */
} cufhe_gate_desc;

typedef struct cufhe_plan cufhe_plan;

/* ---- version / errors --------------------------------------------------- */
const char* cufhe_version(void);
const char* cufhe_last_error(void);

/* ---- setup (reference include/cufhe_gpu.cuh:52-70,
 *      src/cufhe_gates_gpu.cu:38-103) ---------------------------------- */

/* SetGPUNum: only 1 is supported (one process per GPU). */
int cufhe_set_gpu_num(int gpu_num);
/* Initialize(): select the device and build the transform tables. */
int cufhe_initialize(void);
/* Initialize(ek) for the lvl1 gate API: bk is TFHEpp's polynomial-domain
 * BootstrappingKey<lvl01param> (tfhe_bk_words(1, key_bundle) words), ksk is
 * KeySwitchingKey<lvl10param> (tfhe_ksk_words(1) words). The key is
 * converted to the Fourier domain on the GPU. */
int cufhe_initialize_lvl01(const uint32_t* bk, size_t bk_words, int key_bundle,
                           const uint32_t* ksk, size_t ksk_words);

/* Backends and blind-rotation variants of the lvl1 keys (reference build
 * options USE_FFT, USE_KEY_BUNDLE, USE_BLOCK_BINARY, USE_SUBSET_KEY in
 * CMakeLists.txt:18-48). FFT: FP64 negacyclic FFNT; NTT: exact arithmetic
 * mod P = 2147473409 (include/ntt_small_modulus.cuh:38-74). Block-binary
 * needs the NTT backend and a block-binary lvl0 key (tfhe_secret_key_gen). */
#define CUFHE_BACKEND_FFT 0
#define CUFHE_BACKEND_NTT 1
#define CUFHE_BR_PLAIN 0
#define CUFHE_BR_KEY_BUNDLE 1
#define CUFHE_BR_BLOCK_BINARY 2
/* Initialize(ek) with an explicit backend and variant. bk: polynomial-domain
 * BootstrappingKey<lvl01param> (key-bundle layout for CUFHE_BR_KEY_BUNDLE,
 * one TRGSW per lvl0 key bit otherwise); ksk: KeySwitchingKey<lvl10param>,
 * or the subset key (rows j >= n0 only, tfhe_ksk_subset_words() words) when
 * subset_ksk != 0 (reference emplacesubiksk, test/test_block_binary_gpu.cc
 * :153-156). cufhe_initialize_lvl01(bk, w, kb, ksk, w2) is
 * cufhe_initialize_lvl01_ex(bk, w, CUFHE_BACKEND_FFT, kb, ksk, w2, 0). */
int cufhe_initialize_lvl01_ex(const uint32_t* bk, size_t bk_words, int backend, int variant,
                              const uint32_t* ksk, size_t ksk_words, int subset_ksk);
/* Initialize_lvl02(ek, sk): BootstrappingKey<lvl02param> (uint64 words) and
 * KeySwitchingKey<lvl20param>. */
int cufhe_initialize_lvl02(const uint64_t* bk, size_t bk_words, int key_bundle,
                           const uint32_t* ksk, size_t ksk_words);
/* Initialize_lvl02 with an explicit backend for the lvl2 blind rotation:
 * CUFHE_BACKEND_FFT (FP64 FFNT) or CUFHE_BACKEND_NTT (exact arithmetic mod the
 * Goldilocks prime, reference USE_NTT with lvl2 in CMakeLists.txt:18-48,
 * src/bootstrap_gpu.cu:213-340, :356-366); key_bundle selects the key layout as in
 * cufhe_initialize_lvl02, which is this call with CUFHE_BACKEND_FFT. */
int cufhe_initialize_lvl02_ex(const uint64_t* bk, size_t bk_words, int backend, int key_bundle,
                              const uint32_t* ksk, size_t ksk_words);
/* One key of a level at a time, the other kept (reference
 * include/bootstrap_gpu.cuh:43-51 BootstrappingKeyToNTT /
 * BootstrappingKeyBundleToNTT / DeleteBootstrappingKeyNTT[_lvl02],
 * include/keyswitch_gpu.cuh:258-271 KeySwitchingKeyToDevice[_lvl20] /
 * DeleteKeySwitchingKey[_lvl20]). level 1: bk is BootstrappingKey<lvl01param>
 * (uint32 words; variant CUFHE_BR_PLAIN, CUFHE_BR_KEY_BUNDLE, or
 * CUFHE_BR_BLOCK_BINARY on the NTT backend) and ksk KeySwitchingKey<lvl10param>
 * (or the subset key with subset != 0); level 2: bk is
 * BootstrappingKey<lvl02param> (uint64 words; CUFHE_BR_PLAIN or
 * CUFHE_BR_KEY_BUNDLE) and ksk KeySwitchingKey<lvl20param>. A level's gates
 * run once both of its keys are loaded; every load or delete counts as a
 * key change for plans and servers built before. */
int cufhe_bk_load(int level, const void* bk, size_t words, int backend, int variant);
int cufhe_bk_delete(int level);
int cufhe_ksk_load(int level, const uint32_t* ksk, size_t words, int subset);
int cufhe_ksk_delete(int level);
/* CleanUp() / CleanUp_lvl02(). */
int cufhe_cleanup(void);
int cufhe_cleanup_lvl02(void);
/* Synchronize(): wait for all work on the device. */
int cufhe_synchronize(void);

/* ---- streams (reference include/cufhe_gpu.cuh:162-201 Stream,
 *      src/cufhe_gates_gpu.cu:104-114 StreamQuery) ------------------------ */
int cufhe_stream_create(void** stream);
int cufhe_stream_destroy(void* stream);
/* 1 when all work on the stream is complete, 0 when work is pending. */
int cufhe_stream_query(void* stream);
int cufhe_stream_synchronize(void* stream);
/* HIP events for timing kernels on a stream (measurement plumbing; the
 * reference times with cudaEvent directly, e.g. test/test_perf.cc).
 * cufhe_event_elapsed_ms waits for `end` and returns end - start in ms. */
int cufhe_event_create(void** event);
int cufhe_event_destroy(void* event);
int cufhe_event_record(void* event, void* stream);
int cufhe_event_elapsed_ms(void* start, void* end, float* ms);

/* ---- memory (reference Ctxt ctor/dtor include/cufhe_gpu.cuh:78-121 and
 *      CtxtCopyH2D/D2H :203-217) --------------------------------------- */
int cufhe_device_alloc(void** ptr, size_t bytes);
int cufhe_device_free(void* ptr);
int cufhe_host_alloc(void** ptr, size_t bytes); /* pinned */
int cufhe_host_free(void* ptr);
int cufhe_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int cufhe_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int cufhe_memset_d(void* dst, int value, size_t bytes, void* stream);
int cufhe_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);

/* ---- plain per-gate form (reference src/cufhe_gates_gpu.cu:514-537 g*
 *      gates; src/bootstrap_gpu.cu:2416-3884 XxxBootstrap launchers) ----- */

/* lvl1 API, IKS-BR order: one gate per launch on `stream`. All pointers are
 * device lvl1 TLWEs; in1 is unused for NOT/COPY, inc only for MUX/NMUX.
 * out may alias any input. */
int cufhe_gate_lvl1(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                    const uint32_t* inc, void* stream);
/* lvl0 API with the lvl1 keys, BR-IKS order (reference gNand<lvl0param> ...
 * gNMux<lvl0param>, src/cufhe_gates_gpu.cu:198-511 -> bootstrap_gpu.cu
 * :1557-1672): rotation of the pre-added lvl0 inputs at lvl1, SampleExtract,
 * key switch lvl1 -> lvl0 (the subset key when loaded so). 637-word TLWEs. */
int cufhe_gate_lvl0(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                    const uint32_t* inc, void* stream);
/* lvl02 API, BR-IKS order: lvl0 TLWEs in and out, bootstrapped at lvl2
 * (reference Nand_lvl02 ... NMux_lvl02, src/cufhe_gates_gpu.cu:826-977). */
int cufhe_gate_lvl02(int op, uint32_t* out, const uint32_t* in0, const uint32_t* in1,
                     const uint32_t* inc, void* stream);

/* ---- batched form (no reference counterpart: cuFHEpp runs one gate per
 *      stream; this is the batch-amortised path behind the same gates) --- */

/* Plan a batch of n gates over a device pool of lvl1 TLWEs (wire w at
 * d_pool + 1025 w). Outputs of a batch must not be read by the same batch.
 * jobs_per_group: bootstraps per workgroup (1, 2, 4, 6, 8; 0 = default).
 * With the key-bundle keys, 6 and 8 are the kernels that share each key chunk
 * among the workgroup's jobs through LDS (6: two waves per job, 8: one wave
 * per job with the jobs spread over all CUs, the default); with the NTT
 * keys, 6 is the shared-key NTT kernel. use_graph != 0 captures the launches in a HIP graph
 * on the first run and replays it. */
int cufhe_plan_create_lvl1(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                           int jobs_per_group, int use_graph, cufhe_plan** plan);
/* Plan a batch of lvl02 gates (BR-IKS order) over a device pool of lvl0
 * TLWEs (wire w at d_pool + 637 w); same semantics as the lvl02 gates. */
int cufhe_plan_create_lvl02(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                            int use_graph, cufhe_plan** plan);
/* A plan is bound to the keys loaded when it was created (its kernels, job
 * arrays and key-switch workspace are sized for them): after any
 * cufhe_initialize_lvl01* / _lvl02 of its level, cufhe_plan_run and
 * cufhe_plan_run_timed return CUFHE_ERR_STATE without launching; create
 * the plan again. */
int cufhe_plan_run(cufhe_plan* plan, void* stream);
int cufhe_plan_destroy(cufhe_plan* plan);
/* Bootstraps (blind rotations) a plan performs per run. */
int cufhe_plan_bootstraps(const cufhe_plan* plan);
/* Jobs per workgroup the plan launches with (after autotuning). */
int cufhe_plan_jobs_per_group(const cufhe_plan* plan);
/* Run the plan eagerly with HIP events around each kernel on `stream`,
 * wait for it, and return the kernel durations in ms:
 * kernel_ms[0] = key switching, [1] = blind rotation, [2] = combine. */
int cufhe_plan_run_timed(cufhe_plan* plan, void* stream, float* kernel_ms);
/* Autotuner of the batched rotation, no reference counterpart: times every
 * rotation kernel / jobs-per-workgroup choice c valid for the loaded lvl1
 * keys (1, 2, 4: two waves per job; 6: key chunks shared by six jobs
 * through LDS; 8 (key bundle): one wave per job, key shared by eight) on
 * nbootstraps synthetic jobs and records the fastest for that batch size.
 * Plans, one-shot batches and servers created with jobs_per_group 0 then use
 * the recorded choice of the smallest tuned size >= their batch (else the
 * largest tuned size), and the built-in choice when nothing is tuned; a key
 * load clears the record. best (optional) receives the chosen c;
 * times_ms (optional, 9 floats) the rotation ms indexed by c (0 = not
 * valid). Synchronous; allocates scratch for the run. */
int cufhe_autotune_lvl1(int nbootstraps, int* best, float* times_ms, void* stream);
/* Tuning knob of the batched (int8 MFMA) key switch, no reference
 * counterpart: the number of waves the digit dimension of each output tile
 * is split over, their partial sums added atomically (1-64); 0 restores the
 * per-batch choice (about two waves per SIMD). Applies to launches made
 * after the call, and to graphs captured after it. */
int cufhe_set_key_switch_split(int split);
/* Tuning knob of the batched lvl2 FFT rotation, no reference counterpart:
 * lvl2 jobs per workgroup (1 or 2; default 1). Applies to launches made
 * after the call, and to graphs captured after it. */
int cufhe_set_lvl2_jobs_per_group(int c);
/* Convenience: plan + run + destroy (allocates per call). */
int cufhe_gate_batch_lvl1(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                          int jobs_per_group, void* stream);

/* ---- serving loop (no reference counterpart) -----------------------------
 * A server owns a device wire pool of pool_bytes and a thread that runs
 * continuous batching over submitted lvl1 Boolean circuits: a circuit is
 * admitted (FIFO) once its wires fit in the pool, its inputs are copied in
 * and stay resident until it completes; every round launches ONE batch made
 * of the next layer of every admitted circuit (up to max_batch_bootstraps
 * bootstraps; a layer is never split), and circuits that arrive meanwhile
 * join the following round. Results are bit-identical to running each
 * circuit alone (every gate is computed independently of its batch).
 * Uses the lvl1 keys loaded when the server was created: after a reload
 * every pending and later request fails with CUFHE_ERR_STATE (create a new
 * server). */
typedef struct cufhe_server cufhe_server;
typedef struct {
    int n_inputs;                  /* wires [0, n_inputs) are the inputs */
    int n_gates;                   /* gate g writes wire n_inputs + g ('out' ignored) */
    const cufhe_gate_desc* gates;  /* in topological (layer) order */
    int n_layers;                  /* gates [layer_end[l-1], layer_end[l]) form layer l; */
    const int* layer_end;          /*   a gate reads only inputs and earlier layers */
    int n_outputs;
    const uint32_t* outputs;       /* output wire ids */
} cufhe_circuit;
int cufhe_server_create(size_t pool_bytes, int max_batch_bootstraps, int jobs_per_group,
                        cufhe_server** server);
/* Copies the circuit and its n_inputs x 1025-word host inputs; host_outputs
 * (n_outputs x 1025 words) is written when the request completes and must
 * stay valid until cufhe_server_wait reports it done. */
int cufhe_server_submit(cufhe_server* server, const cufhe_circuit* circuit,
                        const uint32_t* host_inputs, uint32_t* host_outputs, uint64_t* request_id);
/* 1 = done, 0 = still running after timeout_ms (< 0: wait forever), < 0 error. */
int cufhe_server_wait(cufhe_server* server, uint64_t request_id, int timeout_ms);
/* times[3] = submit, admit, done in seconds on the server's clock; rounds =
 * batches the request took part in. */
int cufhe_server_request_info(cufhe_server* server, uint64_t request_id, double* times,
                              int* rounds);
/* stats[5] = rounds, bootstraps, gates, GPU-busy seconds, peak pool bytes in use. */
int cufhe_server_stats(cufhe_server* server, double* stats);
/* Finishes all submitted work, then frees the server. */
int cufhe_server_destroy(cufhe_server* server);

/* The same for lvl0 wires with the lvl1 keys (BR-IKS at lvl1, as
 * cufhe_gate_lvl0): wire w at d_pool + 637 w. */
int cufhe_plan_create_lvl01(const cufhe_gate_desc* gates, int n, uint32_t* d_pool,
                            int jobs_per_group, int use_graph, cufhe_plan** plan);

/* ---- components ----------------------------------------------------------
 * IdentityKeySwitch lvl1 -> lvl0 of n extracted TLWEs (reference
 * include/keyswitch_gpu.cuh:128-187 KeySwitchFromTLWE<lvl10param>).
 * in: n x 1025, out: n x 637 words. */
int cufhe_identity_key_switch_lvl10(uint32_t* d_out, const uint32_t* d_in, int n, void* stream);
/* The same key switch for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference. It calls no
 * hipMalloc/hipFree and no allocating library routine: the job descriptors
 * and the int8 MFMA key switch's digit codes live in a fixed device array of
 * the library (7.5 MB per GPU, held until the process ends), so n is at most
 * 4096 (larger n returns CUFHE_ERR_ARG without launching). Asynchronous on
 * `stream`; the output is bit-identical to cufhe_identity_key_switch_lvl10.
 * Calls that overlap on one GPU (concurrent streams or threads) share that
 * array and race: serialise them. */
int cufhe_identity_key_switch_lvl10_noalloc(uint32_t* d_out, const uint32_t* d_in, int n,
                                            void* stream);
/* IdentityKeySwitch lvl2 -> lvl0 with the lvl02 keys (reference
 * include/keyswitch_gpu.cuh:128-187 KeySwitchFromTLWE<lvl20param>, the body
 * rounded from Torus64 to Torus32 as KeySwitchTorusConvert :33-50; its GPU
 * diagnostic test/test_keyswitch_diag.cu:31-127). in: n x 2049 uint64,
 * out: n x 637 uint32 words. */
int cufhe_identity_key_switch_lvl20(uint32_t* d_out, const uint64_t* d_in, int n, void* stream);
/* Bootstrap(out, in, mu, st, gpu) (reference include/bootstrap_gpu.cuh:61-64,
 * src/bootstrap_gpu.cu:1045-1058 and :2346-2355): blind rotation of ONE lvl0
 * TLWE at lvl1 with test vector +-mu, SampleExtract and key switch back to
 * lvl0 (the subset key when loaded so): 637 words in and out, one launch on
 * `stream`, asynchronous like the reference. out may alias in. (The
 * reference launches its kernel with no dynamic LDS and always the plain
 * rotation; here the loaded keys' rotation runs.) */
int cufhe_bootstrap(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, void* stream);
/* Blind rotation of n lvl0 TLWEs with test vector +-mu and sample
 * extraction to lvl1, without the key switch that Bootstrap() adds (the
 * rotation + SampleExtract half of the reference's BR-IKS gates,
 * src/bootstrap_gpu.cu:1248-1273). in: n x 637, out: n x 1025 words. */
int cufhe_bootstrap_lvl01(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, int n,
                          int jobs_per_group, void* stream);
/* TRLWE extras (reference src/cufhe_gates_gpu.cu:134-196): n lvl1 TRLWEs
 * (2 x 1024 words: a then b) -> SampleExtract + key switch -> n lvl0 TLWEs
 * (SampleExtractAndKeySwitch / SEIandKS, src/keyswitch_gpu.cu:61-68); n lvl0
 * TLWEs -> rotated test-vector TRLWEs without extraction
 * (GateBootstrappingTLWE2TRLWElvl01NTT / BootstrapTLWE2TRLWE,
 * src/bootstrap_gpu.cu:2371-2388); TRLWE -> SE + key switch -> rotation ->
 * TRLWE (Refresh / SEIandBootstrap2TRLWE, :2390-2414). */
int cufhe_sample_extract_key_switch_lvl10(uint32_t* d_out, const uint32_t* d_trlwe, int n,
                                          void* stream);
int cufhe_bootstrap_to_trlwe_lvl01(uint32_t* d_out, const uint32_t* d_in, uint32_t mu, int n,
                                   void* stream);
int cufhe_refresh_lvl1(uint32_t* d_out, const uint32_t* d_in, int n, void* stream);
/* The same at lvl2 with the lvl02 keys (reference src/bootstrap_gpu.cu
 * :1248-1273, the rotation + SampleExtract half of the BR-IKS gate): in
 * n x 637 uint32 lvl0 TLWEs, out n x 2049 uint64 lvl2 TLWEs. */
int cufhe_bootstrap_lvl02(uint64_t* d_out, const uint32_t* d_in, uint64_t mu, int n,
                          void* stream);

/* FFNT of count real polynomials of degree N (1024 or 2048). integer != 0:
 * int32 coefficients used as is; else torus words (uint32 for N=1024,
 * uint64 for N=2048) scaled by 2^-w. Spectra: count x N/2 x (double re, im)
 * in the engine's bin order (reference src/bootstrap_gpu.cu:59-112). */
int cufhe_ffnt_forward(int N, const void* d_poly, int integer, void* d_spectra, int count,
                       void* stream);
/* Inverse FFNT incl. 1/(N/2): spectra -> count x N doubles. */
int cufhe_ffnt_inverse(int N, const void* d_spectra, double* d_poly, int count, void* stream);
/* Negacyclic NTT mod P = 2147473409 of count polynomials of N = 1024
 * residues in [0, P): out[k] = sum_i in[i] psi^(i (2k+1)) mod P in natural
 * bin order (psi: the backend's primitive 2048-th root, reported by
 * cufhe_ntt_root()); inverse != 0 maps such spectra back (incl. N^-1)
 * (reference include/ntt_small_modulus.cuh:394-509). */
int cufhe_ntt_transform(const uint32_t* d_in, uint32_t* d_out, int count, int inverse,
                        void* stream);
uint32_t cufhe_ntt_root(void);
/* Negacyclic NTT mod the Goldilocks prime P = 2^64 - 2^32 + 1 of count
 * polynomials of N = 2048 residues in [0, P), the lvl2 NTT backend's
 * transform: out[k] = sum_i in[i] psi^(i (2k+1)) mod P in natural bin order
 * (psi: primitive 4096-th root, reported by cufhe_ntt64_root()); inverse != 0
 * maps such spectra back (incl. N^-1) (reference include/ntt_small_modulus.cuh
 * :60-74, :154-193, src/ntt_small_modulus.cu:85-113). */
int cufhe_ntt64_transform(const uint64_t* d_in, uint64_t* d_out, int count, int inverse,
                          void* stream);
uint64_t cufhe_ntt64_root(void);
/* ---- CMUX / TRGSW API (lvl1, the loaded lvl1 backend) ---------------------
 * TRGSW2NTT (reference src/bootstrap_gpu.cu:165-186, include/cufhe_gpu.cuh:
 * 219): count lvl1 TRGSWs in the polynomial domain (TFHEpp layout, 4 rows x
 * 2 polynomials x 1024 words each) to the transform domain of the loaded
 * lvl1 keys' backend: FFT spectra (65,536 B per TRGSW) or NTT residues
 * (32,768 B), the bootstrapping key's storage form.
 * cufhe_trgsw_transform_bytes_lvl1() is that size (0 before keys load). */
size_t cufhe_trgsw_transform_bytes_lvl1(void);
int cufhe_trgsw_to_transform_lvl1(const uint32_t* d_trgsw, void* d_out, int count, void* stream);
/* CMUXNTT (reference src/cufhe_gates_gpu.cu:116-133, launcher
 * src/bootstrap_gpu.cu:2357-2369, kernels :700-813 FFT and :925-1041 NTT):
 * res = c0 + ExtProd(c1 - c0, cs), i.e. c1 if cs encrypts 1 and c0 if it
 * encrypts 0, for n independent CMUXes in one launch: cs is n transformed
 * TRGSWs, c1 / c0 / res n lvl1 TRLWEs of 2 x 1024 words (a then b). res may
 * alias c0 or c1. Asynchronous on `stream`. */
int cufhe_cmux_lvl1(uint32_t* d_res, const void* d_cs, const uint32_t* d_c1, const uint32_t* d_c0,
                    int n, void* stream);
/* The same with one device descriptor per CMUX (any pointers, e.g. a CMUX
 * tree level whose selectors repeat); the descriptor array is in device
 * memory. */
typedef struct {
    uint32_t* res;
    const void* cs;
    const uint32_t* c1;
    const uint32_t* c0;
} cufhe_cmux_desc;
int cufhe_cmux_batch_lvl1(const cufhe_cmux_desc* d_descs, int n, void* stream);

/* TRGSW (polynomial domain, TFHEpp layout) -> spectra, count TRGSWs. */
int cufhe_trgsw_to_fourier(int N, const void* d_trgsw, void* d_spectra, int count,
                           void* stream);
/* One CMUX accumulate step acc += ExtProd((X^abar - 1) acc, C) on a device
 * TRLWE (k+1 polys) with C in the Fourier domain (reference
 * include/gatebootstrapping_gpu.cuh:69-221 Accumulate). */
int cufhe_cmux_accumulate(int N, void* d_trlwe, const void* d_trgsw_spectra, uint32_t abar,
                          void* stream);

/* ---- Annihilation key switching (FFT backend; level 1: N = 1024, uint32
 * words; level 2: N = 2048, uint64 words) -----------------------------------
 * Key upload (reference include/annihilate_gpu.cuh:33-38,
 * src/annihilate_gpu.cu:631-677 AnnihilateKeyPolynomialToDevice): the key in
 * the polynomial domain as tfhe_annihilate_key_gen writes it (nbit x l x 2
 * polynomials, tfhe_annihilate_key_words(level) words, host memory) is
 * converted to spectra on the device; a new load replaces the old key (it
 * first waits for all work on the device, so key switches in flight finish
 * with the old key). */
int cufhe_annihilate_key_load(int level, const void* ahk, size_t words);
/* The same with the backend: CUFHE_BACKEND_FFT (both levels) or, at level 2,
 * CUFHE_BACKEND_NTT (the Goldilocks prime: the reference's USE_FFT=OFF
 * arithmetic, src/annihilate_gpu.cu:147-178 and :362-434, exact mod P with
 * its torus <-> Z_P conversions). The last load of a level decides the
 * backend of its key switches. */
int cufhe_annihilate_key_load_ex(int level, const void* ahk, size_t words, int backend);
/* DeleteAnnihilateKey (reference src/annihilate_gpu.cu:679-688). */
int cufhe_annihilate_key_delete(int level);
/* AnnihilateKeySwitching (reference src/annihilate_gpu.cu:787-797, :690-725):
 * d_out = the annihilation key switch of the device TRLWE d_in (a then b,
 * 2N words): nbit rounds of c >>= 1, c += EvalAuto_d(c) for
 * d = 2^(bit+1) + 1. The phase keeps the input phase's constant coefficient;
 * the other coefficients sum to ~0. d_out may equal d_in. Asynchronous on
 * `stream`; needs no workspace. */
int cufhe_annihilate_key_switch(int level, void* d_out, const void* d_in, void* stream);
/* AnnihilateKeySwitchingBatchWithWorkspace (reference
 * src/annihilate_gpu.cu:727-785): `batch` TRLWEs at the given strides (torus
 * words, at least 2N) in one launch. Each output may equal its own input
 * (the same pointers and strides); other overlaps are not allowed. */
int cufhe_annihilate_key_switch_batch(int level, void* d_out, size_t out_stride, const void* d_in,
                                      size_t in_stride, size_t batch, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUFHE_AMD_H */
