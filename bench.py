#!/usr/bin/env python3
"""Benchmarks of the gate-bootstrapping path on one MI355X.

Default workload (the batched key-bundle configuration, the headline line):
4096 independent lvl1 binary gates, gate type uniform over {NAND, AND, OR,
XOR, XNOR, NOR}, random input bits from a seed, key bundle KB=2, FFT backend,
n=636 / N=1024 / l=2 / Bg=2^8. One step = one pass of the hot path over the
batch with inputs resident in HBM: identity key switching of all gates, then
the fused blind-rotation + sample-extraction kernel. Keys from seed 0.

Other workloads (--workload), each printing one JSON line of the same shape:
  lvl2mix  batched lvl02 gates (N=2048, Torus64, KB=2, FFT, or with
           --lvl2-backend ntt the Goldilocks NTT): 1536 binary + 512 MUX
           (2 bootstraps each), per-gate-type throughput vs the plain
           per-gate form;
  nttbb    NTT backend, block-binary lvl0 key, subset key switching: 4096
           NAND/XOR gates, bit-exact against the oracle on a sample.

The plain per-gate form (one gate per launch, round-robin over streams, the
reference's form) is measured on a sample beside each batched figure; the CPU
oracle on the host cores is the CPU baseline of the default workload.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The plain per-gate form runs one single-workgroup gate per launch over 64
# streams; HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (default
# 4), so at most 4 such gates run at once. A sweep over 8-32 queues and 64 or
# 128 streams peaked at 20 queues and 128 streams (4,710 gates/s against 987
# with 4 queues; 16: 3,689; 24: 4,417 over 64 streams but 2,842 over 128; 32:
# 1,489-1,871); the batched path uses one stream and is unaffected. Set
# before HIP initialises (no HIP call happens before main); --hw-queues 0
# keeps the environment's value.
def _hw_queues(argv):
    for i, a in enumerate(argv):
        if a == "--hw-queues" and i + 1 < len(argv):
            return argv[i + 1]
        if a.startswith("--hw-queues="):
            return a.split("=", 1)[1]
    return "20"


if _hw_queues(sys.argv) != "0":
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, int(_hw_queues(sys.argv)))))

from cufhe_amd.noise import predicted_lvl1_variance  # noqa: E402  (host-side formula, no GPU)

METRIC = "lvl1 NAND bootstraps/s @batch 4096 vs plain per-gate form; p50/p99 layer latency"
FLOP_PER_BOOTSTRAP_KB = 91502592      # 318 x 287,744 (algorithmic, lvl1 KB=2)
FLOP_PER_BOOTSTRAP_KB_LVL2 = 349077504  # 318 x 1,097,728 (lvl2 KB=2)
BK_BYTES_KB = 62521344                # 318 x 3 TRGSW x 8 polys x 512 bins x 16 B
BK_BYTES_NTT_BB = 20840448            # 636 TRGSW x 8 polys x 1024 bins x 4 B
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0
# one 64-lane VALU instruction per CU-cycle: 256 CUs x 2.4 GHz
VALU_PEAK_GINST = 256 * 2.4
# VALU wave-instructions of one 4096-gate k_br_ntt<2, 2> launch (rocprofv3
# SQ_INSTS_VALU, profiles/br_ntt_bb_lvl1_batch4096_counters.md)
NTT_BB_VALU_INSTS_C2 = 7.609e9
# VALU wave-instructions per lvl2 key-bundle bootstrap of k_br_ntt64<true>
# (rocprofv3 SQ_INSTS_VALU over a 2560-rotation launch,
# profiles/br_ntt64_kb_lvl2_mix2048_counters.md)
NTT64_KB_VALU_INSTS_PER_BOOTSTRAP = 5.281e7
GATE_MIX = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR"]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    i = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[i]


def want_bit(op, x, y, s=0):
    import cufhe_amd as C
    if op == "MUX":
        return y if s else x
    if op == "NMUX":
        return 1 - (y if s else x)
    return C.TRUTH[op](x, y)


def timed_steps(C, plan, steps):
    """K eager steps with per-kernel HIP events; device synchronised on both
    sides. Returns (total s, per-step ms, ks ms, br ms, combine ms)."""
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    step_ms, ks_ms, br_ms, lin_ms = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        k, b, l = plan.run_timed()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        ks_ms.append(k)
        br_ms.append(b)
        lin_ms.append(l)
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    return time.perf_counter() - t0, step_ms, ks_ms, br_ms, lin_ms


def graph_ms(C, gates, dpool, c, level=1, reps=3):
    gplan = C.Plan(gates, dpool, jobs_per_group=c, use_graph=True, level=level)
    gplan.run()
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    g0 = time.perf_counter()
    for _ in range(reps):
        gplan.run()
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    ms = (time.perf_counter() - g0) * 1e3 / reps
    gplan.destroy()
    return ms


def plain_form(C, gate_fn, calls, nstreams, lat_calls=5):
    """Throughput of one-gate-per-launch calls round-robin over streams, and
    single-gate latency on one stream."""
    streams = [C.Stream() for _ in range(nstreams)]
    for i, c in enumerate(calls[:8]):  # warm
        gate_fn(*c, streams[i % nstreams].handle)
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    p0 = time.perf_counter()
    for i, c in enumerate(calls):
        gate_fn(*c, streams[i % nstreams].handle)
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    tput = len(calls) / (time.perf_counter() - p0)
    lat = []
    for c in calls[:lat_calls]:
        l0 = time.perf_counter()
        gate_fn(*c, streams[0].handle)
        streams[0].synchronize()
        lat.append((time.perf_counter() - l0) * 1e3)
    for s in streams:
        s.destroy()
    return tput, pct(lat, 50)


def autotune(C, gates, dpool, cands):
    tune = {}
    for c in cands:
        plan = C.Plan(gates, dpool, jobs_per_group=c)
        plan.run_timed()
        tune[c] = min(plan.run_timed()[1] for _ in range(2))
        log("autotune C=%d: blind rotation %.2f ms" % (c, tune[c]))
        plan.destroy()
    return min(tune, key=tune.get), tune


def verify_chained(C, plan, dpool, ops, bits, B, k1, total):
    """Decryption check over `total` gates outside the timed region: the
    batch is run as consecutive layers, each layer's outputs becoming the
    next layer's first inputs (second inputs fixed), every output decrypted
    and compared with the plaintext evaluation, and the output phase error
    (phase - (+-1/8), torus units) accumulated. Returns (gates, failures,
    layers, noise statistics)."""
    import numpy as np

    if total <= 0:
        return 0, 0, 0, None
    W = C.TLWE1_WORDS
    tt = np.array([[C.TRUTH[o](x, y) for x in (0, 1) for y in (0, 1)] for o in ops], np.uint8)
    xs = bits[:B].astype(np.uint8).copy()
    ys = bits[B:2 * B].astype(np.uint8)
    out = np.empty((B, W), np.uint32)
    done = fails = layers = 0
    s1 = s2 = 0.0
    emax = 0.0
    mu = 0.125  # lvl1 gate output: +-1/8
    while done < total:
        plan.run()
        C.check(C.lib().cufhe_synchronize(), "sync")
        C.check(C.lib().cufhe_memcpy_d2h(C._ptr(out), dpool.at(2 * B * W * 4), out.nbytes, None), "d2h")
        C.check(C.lib().cufhe_stream_synchronize(None), "sync")
        xs = tt[np.arange(B), 2 * xs + ys]
        ph = C.phase(1, k1, out)
        fails += int(np.sum((ph > 0).astype(np.uint8) != xs))
        err = ph - np.where(xs == 1, mu, -mu)
        s1 += float(err.sum())
        s2 += float((err * err).sum())
        emax = max(emax, float(np.abs(err).max()))
        dpool.upload(out)  # outputs -> wires [0, B): the next layer's first inputs
        done += B
        layers += 1
    mean = s1 / done
    var = s2 / done - mean * mean
    noise = {"variance": var, "std": var ** 0.5, "mean": mean, "max_abs": emax,
             "bound": mu, "bound_over_std": mu / var ** 0.5 if var > 0 else None}
    log("chained verification: %d gates over %d layers, %d decryption failures; output noise "
        "std %.3g (variance %.3g), max |e| %.3g against the decryption bound 1/8"
        % (done, layers, fails, noise["std"], var, emax))
    return done, fails, layers, noise


def measured_traffic(workload):
    """HBM bytes per launch of the workload's blind-rotation kernel, from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs summarised under profiles/."""
    tf = os.path.join(ROOT, "profiles", "br_kernel_traffic.json")
    if not os.path.exists(tf):
        return None
    with open(tf) as f:
        return json.load(f).get(workload, {}).get("hbm_bytes_per_launch")


def host_cpu():
    """Host CPU model and logical CPU count (lscpu, os.cpu_count)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                             text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return model, os.cpu_count()


def cpu_baseline(args, C, ops, pool, B, bk, ksk):
    """lvl1 KB gates of the same mix on the host cores: the project's own CPU
    engine with the key precomputed in the Fourier domain (oracle/
    cpu_baseline.cpp), one gate per thread; single-thread latency first."""
    import numpy as np

    import oracle as O
    model, ncpu = host_cpu()
    # one GPU's share of the host: 16 threads, fewer if the host has fewer
    nthr = args.cpu_threads or min(16, ncpu or 1)
    eng = O.CpuBaseline(bk, True)
    try:
        opc = np.array([C.GATES[o] for o in ops], np.int32)
        lat = []
        for i in range(3):
            t0 = time.perf_counter()
            eng.gates(opc[i:i + 1], pool[i:i + 1], pool[B + i:B + i + 1], ksk, 1)
            lat.append((time.perf_counter() - t0) * 1e3)
        done = 0
        c0 = time.perf_counter()
        while time.perf_counter() - c0 < args.cpu_seconds and done < B:
            m = min(nthr, B - done)
            eng.gates(opc[done:done + m], pool[done:done + m], pool[B + done:B + done + m], ksk,
                      nthr)
            done += m
        cpu_s = time.perf_counter() - c0
    finally:
        eng.release()
    cpu = {"value": round(done / cpu_s, 3), "unit": "gates/s", "cores": nthr, "kind": "port",
           "host_cpu_model": model, "host_logical_cpus": ncpu,
           "single_thread_latency_ms": round(statistics.median(lat), 2),
           "sample": "%d gates of the same mix in %.1f s; own C++ CPU engine (bootstrapping key "
                     "precomputed in the Fourier domain, 6 FFNTs per key-bundle step, row-wise "
                     "integer key switch, AVX2/FMA), one gate per thread; not TFHEpp" %
                     (done, cpu_s)}
    log("cpu baseline: %.2f gates/s on %d threads (%s, %s logical CPUs); single-thread latency "
        "%.1f ms" % (cpu["value"], nthr, model, ncpu, cpu["single_thread_latency_ms"]))
    return cpu


def base_result(args, value, ms_per_step, dtype, config):
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic", "config": config,
    }


# ------------------------------------------------------------------ default
def workload_kb4096(args):
    import numpy as np

    import cufhe_amd as C

    B = args.batch
    log("keys (seed 0, key bundle)")
    k0, k1, _ = C.secret_key_gen(0)
    bk = C.bk_gen(1, True, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl01(bk, ksk, backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)

    rng = np.random.default_rng(args.seed)
    ops = [GATE_MIX[i] for i in rng.integers(0, len(GATE_MIX), B)]
    bits = rng.integers(0, 2, 2 * B).astype(np.uint8)
    log("encrypting %d inputs" % (2 * B))
    pool = np.zeros((3 * B, C.TLWE1_WORDS), np.uint32)
    pool[: 2 * B] = C.encrypt_bits(1, 1000 + args.seed, k1, bits)
    dpool = C.DeviceBuffer.from_numpy(pool)
    gates = [(ops[i], 2 * B + i, i, B + i, 0) for i in range(B)]
    want = np.array([C.TRUTH[o](int(bits[i]), int(bits[B + i])) for i, o in enumerate(ops)],
                    np.uint8)

    if args.jobs_per_group:
        best_c, tune = autotune(C, gates, dpool, [args.jobs_per_group])
    else:
        # the library's autotuner times every rotation kernel on B synthetic
        # jobs and records the fastest; jobs_per_group 0 then uses it
        best_c, tune = C.autotune_lvl1(B)
        for c, t in sorted(tune.items()):
            log("autotune C=%d: blind rotation %.2f ms" % (c, t))

    plan = C.Plan(gates, dpool, jobs_per_group=0 if not args.jobs_per_group else best_c)
    assert plan.jobs_per_group == best_c
    nbr = plan.bootstraps
    log("warmup %d steps (C=%d, %d bootstraps/step)" % (args.warmup, best_c, nbr))
    for _ in range(args.warmup):
        plan.run_timed()
    out = dpool.download(np.uint32, (3 * B, C.TLWE1_WORDS))[2 * B:]
    failures = int(np.sum(C.decrypt_bits(1, k1, out) != want))
    log("decryption failures after warmup: %d / %d" % (failures, B))

    total, step_ms, ks_ms, br_ms, _ = timed_steps(C, plan, args.steps)
    value = B * args.steps / total
    ms_per_step = total * 1e3 / args.steps
    log("batched: %.1f gates/s, %.2f ms/step (ks %.2f ms, br %.2f ms)" %
        (value, ms_per_step, statistics.mean(ks_ms), statistics.mean(br_ms)))
    gms = graph_ms(C, gates, dpool, best_c)
    # layer-latency distribution over more layers than the timed steps (outside
    # the timed region): p99 from --latency-layers eager layers
    lat_ms = step_ms
    if args.latency_layers > len(step_ms):
        _, extra, _, _, _ = timed_steps(C, plan, args.latency_layers - len(step_ms))
        lat_ms = step_ms + extra
    verified, vfail, vlayers, vnoise = verify_chained(C, plan, dpool, ops, bits, B, k1,
                                                      args.verify_gates)

    P = min(args.plain_sample, B)
    W = C.TLWE1_WORDS * 4

    def gate_fn(i, st):
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[ops[i]], dpool.at((2 * B + i) * W),
                                        dpool.at(i * W), dpool.at((B + i) * W), None, st), "gate")
    plain_tput, plain_lat = plain_form(C, gate_fn, [(i,) for i in range(P)], args.plain_streams)
    log("plain form: %.1f gates/s over %d streams; single-gate latency p50 %.2f ms" %
        (plain_tput, args.plain_streams, plain_lat))

    br_avg_s = statistics.mean(br_ms) / 1e3
    achieved = FLOP_PER_BOOTSTRAP_KB * nbr / br_avg_s / 1e12
    groups = -(-nbr // best_c)
    key_gbps = BK_BYTES_KB * groups / br_avg_s / 1e9
    traffic = measured_traffic("kb4096")

    cpu = None if args.no_cpu else cpu_baseline(args, C, ops, pool, B, bk, ksk)

    res = base_result(args, value, ms_per_step, "f64", {
        "workload": "batched lvl1 binary gates, KB=2, FFT, 4096 gates uniform over "
                    "NAND/AND/OR/XOR/XNOR/NOR, inputs resident in HBM",
        "global_batch": B, "n": 636, "N": 1024, "l": 2, "Bgbit": 8,
        "jobs_per_group": best_c, "autotune_br_ms": {str(k): round(v, 3) for k, v in tune.items()},
        "parallelism": "single GPU",
    })
    res.update({
        "p50_layer_ms": round(pct(lat_ms, 50), 3),
        "p99_layer_ms": round(pct(lat_ms, 99), 3),
        "layer_latency_samples": len(lat_ms),
        "graph_ms_per_step": round(gms, 3),
        "kernel_ms": {"key_switch": round(statistics.mean(ks_ms), 3),
                      "blind_rotate": round(statistics.mean(br_ms), 3)},
        "plain_form": {"gates_per_s": round(plain_tput, 2), "streams": args.plain_streams,
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                       "sample": P, "latency_p50_ms": round(plain_lat, 3)},
        "batched_over_plain": round(value / plain_tput, 2),
        "decrypt_failures": failures,
        "verified_chain": {"gates": verified, "layers": vlayers, "decrypt_failures": vfail,
                           "output_noise": vnoise,
                           "predicted_variance": predicted_lvl1_variance("kb"),
                           "predicted_from": "318 key-bundle steps, N=1024, l=2, Bg=2^8, "
                                             "alpha_bk=2^-24.8, 6 key-noise terms per step (cufhe_amd/noise.py)"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "basis": "algorithmic FP64 flops (FP64 VALU FMAs: the peak is MI355X's FP64 "
                              "figure, the same for vector and matrix)"},
        "key_stream": {"algorithmic_GBps": round(key_gbps, 1), "peak_GBps": HBM_PEAK_GBPS},
        "cpu_baseline": cpu,
    })
    return res


# ------------------------------------------------------------------ lvl2 mix
def workload_lvl2mix(args):
    import numpy as np

    import cufhe_amd as C

    nbin, nmux = 1536, 512
    ntt = args.lvl2_backend == "ntt"
    log("keys (seed 0, lvl02 key bundle, %s backend)" % args.lvl2_backend)
    k0, _, k2 = C.secret_key_gen(0)
    bk = C.bk_gen(2, True, 0, k0, k2)
    ksk = C.ksk_gen(2, 0, k2, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl02(bk, ksk, backend=C.BACKEND_NTT if ntt else C.BACKEND_FFT, key_bundle=True)
    if args.lvl2_jobs_per_group:
        C.check(C.lib().cufhe_set_lvl2_jobs_per_group(args.lvl2_jobs_per_group), "lvl2 jobs")
    rng = np.random.default_rng(args.seed)
    ops = [GATE_MIX[i] for i in rng.integers(0, len(GATE_MIX), nbin)] + ["MUX"] * nmux
    perm = rng.permutation(len(ops))
    ops = [ops[i] for i in perm]  # uneven mix, interleaved
    n = len(ops)
    bits = rng.integers(0, 2, 3 * n).astype(np.uint8)
    W0 = C.TLWE0_WORDS
    pool = np.zeros((4 * n, W0), np.uint32)
    pool[: 3 * n] = C.encrypt_bits(0, 2000 + args.seed, k0, bits)
    want = np.array([want_bit(o, int(bits[i]), int(bits[n + i]), int(bits[2 * n + i]))
                     for i, o in enumerate(ops)], np.uint8)
    dpool = C.DeviceBuffer.from_numpy(pool)

    def run_subset(idx, label):
        gates = [(ops[i], 3 * n + i, i, n + i, 2 * n + i) for i in idx]
        plan = C.Plan(gates, dpool, level=2)
        for _ in range(args.warmup):
            plan.run_timed()
        total, step_ms, ks_ms, br_ms, lin_ms = timed_steps(C, plan, args.steps)
        nb = plan.bootstraps
        plan.destroy()
        r = {"gates": len(idx), "bootstraps": nb,
             "gates_per_s": round(len(idx) * args.steps / total, 2),
             "bootstraps_per_s": round(nb * args.steps / total, 2),
             "ms_per_step": round(total * 1e3 / args.steps, 3),
             "p50_layer_ms": round(pct(step_ms, 50), 3), "p99_layer_ms": round(pct(step_ms, 99), 3),
             "kernel_ms": {"blind_rotate": round(statistics.mean(br_ms), 3),
                           "key_switch": round(statistics.mean(ks_ms), 3)}}
        log("lvl2 %s: %s" % (label, r))
        return r, statistics.mean(br_ms), nb

    mix, br_mix_ms, nb_mix = run_subset(list(range(n)), "mix")
    out = dpool.download(np.uint32, pool.shape)[3 * n:]
    failures = int(np.sum(C.decrypt_bits(0, k0, out) != want))
    binr, _, _ = run_subset([i for i in range(n) if ops[i] != "MUX"], "binary")
    muxr, _, _ = run_subset([i for i in range(n) if ops[i] == "MUX"], "mux")

    Wb = W0 * 4

    def gate_fn(i, st):
        C.check(C.lib().cufhe_gate_lvl02(C.GATES[ops[i]], dpool.at((3 * n + i) * Wb),
                                         dpool.at(i * Wb), dpool.at((n + i) * Wb),
                                         dpool.at((2 * n + i) * Wb), st), "gate_lvl02")
    sample = max(16, args.plain_sample // 4)
    bin_idx = [i for i in range(n) if ops[i] != "MUX"][:sample]
    mux_idx = [i for i in range(n) if ops[i] == "MUX"][:sample]
    pb, pb_lat = plain_form(C, gate_fn, [(i,) for i in bin_idx], args.plain_streams)
    pm, pm_lat = plain_form(C, gate_fn, [(i,) for i in mux_idx], args.plain_streams)
    if ntt:
        # integer-only kernel bound by VALU issue: measured VALU wave-instructions
        # per bootstrap over the live launch time
        valu_g = NTT64_KB_VALU_INSTS_PER_BOOTSTRAP * nb_mix / (br_mix_ms / 1e3) / 1e9
        roof = {"bound": "valu", "achieved": round(valu_g, 1), "peak": VALU_PEAK_GINST,
                "unit": "G wave-instr/s", "frac": round(valu_g / VALU_PEAK_GINST, 4),
                "traffic": measured_traffic("lvl2mix_ntt"),
                "basis": "measured SQ_INSTS_VALU per bootstrap (no FP64; integer issue bound)"}
    else:
        achieved = FLOP_PER_BOOTSTRAP_KB_LVL2 * nb_mix / (br_mix_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                "traffic": measured_traffic("lvl2mix")}
    res = base_result(args, mix["gates_per_s"], mix["ms_per_step"], "u64 mod P" if ntt else "f64", {
        "workload": "batched lvl02 gates (lvl0 in/out, BR at N=2048 Torus64, KB=2, %s): "
                    "1536 binary (uniform over NAND/AND/OR/XOR/XNOR/NOR) + 512 MUX, interleaved"
                    % ("NTT mod 2^64-2^32+1" if ntt else "FFT"),
        "global_batch": n, "n": 636, "N": 2048, "l": 4, "Bgbit": 10,
        "parallelism": "single GPU",
    })
    res["metric"] = "lvl02 gates/s, 1536 binary + 512 MUX, batched vs plain per-gate form"
    res.update({
        "decrypt_failures": failures,
        "per_type": {"mix": mix, "binary": binr, "mux": muxr},
        "plain_form": {"binary_gates_per_s": round(pb, 2), "mux_gates_per_s": round(pm, 2),
                       "binary_latency_p50_ms": round(pb_lat, 3),
                       "mux_latency_p50_ms": round(pm_lat, 3), "streams": args.plain_streams,
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
        "batched_over_plain": {"binary": round(binr["gates_per_s"] / pb, 2),
                               "mux": round(muxr["gates_per_s"] / pm, 2)},
        "roofline": roof,
        "cpu_baseline": None,
    })
    return res


# ------------------------------------------------------------------ NTT block-binary
def workload_nttbb(args):
    import numpy as np

    import cufhe_amd as C

    B = args.batch
    log("keys (seed 0, block-binary lvl0 key, subset KSK)")
    k0, k1, _ = C.secret_key_gen(0, block_binary=True)
    bk = C.bk_gen(1, False, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0, subset=True)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl01(bk, ksk, backend=C.BACKEND_NTT, variant=C.BR_BLOCK_BINARY, subset=True)
    rng = np.random.default_rng(args.seed)
    ops = ["NAND" if v == 0 else "XOR" for v in rng.integers(0, 2, B)]
    bits = rng.integers(0, 2, 2 * B).astype(np.uint8)
    pool = np.zeros((3 * B, C.TLWE1_WORDS), np.uint32)
    pool[: 2 * B] = C.encrypt_bits(1, 3000 + args.seed, k1, bits)
    dpool = C.DeviceBuffer.from_numpy(pool)
    gates = [(ops[i], 2 * B + i, i, B + i, 0) for i in range(B)]
    want = np.array([C.TRUTH[o](int(bits[i]), int(bits[B + i])) for i, o in enumerate(ops)],
                    np.uint8)
    if args.jobs_per_group:
        best_c, tune = autotune(C, gates, dpool, [args.jobs_per_group])
    else:
        best_c, tune = C.autotune_lvl1(B)
        for c, t in sorted(tune.items()):
            log("autotune C=%d: blind rotation %.2f ms" % (c, t))
    plan = C.Plan(gates, dpool, jobs_per_group=best_c)
    nbr = plan.bootstraps
    for _ in range(args.warmup):
        plan.run_timed()
    out = dpool.download(np.uint32, (3 * B, C.TLWE1_WORDS))[2 * B:]
    failures = int(np.sum(C.decrypt_bits(1, k1, out) != want))
    bit_exact = None
    if not args.no_cpu:
        import oracle as O
        sample = rng.choice(B, 32, replace=False)
        codes = np.array([C.GATES[ops[i]] for i in sample], np.int32)
        ref = O.gate_batch_lvl1(O.NTT, O.BR_BB | O.SUBSET, codes, pool[sample], pool[B + sample],
                                bk, ksk, min(16, os.cpu_count() or 1))
        bit_exact = int(sum(np.array_equal(out[i], ref[j]) for j, i in enumerate(sample)))
        log("bit-exact vs oracle: %d / %d" % (bit_exact, len(sample)))
    total, step_ms, ks_ms, br_ms, _ = timed_steps(C, plan, args.steps)
    value = B * args.steps / total
    gms = graph_ms(C, gates, dpool, best_c)
    verified, vfail, vlayers, vnoise = verify_chained(C, plan, dpool, ops, bits, B, k1,
                                                      args.verify_gates)
    W = C.TLWE1_WORDS * 4

    def gate_fn(i, st):
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[ops[i]], dpool.at((2 * B + i) * W),
                                        dpool.at(i * W), dpool.at((B + i) * W), None, st), "gate")
    plain_tput, plain_lat = plain_form(C, gate_fn, [(i,) for i in range(min(args.plain_sample, B))],
                                       args.plain_streams)
    br_avg_s = statistics.mean(br_ms) / 1e3
    groups = -(-nbr // best_c)
    key_gbps = BK_BYTES_NTT_BB * groups / br_avg_s / 1e9
    valu_g = (NTT_BB_VALU_INSTS_C2 / br_avg_s / 1e9) if best_c == 2 else float("nan")
    res = base_result(args, value, total * 1e3 / args.steps, "u32 mod P", {
        "workload": "batched lvl1 NAND/XOR gates, NTT backend (P=2147473409), block-binary "
                    "lvl0 key (blocks of 4), subset key switching, 4096 gates",
        "global_batch": B, "n": 636, "N": 1024, "l": 2, "Bgbit": 8, "ell": 4,
        "jobs_per_group": best_c, "autotune_br_ms": {str(k): round(v, 3) for k, v in tune.items()},
        "parallelism": "single GPU",
    })
    res["metric"] = "lvl1 NAND/XOR bootstraps/s @batch 4096, NTT block-binary"
    res.update({
        "p50_layer_ms": round(pct(step_ms, 50), 3), "p99_layer_ms": round(pct(step_ms, 99), 3),
        "graph_ms_per_step": round(gms, 3),
        "kernel_ms": {"key_switch": round(statistics.mean(ks_ms), 3),
                      "blind_rotate": round(statistics.mean(br_ms), 3)},
        "plain_form": {"gates_per_s": round(plain_tput, 2), "streams": args.plain_streams,
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                       "latency_p50_ms": round(plain_lat, 3)},
        "batched_over_plain": round(value / plain_tput, 2),
        "decrypt_failures": failures,
        "verified_chain": {"gates": verified, "layers": vlayers, "decrypt_failures": vfail,
                           "output_noise": vnoise,
                           "predicted_variance": predicted_lvl1_variance("bb"),
                           "predicted_from": "159 block-binary blocks of 4, N=1024, l=2, Bg=2^8, "
                                             "alpha_bk=2^-24.8, 8 key-noise terms per block (cufhe_amd/noise.py)"},
        "bit_exact_sample": bit_exact,
        # integer-only kernel bound by VALU issue: VALU wave-instructions per
        # launch (profiles/br_ntt_bb_lvl1_batch4096_counters.md, C=2) over
        # the live launch time, against one wave-instruction per CU-cycle
        "roofline": {"bound": "valu", "achieved": round(valu_g, 1), "peak": VALU_PEAK_GINST,
                     "unit": "G wave-instr/s", "frac": round(valu_g / VALU_PEAK_GINST, 4),
                     "traffic": measured_traffic("nttbb"),
                     "basis": "measured SQ_INSTS_VALU per launch (no FP64; integer issue bound)"},
        "key_stream": {"algorithmic_GBps": round(key_gbps, 1), "peak_GBps": HBM_PEAK_GBPS},
        "cpu_baseline": None,
    })
    return res


# ------------------------------------------------------------------ annihilation
# Algorithmic FP64 flops of one annihilation key switch: nbit rounds x (l
# forward + 2 inverse H-point FFTs at 5 H log2 H, + 2 l H complex MACs at 8).
FLOP_PER_ANNIHILATE = {1: 10 * (4 * 5 * 512 * 9 + 2 * 2 * 512 * 8),      # 1,085,440
                       2: 11 * (6 * 5 * 1024 * 10 + 2 * 4 * 1024 * 8)}  # 4,100,096


def _annihilate_level(C, args, level, B, key, backend=0):
    import numpy as np

    N = 1024 if level == 1 else 2048
    T = np.uint32 if level == 1 else np.uint64
    w = 32 if level == 1 else 64
    ahk = C.annihilate_key_gen(level, args.seed, key)
    C.annihilate_key_load(level, ahk, backend=backend)
    rng = np.random.default_rng(args.seed + level)
    bits = rng.integers(0, 2, B)
    m = rng.integers(0, 2**63, (B, N), dtype=np.uint64)
    m = m.astype(T) if level == 2 else (m >> np.uint64(31)).astype(T)
    mu = T(1) << T(w - 3)
    m[:, 0] = np.where(bits == 1, mu, T(0) - mu)  # the bit at X^0, junk elsewhere
    log("level %d: encrypting %d TRLWEs" % (level, B))
    ct = C.trlwe_encrypt_level(level, 900 + args.seed, key, m)
    din = C.DeviceBuffer.from_numpy(ct)
    dout = C.DeviceBuffer(ct.nbytes)
    st = C.Stream()

    def launch():
        C.check(C.lib().cufhe_annihilate_key_switch_batch(level, dout.ptr, 2 * N, din.ptr, 2 * N,
                                                          B, st.handle), "annihilate")
    for _ in range(args.warmup):
        launch()
    st.synchronize()
    evs = [(C.Event(), C.Event()) for _ in range(args.steps)]
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(st)
        launch()
        e1.record(st)
    st.synchronize()
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    total = time.perf_counter() - t0
    kms = [e0.elapsed_ms(e1) for e0, e1 in evs]
    for e0, e1 in evs:
        e0.destroy()
        e1.destroy()
    st.destroy()
    out = dout.download(T, (B, 2 * N))
    ph0 = C.trlwe_phase(level, key, out)[:, 0]
    ph0 = ph0.view(np.int32) if level == 1 else ph0.view(np.int64)
    failures = int(np.sum((ph0 > 0).astype(np.int64) != bits))
    dev = (np.abs(ph0.astype(np.float64)) - float(mu)) / 2.0 ** w
    oracle_cmp = None
    if not args.no_cpu:
        import oracle as O
        worst, same = 0, 0
        mode = O.NTT if backend else O.EXACT
        for i in (0, B - 1):
            ref = O.annihilate(level, mode, ct[i], ahk)
            same += int(np.array_equal(out[i], ref))
            d = C.trlwe_phase(level, key, out[i])[0] - C.trlwe_phase(level, key, ref)[0]
            d = d.view(np.int32) if level == 1 else d.view(np.int64)
            worst = max(worst, int(np.abs(d.astype(np.float64)).max()))
        oracle_cmp = {"oracle_mode": "ntt" if backend else "exact", "sampled": 2,
                      "identical_words": same, "max_phase_diff": worst / 2.0 ** w}
    k_s = statistics.mean(kms) / 1e3
    achieved = FLOP_PER_ANNIHILATE[level] * B / k_s / 1e12
    log("level %d%s: %d TRLWEs, kernel %.3f ms, %.0f /s, %.2f TFLOP/s FP64, %d decrypt failures, "
        "vs oracle %s" % (level, " (NTT)" if backend else "", B, k_s * 1e3, B / k_s, achieved,
                          failures, oracle_cmp))
    return {
        "trlwes": B, "N": N, "annihilations_per_s": round(B * args.steps / total, 2),
        "ms_per_step": round(total * 1e3 / args.steps, 3), "kernel_ms": round(k_s * 1e3, 3),
        "decrypt_failures": failures, "max_const_coef_error": float(np.abs(dev).max()),
        "vs_oracle": oracle_cmp,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic("annihilate") if level == 1 and not backend else None,
                     "basis": "algorithmic FP64 flops %d per annihilation (FP64 VALU)"
                              % FLOP_PER_ANNIHILATE[level]},
    }, ct[:64], ahk


def workload_annihilate(args):
    import numpy as np

    import cufhe_amd as C

    B = args.batch
    log("keys (seed 0)")
    _, k1, k2 = C.secret_key_gen(0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    l1, sample1, ahk1 = _annihilate_level(C, args, 1, B, k1)
    l2, _, _ = _annihilate_level(C, args, 2, max(1, B // 4), k2)
    l2n, _, _ = _annihilate_level(C, args, 2, max(1, B // 4), k2, backend=C.BACKEND_NTT)
    l2n["backend"] = "Goldilocks NTT (exact mod P)"
    del l2n["roofline"]
    cpu = None
    if not args.no_cpu:
        import oracle as O
        log("cpu baseline: oracle FFT mode, one thread")
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            O.annihilate(1, O.FFT, sample1, ahk1)
            n += len(sample1)
        el = time.perf_counter() - t0
        cpu = {"value": round(n / el, 2), "unit": "annihilations/s", "cores": 1, "kind": "port",
               "sample": "%d level-1 annihilations (N=1024) in %.1f s, batches of %d: the "
                         "oracle's FP64 FFT mode (oracle/oracle.cpp annihilate: key spectra once "
                         "per batch, radix-2 FFT), one thread" % (n, el, len(sample1)),
               "host_cpu_model": host_cpu()[0]}
    res = base_result(args, l1["annihilations_per_s"], l1["ms_per_step"], "f64", {
        "workload": "annihilation key switching of 4096 lvl1 TRLWEs (N=1024, l=2, Bgbit=8; "
                    "10 rounds) in one launch, inputs resident in HBM; lvl2 (N=2048, l=4, "
                    "Bgbit=10; 11 rounds) at a quarter of the batch",
        "global_batch": B, "parallelism": "single GPU"})
    res["metric"] = "lvl1 annihilation key switches/s @batch 4096 (TRLWE N=1024)"
    res["unit"] = "annihilations/s"
    res.update({"kernel_ms": l1["kernel_ms"], "decrypt_failures": l1["decrypt_failures"],
                "lvl1": l1, "lvl2": l2, "lvl2_ntt": l2n, "roofline": l1["roofline"],
                "cpu_baseline": cpu})
    del l1["roofline"]
    return res


# ------------------------------------------------------------------ AES circuit
def workload_aes(args):
    """AES-128 decryption of args.aes_blocks blocks (each its own key) as gate
    circuits on the serving loop, all submitted at once."""
    import numpy as np

    import cufhe_amd as C
    from cufhe_amd import aes as A

    log("keys (seed 0)")
    k0, k1, _ = C.secret_key_gen(0)
    bk = C.bk_gen(1, True, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl01(bk, ksk)
    circ = A.decrypt_circuit()
    nb = args.aes_blocks
    rng = np.random.default_rng(args.seed)
    keys = [bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist()) for _ in range(nb)]
    cts = [bytes(rng.integers(0, 256, 16, dtype=np.uint8).tolist()) for _ in range(nb)]
    expect = None
    if not args.no_cpu:
        from oracle import aes_ref as R  # the expected plaintexts (checker only)
        expect = [R.decrypt_block(k, c) for k, c in zip(keys, cts)]
    ins = []
    for i in range(nb):
        # the key owner expands the key (FIPS-197 5.2) before encrypting it
        rk = A.round_keys(keys[i])
        bits = A.input_bits(cts[i], rk)
        ins.append(np.concatenate([C.encrypt_bits(1, 4000 + i, k1, bits[:-2]),
                                   A.trivial_bits(bits[-2:])]))
    wires = circ.n_inputs + circ.n_gates
    srv = C.Server(pool_bytes=(nb + 1) * wires * C.TLWE1_WORDS * 4,
                   max_batch_bootstraps=args.aes_batch)
    try:
        warm = srv.submit(circ, ins[0])
        srv.result(warm)
        s0 = srv.stats()
        C.check(C.lib().cufhe_synchronize(), "synchronize")
        t0 = time.perf_counter()
        rids = [srv.submit(circ, x) for x in ins]
        outs = [srv.result(r) for r in rids]
        total = time.perf_counter() - t0
        lat = [srv.info(r)["done"] - srv.info(r)["submit"] for r in rids]
        s1 = srv.stats()
    finally:
        srv.destroy()
    got = [A.from_bits(C.decrypt_bits(1, k1, o)) for o in outs]
    wrong = None if expect is None else int(sum(g != e for g, e in zip(got, expect)))
    nbs = s1["bootstraps"] - s0["bootstraps"]
    busy = s1["busy_s"] - s0["busy_s"]
    log("aes: %d blocks in %.2f s (%d rounds, %d bootstraps, GPU busy %.2f s), %s wrong"
        % (nb, total, s1["rounds"] - s0["rounds"], nbs, busy, wrong))
    res = base_result(args, nb / total, total * 1e3, "f64", {
        "workload": "AES-128 decryption as a Boolean circuit of lvl1 gates (%d gates, %d "
                    "bootstraps, %d layers per block), %d blocks with their own keys "
                    "submitted together to the serving loop" % (circ.n_gates, circ.bootstraps,
                                                                 circ.depth, nb),
        "blocks": nb, "parallelism": "single GPU"})
    res["metric"] = "AES-128 decryption blocks/s (gate circuit, serving loop)"
    res["unit"] = "blocks/s"
    res["steps"] = 1
    res.update({"bootstraps_per_s": round(nbs / total, 1), "rounds": s1["rounds"] - s0["rounds"],
                "gpu_busy_s": round(busy, 3), "block_latency_p50_s": round(pct(lat, 50), 3),
                "wrong_blocks": wrong, "cpu_baseline": None, "roofline": None})
    return res


# ------------------------------------------------------------------ ASCON circuit
def workload_ascon(args):
    """ASCON-XOF (args.ascon_in_bytes -> args.ascon_out_bytes) of
    args.ascon_messages messages as gate circuits on the serving loop, all
    submitted at once."""
    import numpy as np

    import cufhe_amd as C
    from cufhe_amd import aes as A
    from cufhe_amd import ascon as S

    log("keys (seed 0)")
    k0, k1, _ = C.secret_key_gen(0)
    bk = C.bk_gen(1, True, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl01(bk, ksk)
    circ = S.xof_circuit(args.ascon_in_bytes, args.ascon_out_bytes)
    nm = args.ascon_messages
    rng = np.random.default_rng(args.seed)
    msgs = [bytes(rng.integers(0, 256, args.ascon_in_bytes, dtype=np.uint8).tolist())
            for _ in range(nm)]
    expect = None
    if not args.no_cpu:
        from oracle import ascon_ref as R  # the expected outputs (checker only)
        expect = [R.ascon_xof(m, args.ascon_out_bytes) for m in msgs]
    ins = [np.concatenate([C.encrypt_bits(1, 5000 + i, k1, A.bits(m)), A.trivial_bits([0, 1])])
           for i, m in enumerate(msgs)]
    wires = circ.n_inputs + circ.n_gates
    srv = C.Server(pool_bytes=(nm + 1) * wires * C.TLWE1_WORDS * 4,
                   max_batch_bootstraps=args.aes_batch)
    try:
        srv.result(srv.submit(circ, ins[0]))
        s0 = srv.stats()
        C.check(C.lib().cufhe_synchronize(), "synchronize")
        t0 = time.perf_counter()
        rids = [srv.submit(circ, x) for x in ins]
        outs = [srv.result(r) for r in rids]
        total = time.perf_counter() - t0
        lat = [srv.info(r)["done"] - srv.info(r)["submit"] for r in rids]
        s1 = srv.stats()
    finally:
        srv.destroy()
    got = [A.from_bits(C.decrypt_bits(1, k1, o)) for o in outs]
    wrong = None if expect is None else int(sum(g != e for g, e in zip(got, expect)))
    nbs = s1["bootstraps"] - s0["bootstraps"]
    log("ascon: %d messages in %.2f s (%d rounds, %d bootstraps), %s wrong"
        % (nm, total, s1["rounds"] - s0["rounds"], nbs, wrong))
    res = base_result(args, nm / total, total * 1e3, "f64", {
        "workload": "ASCON-XOF (%d-byte messages, %d-byte outputs) as a Boolean circuit of "
                    "lvl1 gates (%d gates, %d bootstraps, %d layers per message), %d messages "
                    "submitted together to the serving loop"
                    % (args.ascon_in_bytes, args.ascon_out_bytes, circ.n_gates,
                       circ.bootstraps, circ.depth, nm),
        "messages": nm, "parallelism": "single GPU"})
    res["metric"] = "ASCON-XOF messages/s (gate circuit, serving loop)"
    res["unit"] = "messages/s"
    res["steps"] = 1
    res.update({"bootstraps_per_s": round(nbs / total, 1), "rounds": s1["rounds"] - s0["rounds"],
                "gpu_busy_s": round(s1["busy_s"] - s0["busy_s"], 3),
                "message_latency_p50_s": round(pct(lat, 50), 3), "wrong_outputs": wrong,
                "cpu_baseline": None, "roofline": None})
    return res


# ------------------------------------------------------------------ serving
def workload_serving(args):
    import threading

    import numpy as np

    import cufhe_amd as C
    from cufhe_amd.circuits import random_circuit

    log("keys (seed 0, key bundle)")
    k0, k1, _ = C.secret_key_gen(0)
    bk = C.bk_gen(1, True, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.initialize_lvl01(bk, ksk, backend=C.BACKEND_FFT, variant=C.BR_KEY_BUNDLE)
    rng = np.random.default_rng(args.seed)

    def make(n, tag):
        circs, ins, bits = [], [], []
        for i in range(n):
            c = random_circuit(rng, int(rng.integers(64, 1025)), int(rng.integers(4, 33)))
            b = rng.integers(0, 2, c.n_inputs).astype(np.uint8)
            circs.append(c)
            bits.append(b)
            ins.append(C.encrypt_bits(1, 10000 * tag + i, k1, b))
        return circs, ins, bits

    def serve(circs, ins, arrivals):
        srv = C.Server(pool_bytes=int(args.serve_pool_gb * (1 << 30)),
                       jobs_per_group=args.jobs_per_group)
        ids = [None] * len(circs)
        t_start = time.perf_counter()

        def client():
            for i, c in enumerate(circs):
                dt = t_start + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(c, ins[i])

        th = threading.Thread(target=client)
        th.start()
        th.join()
        outs = [srv.result(r) for r in ids]
        infos = [srv.info(r) for r in ids]
        st = srv.stats()
        srv.destroy()
        span = max(f["done"] for f in infos) - min(f["submit"] for f in infos)
        return outs, infos, st, span

    # capacity: a burst of circuits submitted at once (the server runs at
    # its full co-batching depth); circuits per second it sustains
    circs, ins, _ = make(args.serve_burst, 1)
    _, _, st, span = serve(circs, ins, np.zeros(len(circs)))
    cap_gates = sum(c.n_gates for c in circs) / span
    cap = len(circs) / span
    log("capacity (burst of %d circuits): %.1f circuits/s, %.0f gates/s, GPU busy %.2f"
        % (len(circs), cap, cap_gates, st["busy_s"] / span))
    if args.serve_rates:
        rates = [float(r) for r in args.serve_rates.split(",")]
    else:
        rates = [round(f * cap, 2) for f in (float(x) for x in args.serve_levels.split(","))]
    results = []
    for li, rate in enumerate(rates):
        circs, ins, bits = make(args.serve_circuits, li + 2)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(circs)))
        outs, infos, st, span = serve(circs, ins, arrivals)
        lat = [1e3 * (f["done"] - f["submit"]) for f in infos]
        gates = sum(c.n_gates for c in circs)
        fails = sum(int(np.sum(C.decrypt_bits(1, k1, o) != c.evaluate(b)))
                    for o, c, b in zip(outs, circs, bits))
        r = {"rate_circuits_per_s": rate, "load_vs_capacity": round(rate / cap, 3),
             "circuits": len(circs), "gates": gates,
             "gates_per_s": round(gates / span, 1),
             "p50_latency_ms": round(pct(lat, 50), 1), "p99_latency_ms": round(pct(lat, 99), 1),
             "max_latency_ms": round(max(lat), 1),
             "rounds": st["rounds"], "mean_gates_per_round": round(st["gates"] / st["rounds"], 1),
             "gpu_busy_frac": round(st["busy_s"] / span, 3),
             "peak_pool_MiB": round(st["peak_pool_bytes"] / 2**20, 1),
             "decrypt_failures": fails}
        log("serving %s" % r)
        results.append(r)
    top = results[-1]
    res = base_result(args, top["gates_per_s"], 1e3 / max(top["gates_per_s"], 1e-9), "f64", {
        "workload": "serving: Poisson arrivals of random layered lvl1 circuits (width U[64,1024], "
                    "depth U[4,32], 10%% MUX, binary gates uniform), KB=2 FFT, device-memory "
                    "admission (%g GiB pool), layers of all admitted circuits co-batched per round, "
                    "two rounds in flight" % args.serve_pool_gb,
        "rates_circuits_per_s": rates, "circuits_per_rate": args.serve_circuits,
        "parallelism": "single GPU",
    })
    res["metric"] = "serving gates/s and p50/p99 per-circuit latency (lvl1 circuits, KB=2)"
    res.update({"capacity": {"circuits_per_s": round(cap, 2), "gates_per_s": round(cap_gates, 1),
                             "burst_circuits": args.serve_burst},
                "levels": results,
                "p50_latency_ms": top["p50_latency_ms"], "p99_latency_ms": top["p99_latency_ms"],
                "decrypt_failures": sum(r["decrypt_failures"] for r in results),
                "roofline": None, "cpu_baseline": None})
    return res


WORKLOADS = {"kb4096": workload_kb4096, "lvl2mix": workload_lvl2mix, "nttbb": workload_nttbb,
             "serving": workload_serving, "annihilate": workload_annihilate, "aes": workload_aes,
             "ascon": workload_ascon}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="kb4096")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--latency-layers", type=int, default=200,
                    help="eager layers (timed steps included) behind p50/p99_layer_ms; the "
                         "extra layers run after the timed region")
    ap.add_argument("--hw-queues", type=int, default=20,
                    help="GPU_MAX_HW_QUEUES for this process (1-32; 0 keeps the environment's)")
    ap.add_argument("--lvl2-backend", choices=["fft", "ntt"], default="fft",
                    help="lvl2mix: blind rotation on the FP64 FFT or the Goldilocks NTT backend")
    ap.add_argument("--lvl2-jobs-per-group", type=int, default=0,
                    help="lvl2mix, FFT backend: lvl2 jobs per workgroup (1 or 2; 0 = library default)")
    ap.add_argument("--jobs-per-group", type=int, default=0,
                    help="bootstraps per workgroup; 0 = autotune over {1, 2, 4, 6, 8} "
                         "(6: key shared through LDS, FFT key bundle only)")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--aes-blocks", type=int, default=8, help="aes: blocks submitted together")
    ap.add_argument("--aes-batch", type=int, default=16384,
                    help="aes, ascon: the server's bootstraps per round")
    ap.add_argument("--ascon-messages", type=int, default=32)
    ap.add_argument("--ascon-in-bytes", type=int, default=16)
    ap.add_argument("--ascon-out-bytes", type=int, default=32)
    ap.add_argument("--plain-streams", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--verify-gates", type=int, default=1 << 20,
                    help="kb4096, nttbb: gates decrypted and checked after timing, as chained "
                         "layers (0 = off)")
    ap.add_argument("--serve-rates", default="",
                    help="circuits/s, one level each (default: --serve-levels of the capacity)")
    ap.add_argument("--serve-levels", default="0.3,0.6,0.9",
                    help="arrival rates as fractions of the measured burst capacity")
    ap.add_argument("--serve-burst", type=int, default=32, help="circuits of the capacity burst")
    ap.add_argument("--serve-circuits", type=int, default=64, help="circuits per rate level")
    ap.add_argument("--serve-pool-gb", type=float, default=4.0)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the workload, copy this many MiB device-to-device (a kernel of "
                         "known traffic for scaling HBM counters); not part of any figure")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the gate batch is one GPU's share; "
              "multi-GPU sharding (independent gates, replicated keys) is not part of this "
              "build. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    res = WORKLOADS[args.workload](args)
    if args.calibrate_copy_mib > 0:
        import cufhe_amd as C
        n = args.calibrate_copy_mib << 20
        a, b = C.DeviceBuffer(n), C.DeviceBuffer(n)
        C.check(C.lib().cufhe_memset_d(a.ptr, 1, n, None), "memset")
        C.check(C.lib().cufhe_synchronize(), "synchronize")
        C.check(C.lib().cufhe_memcpy_d2d(b.ptr, a.ptr, n, None), "copy")
        C.check(C.lib().cufhe_synchronize(), "synchronize")
        log("calibration copy: %d bytes device-to-device" % n)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
