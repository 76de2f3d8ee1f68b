#!/usr/bin/env python3
"""Benchmark of the batched lvl1 gate-bootstrapping path on one MI355X.

Workload (the batched key-bundle configuration): 4096 independent lvl1
binary gates, gate type uniform over {NAND, AND, OR, XOR, XNOR, NOR}, random
input bits from a seed, key bundle KB=2, FFT backend, n=636 / N=1024 / l=2 /
Bg=2^8. One step = one pass of the hot path over the batch with inputs
resident in HBM: identity key switching of all gates, then the fused
blind-rotation + sample-extraction kernel. Keys are generated from seed 0.

Prints ONE JSON line (rank 0). The plain per-gate form (one gate per launch,
round-robin over streams, the reference's form) is measured on a sample and
reported beside the batched figure, with the CPU oracle as the CPU baseline.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "lvl1 NAND bootstraps/s @batch 4096 vs plain per-gate form; p50/p99 layer latency"
FLOP_PER_BOOTSTRAP_KB = 91502592      # 318 x 287,744 (algorithmic, lvl1 KB=2)
BK_BYTES_KB = 62521344                # 318 x 3 TRGSW x 8 polys x 512 bins x 16 B
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0
GATE_MIX = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR"]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    i = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[i]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--jobs-per-group", type=int, default=0,
                    help="bootstraps per workgroup; 0 = autotune over {1, 2, 4, 8}")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--plain-streams", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the gate batch is one GPU's share; "
              "multi-GPU sharding (independent gates, replicated keys) is not part of this "
              "build. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cufhe_amd as C

    B = args.batch
    log("keys (seed 0, key bundle)")
    k0, k1, _ = C.secret_key_gen(0)
    bk = C.bk_gen(1, True, 0, k0, k1)
    ksk = C.ksk_gen(1, 0, k1, k0)
    C.check(C.lib().cufhe_initialize(), "initialize")
    C.check(C.lib().cufhe_initialize_lvl01(C._ptr(bk), bk.size, 1, C._ptr(ksk), ksk.size),
            "initialize_lvl01")

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

    # ---- autotune jobs per workgroup
    cands = [args.jobs_per_group] if args.jobs_per_group else [1, 2, 4, 8]
    tune = {}
    for c in cands:
        plan = C.Plan(gates, dpool, jobs_per_group=c)
        plan.run_timed()
        t = [plan.run_timed()[1] for _ in range(2)]
        tune[c] = min(t)
        log("autotune C=%d: blind rotation %.2f ms" % (c, tune[c]))
        plan.destroy()
    best_c = min(tune, key=tune.get)

    plan = C.Plan(gates, dpool, jobs_per_group=best_c)
    nbr = plan.bootstraps
    log("warmup %d steps (C=%d, %d bootstraps/step)" % (args.warmup, best_c, nbr))
    for _ in range(args.warmup):
        plan.run_timed()
    out = dpool.download(np.uint32, (3 * B, C.TLWE1_WORDS))[2 * B:]
    dec = C.decrypt_bits(1, k1, out)
    failures = int(np.sum(dec != want))
    log("decryption failures after warmup: %d / %d" % (failures, B))

    # ---- timed region: exactly K steps, device synchronised on both sides
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    step_ms, ks_ms, br_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        k, b, _ = plan.run_timed()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        ks_ms.append(k)
        br_ms.append(b)
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    total = time.perf_counter() - t0
    value = B * args.steps / total
    ms_per_step = total * 1e3 / args.steps
    log("batched: %.1f gates/s, %.2f ms/step (ks %.2f ms, br %.2f ms)" %
        (value, ms_per_step, statistics.mean(ks_ms), statistics.mean(br_ms)))

    # graph replay of the same plan (launch overhead removed)
    gplan = C.Plan(gates, dpool, jobs_per_group=best_c, use_graph=True)
    gplan.run()
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    g0 = time.perf_counter()
    for _ in range(3):
        gplan.run()
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    graph_ms = (time.perf_counter() - g0) * 1e3 / 3
    gplan.destroy()

    # ---- plain per-gate form on a sample (one gate per launch, many streams)
    P = min(args.plain_sample, B)
    streams = [C.Stream() for _ in range(args.plain_streams)]
    W = C.TLWE1_WORDS * 4
    for i in range(min(8, P)):  # warm
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[ops[i]], dpool.at((2 * B + i) * W),
                                        dpool.at(i * W), dpool.at((B + i) * W), None,
                                        streams[i % len(streams)].handle), "gate")
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    p0 = time.perf_counter()
    for i in range(P):
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[ops[i]], dpool.at((2 * B + i) * W),
                                        dpool.at(i * W), dpool.at((B + i) * W), None,
                                        streams[i % len(streams)].handle), "gate")
    C.check(C.lib().cufhe_synchronize(), "synchronize")
    plain_tput = P / (time.perf_counter() - p0)
    lat = []
    for i in range(5):
        l0 = time.perf_counter()
        C.check(C.lib().cufhe_gate_lvl1(C.GATES[ops[i]], dpool.at((2 * B + i) * W),
                                        dpool.at(i * W), dpool.at((B + i) * W), None,
                                        streams[0].handle), "gate")
        streams[0].synchronize()
        lat.append((time.perf_counter() - l0) * 1e3)
    for s in streams:
        s.destroy()
    log("plain form: %.1f gates/s over %d streams; single-gate latency p50 %.2f ms" %
        (plain_tput, len(streams), pct(lat, 50)))

    # ---- roofline of the dominant kernel (blind rotation)
    br_avg_s = statistics.mean(br_ms) / 1e3
    achieved = FLOP_PER_BOOTSTRAP_KB * nbr / br_avg_s / 1e12
    groups = -(-nbr // best_c)
    key_gbps = BK_BYTES_KB * groups / br_avg_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "br_kernel_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # ---- CPU baseline: the oracle (FP64 FFNT mirror + integer KS) on host cores
    cpu = None
    if not args.no_cpu:
        import oracle as O
        nthr = args.cpu_threads or min(16, os.cpu_count() or 1)
        done = 0
        c0 = time.perf_counter()
        while time.perf_counter() - c0 < args.cpu_seconds and done < B:
            m = min(nthr, B - done)
            opcodes = np.array([C.GATES[o] for o in ops[done:done + m]], np.int32)
            O.gate_batch_lvl1(O.FFT, True, opcodes, pool[done:done + m],
                              pool[B + done:B + done + m], bk, ksk, nthr)
            done += m
        cpu_s = time.perf_counter() - c0
        cpu = {"value": round(done / cpu_s, 3), "unit": "gates/s", "cores": nthr, "kind": "port",
               "sample": "%d gates of the same mix, own C++ CPU restatement (FP64 FFNT, integer "
                         "key switch), one gate per thread; not TFHEpp" % done}
        log("cpu baseline: %.2f gates/s on %d threads" % (cpu["value"], nthr))

    res = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched lvl1 binary gates, KB=2, FFT, 4096 gates uniform over "
                        "NAND/AND/OR/XOR/XNOR/NOR, inputs resident in HBM",
            "global_batch": B, "n": 636, "N": 1024, "l": 2, "Bgbit": 8,
            "jobs_per_group": best_c, "autotune_br_ms": {str(k): round(v, 3) for k, v in tune.items()},
            "parallelism": "single GPU",
        },
        "p50_layer_ms": round(pct(step_ms, 50), 3),
        "p99_layer_ms": round(pct(step_ms, 99), 3),
        "graph_ms_per_step": round(graph_ms, 3),
        "kernel_ms": {"key_switch": round(statistics.mean(ks_ms), 3),
                      "blind_rotate": round(statistics.mean(br_ms), 3)},
        "plain_form": {"gates_per_s": round(plain_tput, 2), "streams": args.plain_streams,
                       "sample": P, "latency_p50_ms": round(pct(lat, 50), 3)},
        "batched_over_plain": round(value / plain_tput, 2),
        "decrypt_failures": failures,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic},
        "key_stream": {"algorithmic_GBps": round(key_gbps, 1), "peak_GBps": HBM_PEAK_GBPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
