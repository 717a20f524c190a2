#!/usr/bin/env python3
"""Benchmark of the MI355X Groth16 prover back end (libg16hip).

Headline metric: Groth16 proofs/s with p50/p99 latency at the 2^20-constraint class, plus
G1 MSM points/s at 2^20. A step is one full prove (stages 0-11: gather, six NTTs, H, five
MSMs, host assembly with fixed blinders r, s) of a synthetic BN254 circuit of 1,100,000
constraints (domain 2^21, ~1.2M wires, 3 nonzeros per A/B row, 40% boolean witness), with the
proving key resident in HBM (prepared outside the timed region, the reference's "warm"
mode, bin/snarkrs/src/bench.rs:153-190) and the witness already in device memory.

The CPU baseline is the repository's CPU oracle (the reference CPU back end's algorithm
restated in C++; the Rust original cannot be built here) proving the same key and witness on
the host cores, and its proof is compared with the GPU's bit for bit.

    python bench.py [--steps K] [--warmup W] [--workload large|plain] [--no-cpu] [--no-msm]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Groth16 proofs/sec + p50/p99 latency at 2^20 constraints; G1 MSM points/sec"
# MI355X VALU 32x32->64 multiply-add peak: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz at quarter
# rate (v_mad_u64_u32).
PEAK_MUL32_PER_S = 256 * 4 * 16 * 2.4e9 / 4
# Algorithmic work per bucket addition (mixed add, madd-2008-s): 10 Fq products (G1),
# 28 (G2: 3 Fq products per Fq2 product); one Fq product = 136 32x32 partial products (CIOS).
MUL32_PER_MADD = {1: 10 * 136, 2: 28 * 136}
# Reference, CUDA back end on a Tesla T4: single prove, synthetic key, domain 2^21, warm:
# 1263 ms (bench/results/cuda-size-sweep-2026-11-10.md:26).
REFERENCE_T4_PROVE_MS_2_21 = 1263.0
R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def cpu_baseline(key, wit, gpu_proof, threads):
    """The oracle's full prove of the same key and witness on the host cores (one proof)."""
    from oracle.oracle import Oracle, build

    build()
    orc = Oracle()
    orc.set_threads(threads)
    t0 = time.perf_counter()
    _, proof = orc.prove(key, wit, R_BLIND, S_BLIND)
    dt = time.perf_counter() - t0
    return {
        "value": 1.0 / dt,
        "unit": "proofs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one full prove of the same workload (key, witness, r, s) by the C++ oracle "
                  f"(reference CPU algorithm restated) on {threads} host threads: {dt:.2f} s",
        "seconds": dt,
        "proof_bit_identical_to_gpu": bool((proof == gpu_proof).all()),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["large", "plain"], default="large")
    ap.add_argument("--msm-log-n", type=int, default=20)
    ap.add_argument("--msm-reps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-msm", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one, and this prover back end "
              "runs one proof per device (multi-GPU sharding is described in DESIGN.md, not "
              "built). Run with --gpus 1.", file=sys.stderr)
        return 2

    from snarkrs_amd.devmem import DeviceArray, synchronize
    from snarkrs_amd.native import Device, Timings
    from snarkrs_amd.synth import LARGE_1_1M, PLAIN_2_16, Synth, scalars
    from snarkrs_amd.wire import fr_array

    dev = Device(0)
    wl = LARGE_1_1M if args.workload == "large" else PLAIN_2_16
    t = time.time()
    syn = Synth(wl)
    log(f"circuit: {syn.n_constraints} constraints, {syn.n_vars} wires, domain {syn.domain}, "
        f"nnz A/B {syn.nnz_a}/{syn.nnz_b} ({time.time() - t:.1f} s)")
    t = time.time()
    key, wit = syn.key(dev)
    syn.close()
    log(f"proving key built (GPU fixed-base setup) in {time.time() - t:.1f} s")
    t = time.time()
    circ = dev.prepare(key)
    log(f"prepared: {circ.resident_bytes() / 2**30:.2f} GiB resident ({time.time() - t:.1f} s)")
    d_wit = DeviceArray(wit)
    r, s = fr_array([R_BLIND]), fr_array([S_BLIND])

    for i in range(args.warmup):
        circ.prove_with_blinders_device(d_wit.addr, circ.n_vars, r, s)
    log(f"warmup: {args.warmup} proofs")
    dev.kernel_stats_reset()
    stage = Timings()
    lat = []
    proof = None
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        proof = circ.prove_with_blinders_device(d_wit.addr, circ.n_vars, r, s, stage)
        lat.append(time.perf_counter() - ts)
    synchronize()
    total = time.perf_counter() - t0
    ks = dev.kernel_stats()
    proofs_per_s = args.steps / total
    log(f"timed: {args.steps} proofs in {total:.3f} s -> {proofs_per_s:.2f} proofs/s, "
        f"p50 {percentile(lat, 50) * 1e3:.1f} ms")

    # Roofline of the dominant kernel, the G1 bucket accumulation, over the timed proofs.
    launches, entries, acc_ms = ks["launches"][0], ks["entries"][0], ks["accumulate_ms"][0]
    achieved = entries * MUL32_PER_MADD[1] / (acc_ms / 1e3) if acc_ms else 0.0
    roofline = {
        "kernel": "k_accumulate<Fq> (G1 bucket accumulation)",
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": PEAK_MUL32_PER_S / 1e12,
        "unit": "Tmul32/s",
        "frac": achieved / PEAK_MUL32_PER_S,
        "traffic": None,
        "launches": launches,
        "avg_launch_ms": acc_ms / launches if launches else None,
        "entries_per_launch": entries / launches if launches else None,
        "g2_avg_launch_ms": (ks["accumulate_ms"][1] / ks["launches"][1]) if ks["launches"][1] else None,
    }

    msm = None
    if not args.no_msm:
        n = 1 << args.msm_log_n
        t = time.time()
        ks_bases = scalars(4242, n, 0)
        bases = dev.fixed_base(1, ks_bases)
        sc = scalars(2020, n, 300)
        d_b, d_s = DeviceArray(bases), DeviceArray(sc)
        log(f"MSM inputs: 2^{args.msm_log_n} G1 bases k_i G, 30% of scalars in {{0,1}} "
            f"({time.time() - t:.1f} s)")
        dev.msm_device(1, d_b.addr, d_s.addr, n)
        times, stats = [], []
        for _ in range(args.msm_reps):
            synchronize()
            ts = time.perf_counter()
            dev.msm_device(1, d_b.addr, d_s.addr, n)
            times.append(time.perf_counter() - ts)
            stats.append(dev.msm_stats())
        med = float(np.median(times))
        msm = {
            "n": n,
            "points_per_s": n / med,
            "ms_median": med * 1e3,
            "c": stats[-1]["c"],
            "entries": stats[-1]["entries"],
            "phase_ms": {k: float(np.median([st[k] for st in stats]))
                         for k in ("sort_ms", "accumulate_ms", "merge_ms", "reduce_ms", "total_ms")},
        }
        log(f"MSM G1 2^{args.msm_log_n}: {med * 1e3:.2f} ms -> {n / med / 1e6:.1f} M points/s")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU baseline: oracle prove on {threads} threads ...")
        cpu = cpu_baseline(key, wit, proof, threads)
        log(f"CPU baseline: {cpu['seconds']:.2f} s per proof, bit-identical: "
            f"{cpu['proof_bit_identical_to_gpu']}")

    line = {
        "metric": METRIC,
        "value": proofs_per_s,
        "unit": "proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (proofs_per_s * REFERENCE_T4_PROVE_MS_2_21 / 1e3
                        if args.workload == "large" else None),
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": ("single Groth16 prove, BN254, synthetic R1CS of "
                         f"{wl.n_constraints} constraints (domain {key.domain_size}), "
                         f"{key.n_vars} wires, {wl.terms} nonzeros per A/B row, "
                         f"{wl.bool_permille / 10:.0f}% boolean witness, fixed r,s, key resident "
                         "in HBM, witness resident in HBM"),
            "constraints": wl.n_constraints,
            "domain": key.domain_size,
            "n_vars": key.n_vars,
        },
        "latency_ms": {"p50": percentile(lat, 50) * 1e3, "p99": percentile(lat, 99) * 1e3,
                       "min": min(lat) * 1e3, "max": max(lat) * 1e3},
        "stage_us_per_proof": {k: v / args.steps for k, v in stage.as_dict().items()},
        "msm_g1": msm,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
