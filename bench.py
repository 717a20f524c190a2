#!/usr/bin/env python3
"""Benchmark of the MI355X Groth16 prover back end (libg16hip).

Headline metric: Groth16 proofs/s with p50/p99 latency at the 2^20-constraint class, plus
G1 MSM points/s at 2^20. A step is one full prove (stages 0-11: gather, six NTTs, H, five
MSMs, host assembly with fixed blinders r, s) of a synthetic BN254 circuit of 1,100,000
constraints (domain 2^21, ~1.2M wires, 3 nonzeros per A/B row, 40% boolean witness), with the
proving key resident in HBM (prepared outside the timed region, the reference's "warm"
mode, bin/snarkrs/src/bench.rs:153-190) and the witness in host memory, packed into
page-locked staging and uploaded inside every timed proof (bench.rs:475-478,
crates/cuda/src/stages.rs:946-985). The same proof with the witness already in HBM is
reported beside it (`device_witness`).

The CPU baseline is the repository's CPU oracle (the reference CPU back end's algorithm
restated in C++; the Rust original cannot be built here) proving the same key and witness on
every host core the process may use, median of --cpu-reps proofs, each compared with the
GPU's proof bit for bit.

    python bench.py [--steps K] [--warmup W] [--workload large|plain] [--no-cpu] [--no-msm]
    python bench.py --mode batch   # 64 proofs over 4 circuits of 2^14..2^20 constraints
    python bench.py --mode serve   # Poisson serving over 8 circuits of 2^16..2^22
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
# MI355X VALU 32x32->64 multiply-add peak: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz, one
# v_mad_u64_u32 per lane per cycle (a dependent-chain probe measures 2.99e13/s, 76% of it).
PEAK_MUL32_PER_S = 256 * 4 * 16 * 2.4e9
# Algorithmic work per bucket addition (mixed add, madd-2008-s): 10 Fq products (G1),
# 28 (G2: 3 Fq products per Fq2 product); one Fq product = 136 32x32 partial products (CIOS).
MUL32_PER_MADD = {1: 10 * 136, 2: 28 * 136}
# HBM bytes per k_accumulate<Fm> launch in the 1.1M-constraint prove (FETCH_SIZE + WRITE_SIZE,
# scaled by the counters' reading of a 256 MiB device copy: fetch x2, write x1), from the
# rocprofv3 --pmc runs in profiles/k_accumulate_traffic_1_1M.txt.
ACCUMULATE_TRAFFIC_BYTES_1_1M = 3.39e9
# Reference, CUDA back end on a Tesla T4: single prove, synthetic key, domain 2^21, warm:
# 1263 ms (bench/results/cuda-size-sweep-2026-11-10.md:26).
REFERENCE_T4_PROVE_MS_2_21 = 1263.0
R_BLIND = 0x1234567890ABCDEF1234567890ABCDEF
S_BLIND = 0xFEDCBA0987654321FEDCBA0987654321


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def reference_window(m: int) -> int:
    """The reference's window model (crates/msm/src/lib.rs:160-172): the c in [3, 16]
    minimising ceil(255 / c) * (m + 3 * 2^(c - 1)) on the general scalar count m."""
    return min(range(3, 17), key=lambda c: (-(-255 // c)) * (m + 3 * (1 << (c - 1))))


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(key, wit, gpu_proof, reps: int):
    """The oracle's full prove of the same key and witness on every host core this process
    may run on (its CPU affinity), median of `reps` proofs."""
    from oracle.oracle import Oracle, build

    build()
    orc = Oracle()
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    orc.set_threads(usable)
    times, same = [], True
    for _ in range(reps):
        t0 = time.perf_counter()
        _, proof = orc.prove(key, wit, R_BLIND, S_BLIND)
        times.append(time.perf_counter() - t0)
        same = same and bool((proof == gpu_proof).all())
    dt = float(np.median(times))
    return {
        "value": 1.0 / dt,
        "unit": "proofs/s",
        "cores": usable,
        "nproc": nproc,
        "model": _cpu_model(),
        "reps": reps,
        "kind": "port",
        "sample": f"median of {reps} full proves of the same workload (key, witness, r, s) by "
                  f"the C++ oracle (reference CPU algorithm restated) on {usable} host threads "
                  f"(the process's CPU affinity; nproc {nproc}): {dt:.2f} s per proof",
        "seconds": dt,
        "seconds_each": times,
        "proof_bit_identical_to_gpu": same,
    }


def _mixed_workloads(which: str):
    from snarkrs_amd import synth

    return synth.mixed_workloads(**(synth.BATCH_WORKLOADS if which == "batch"
                                    else synth.SERVE_WORKLOADS))


def _build_circuits(dev, workloads, witnesses_per_circuit: int):
    from snarkrs_amd.synth import Synth

    circs = []
    for i, wl in enumerate(workloads):
        t = time.time()
        syn = Synth(wl)
        key, wit = syn.key(dev)
        wits = [wit] + [syn.witness(wl.witness_seed + 1000 * (k + 1))
                        for k in range(witnesses_per_circuit - 1)]
        syn.close()
        circ = dev.prepare(key)
        circs.append((wl, circ, wits))
        log(f"circuit {i}: {wl.n_constraints} constraints, domain {key.domain_size}, "
            f"{key.n_vars} wires ({time.time() - t:.1f} s)")
    return circs


def _blinders(k: int):
    from snarkrs_amd.wire import fr_array

    return fr_array([R_BLIND + 7 * k]), fr_array([S_BLIND + 11 * k])


def run_batch(args) -> int:
    """64 proofs with independent witnesses over 4 resident circuits of 2^14..2^20
    constraints, proved as one batch (each circuit's proofs in waves proved together, the wave
    size chosen per circuit by the library's autotuner) and, for comparison, one at a time
    through the plain path, with the same host witnesses."""
    from snarkrs_amd import native
    from snarkrs_amd.native import Device

    dev = Device(0)
    wls = _mixed_workloads("batch")
    per = args.batch_proofs // len(wls)
    circs = _build_circuits(dev, wls, per)
    tuned = []
    for wl, circ, wits in circs:
        if args.batch_cap:
            circ.set_batch_cap(args.batch_cap)
            tuned.append({"best": args.batch_cap, "tried": {}})
        else:
            t = time.time()
            tuned.append(circ.batch_autotune(wits[0], 32))
            log(f"autotune {wl.n_constraints} constraints: {tuned[-1]} ({time.time() - t:.1f} s)")
    items = []
    for k in range(args.batch_proofs):
        wl, circ, wits = circs[k % len(circs)]
        r, s = _blinders(k)
        items.append((circ, wits[(k // len(circs)) % len(wits)], r, s))
    # warmup: one batch, and the plain path once per circuit
    native.prove_batch(items)
    for it in items[: len(circs)]:
        it[0].prove_with_blinders(it[1], it[2], it[3])
    t_batch = []
    for _ in range(args.batch_reps):
        t0 = time.perf_counter()
        got = native.prove_batch(items)
        t_batch.append(time.perf_counter() - t0)
    t_b = float(np.median(t_batch))
    t0 = time.perf_counter()
    plain = [c.prove_with_blinders(w, r, s) for c, w, r, s in items]
    t_plain = time.perf_counter() - t0
    per_circuit_plain_ms = []
    for wl, circ, wits in circs:
        r, s = _blinders(0)
        ts = time.perf_counter()
        for _ in range(4):
            circ.prove_with_blinders(wits[0], r, s)
        per_circuit_plain_ms.append((time.perf_counter() - ts) / 4 * 1e3)
    same = all((a == b).all() for a, b in zip(got, plain))
    log(f"batch: {len(items)} proofs in {t_b * 1e3:.1f} ms (median of {args.batch_reps}) -> "
        f"{len(items) / t_b:.1f} proofs/s; plain path {t_plain * 1e3:.1f} ms -> "
        f"{len(items) / t_plain:.1f} proofs/s; bit-identical {same}")
    verified = _verify_all(items, got)
    log(f"every batch proof verifies: {verified}")
    line = {
        "metric": "Groth16 batch proofs/s (64 proofs, 4 resident circuits of 2^14..2^20 constraints)",
        "value": len(items) / t_b,
        "unit": "proofs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "batch of independent proofs over resident circuits",
            "circuits": [{"constraints": wl.n_constraints, "domain": c.domain_size,
                          "wave_cap": tu["best"], "autotune_proofs_per_s": tu["tried"],
                          "plain_ms_per_proof": pm}
                         for (wl, c, _), tu, pm in zip(circs, tuned, per_circuit_plain_ms)],
            "proofs": len(items),
            "witness": "host memory (staged and uploaded inside the timed region, both paths)",
            "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        },
        "batch_ms": t_b * 1e3,
        "batch_ms_each": [t * 1e3 for t in t_batch],
        "plain_path": {"value": len(items) / t_plain, "ms": t_plain * 1e3},
        "speedup_vs_plain": t_plain / t_b,
        "bit_identical_to_plain": same,
        "every_proof_verifies": verified,
    }
    print(json.dumps(line), flush=True)
    return 0 if same and verified else 1


def _verify_all(items, proofs) -> bool:
    """Each proof checked against its circuit's verifying key (host pairing verifier)."""
    from snarkrs_amd import native
    from snarkrs_amd.verify import VerifyingKey, verify

    vks = {}
    for (circ, w, _, _), proof in zip(items, proofs):
        if id(circ) not in vks:
            vks[id(circ)] = VerifyingKey.from_bytes(circ.vk_bytes())
        vk = vks[id(circ)]
        try:
            verify(vk, np.asarray(w).reshape(-1, 32)[1:1 + vk.n_public], proof)
        except native.VerifyError:
            return False
    return True


def run_serve(args) -> int:
    """Poisson arrivals over 8 resident circuits of 2^16..2^22 constraints: the serving loop
    (continuous batching, slots admitted by free device memory) against the plain path (one
    proof at a time, first come first served) on the same arrival schedule."""
    import math
    import threading

    from snarkrs_amd import native
    from snarkrs_amd.native import Device
    from snarkrs_amd.wire import SplitMix64

    dev = Device(0)
    wls = _mixed_workloads("serve")
    circs = _build_circuits(dev, wls, 2)
    rng = SplitMix64(41)
    reqs = []
    for k in range(args.requests):
        ci = rng.next_u64() % len(circs)
        wl, circ, wits = circs[ci]
        r, s = _blinders(k)
        reqs.append((ci, circ, wits[k % len(wits)], r, s))
    # service time of each circuit on the plain path (after one warm proof each)
    svc = []
    for wl, circ, wits in circs:
        r, s = _blinders(0)
        circ.prove_with_blinders(wits[0], r, s)
        t0 = time.perf_counter()
        circ.prove_with_blinders(wits[0], r, s)
        svc.append(time.perf_counter() - t0)
    mean_svc = sum(svc[ci] for ci, *_ in reqs) / len(reqs)
    # unit-rate exponential gaps, scaled to each load's rate
    gaps = [-math.log((rng.next_u64() + 1) / 2.0**64) for _ in reqs]

    def plain_path(arrivals):
        lat, out = [], []
        t0 = time.perf_counter()
        for (ci, circ, w, r, s), a in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if now < a:
                time.sleep(a - now)
            out.append(circ.prove_with_blinders(w, r, s))
            lat.append(time.perf_counter() - t0 - a)
        return lat, out, time.perf_counter() - t0

    def served(arrivals):
        srv = native.Server(dev, reserve_bytes=args.reserve_gib << 30, lanes=args.lanes)
        tickets = [None] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for k, ((ci, circ, w, r, s), a) in enumerate(zip(reqs, arrivals)):
                now = time.perf_counter() - t0
                if now < a:
                    time.sleep(a - now)
                tickets[k] = srv.submit(circ, w, r, s)

        th = threading.Thread(target=submitter)
        th.start()
        lat, out = [], []
        for k in range(len(reqs)):
            while tickets[k] is None:
                time.sleep(0.0005)
            p, ms = srv.wait(tickets[k])
            out.append(p)
            lat.append(ms / 1e3)
        th.join()
        wall = time.perf_counter() - t0
        srv.close()
        return lat, out, wall

    res = lambda lat, wall: {"proofs_per_s": len(reqs) / wall,
                             "p50_ms": percentile(lat, 50) * 1e3,
                             "p99_ms": percentile(lat, 99) * 1e3}
    runs, same = [], True
    for load in [float(x) for x in args.loads.split(",")]:
        rate = load / mean_svc
        arrivals, t = [], 0.0
        for g in gaps:
            t += g / rate
            arrivals.append(t)
        log(f"serve: {len(reqs)} requests, mean plain service {mean_svc * 1e3:.1f} ms, Poisson "
            f"rate {rate:.1f}/s (load {load} of the plain path)")
        lat_p, out_p, wall_p = plain_path(arrivals)
        lat_s, out_s, wall_s = served(arrivals)
        same = same and all((a == b).all() for a, b in zip(out_p, out_s))
        sv, pl = res(lat_s, wall_s), res(lat_p, wall_p)
        log(f"load {load}: served {sv}; plain path {pl}")
        runs.append({"load_of_plain_path": load, "rate_per_s": rate, "served": sv,
                     "plain_path": pl})
    verified = _verify_all([(circ, w, r, s) for _, circ, w, r, s in reqs], out_s)
    log(f"bit-identical {same}; every served proof verifies: {verified}")
    # Past capacity both paths run flat out: the highest load's throughputs are the saturation
    # figures (the plain path serves one request at a time, first come first served).
    top = max(runs, key=lambda x: x["load_of_plain_path"])
    saturation = {"load_of_plain_path": top["load_of_plain_path"],
                  "served_proofs_per_s": top["served"]["proofs_per_s"],
                  "plain_proofs_per_s": top["plain_path"]["proofs_per_s"],
                  "ratio": top["served"]["proofs_per_s"] / top["plain_path"]["proofs_per_s"]}
    saturation["target_3x_met"] = saturation["ratio"] >= 3.0
    log(f"saturation at load {top['load_of_plain_path']}: served/plain = {saturation['ratio']:.2f}")
    line = {
        "metric": "Groth16 serving proofs/s + p50/p99 latency (Poisson arrivals, 8 resident "
                  "circuits of 2^16..2^22 constraints)",
        "value": runs[-1]["served"]["proofs_per_s"],
        "unit": "proofs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "Poisson arrivals, circuit drawn uniformly per request; value = the "
                        "highest load's served proofs/s (saturation: arrivals at 4x the plain "
                        "path's capacity)",
            "circuits": [{"constraints": wl.n_constraints, "domain": c.domain_size,
                          "plain_service_ms": t * 1e3} for (wl, c, _), t in zip(circs, svc)],
            "requests": len(reqs),
            "lanes": args.lanes,
            "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "graph_replay": os.environ.get("G16HIP_GRAPHS") == "1",
        },
        "by_load": runs,
        "saturation": saturation,
        "bit_identical_to_plain": same,
        "every_proof_verifies": verified,
    }
    print(json.dumps(line), flush=True)
    return 0 if same and verified else 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["large", "plain"], default="large")
    ap.add_argument("--msm-log-n", type=int, default=20)
    ap.add_argument("--msm-reps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5, help="oracle proves timed (median)")
    ap.add_argument("--no-msm", action="store_true")
    ap.add_argument("--mode", choices=["prove", "batch", "serve"], default="prove",
                    help="prove: the single-proof headline (default); batch: 64 proofs over 4 "
                         "circuits; serve: Poisson serving over 8 circuits")
    ap.add_argument("--batch-proofs", type=int, default=64)
    ap.add_argument("--slots", type=int, default=4, help="proof slots per circuit (serve)")
    ap.add_argument("--batch-cap", type=int, default=0,
                    help="proofs per wave for every circuit (batch; 0 = autotune per circuit)")
    ap.add_argument("--batch-reps", type=int, default=3, help="timed batches (median)")
    ap.add_argument("--lanes", type=int, default=4, help="proofs in flight (serve)")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--loads", default="0.8,1.25,2,3,4",
                    help="arrival rates as multiples of the plain path's capacity (serve)")
    ap.add_argument("--reserve-gib", type=int, default=16)
    ap.add_argument("--constant-work", action="store_true",
                    help="constant-work MSMs (every digit an entry), opt in")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for this process (batch, serve; 0 = leave as set)")
    args = ap.parse_args()
    if args.hw_queues:
        # HIP reads it when the runtime starts, which has not happened yet in this process.
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one, and this prover back end "
              "runs one proof per device (multi-GPU sharding is described in DESIGN.md, not "
              "built). Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)

    from snarkrs_amd.devmem import DeviceArray, synchronize
    from snarkrs_amd.native import Device, Timings
    from snarkrs_amd.synth import LARGE_1_1M, PLAIN_2_16, Synth, scalars
    from snarkrs_amd.wire import R_MOD, fr_array

    dev = Device(0)
    if args.constant_work:
        dev.set_constant_work(True)
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

    # The headline is the reference's warm rep (bin/snarkrs/src/bench.rs:475-478): the witness
    # in host memory, staged through page-locked memory and uploaded inside every timed proof.
    for i in range(args.warmup):
        circ.prove_with_blinders(wit, r, s)
    log(f"warmup: {args.warmup} proofs")
    dev.kernel_stats_reset()
    stage = Timings()
    lat = []
    proof = None
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        proof = circ.prove_with_blinders(wit, r, s, stage)
        lat.append(time.perf_counter() - ts)
    synchronize()
    total = time.perf_counter() - t0
    ks = dev.kernel_stats()
    proofs_per_s = args.steps / total
    log(f"timed (host witness): {args.steps} proofs in {total:.3f} s -> {proofs_per_s:.2f} "
        f"proofs/s, p50 {percentile(lat, 50) * 1e3:.1f} ms")

    # The same proof with the witness already in HBM (no upload), for comparison.
    for i in range(args.warmup):
        circ.prove_with_blinders_device(d_wit.addr, circ.n_vars, r, s)
    lat_d = []
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        proof_d = circ.prove_with_blinders_device(d_wit.addr, circ.n_vars, r, s)
        lat_d.append(time.perf_counter() - ts)
    synchronize()
    total_d = time.perf_counter() - t0
    assert (proof_d == proof).all(), "device-witness proof differs from the host-witness proof"
    device_witness = {"proofs_per_s": args.steps / total_d,
                      "p50_ms": percentile(lat_d, 50) * 1e3,
                      "p99_ms": percentile(lat_d, 99) * 1e3}
    log(f"timed (device witness): {args.steps / total_d:.2f} proofs/s")

    # Every timed proof is the same (same witness, r and s): check it once against the
    # circuit's verifying key with the host pairing verifier, outside the timed region.
    from snarkrs_amd.verify import VerifyingKey, verify
    tv = time.perf_counter()
    verify(VerifyingKey.from_bytes(circ.vk_bytes()),
           np.asarray(wit).reshape(-1, 32)[1:1 + key.n_public], proof)
    verify_ms = (time.perf_counter() - tv) * 1e3
    log(f"the proof verifies (host pairing check, {verify_ms:.1f} ms)")

    # Roofline of the dominant kernel, the G1 bucket accumulation, over the timed proofs.
    launches, entries, acc_ms = ks["launches"][0], ks["entries"][0], ks["accumulate_ms"][0]
    achieved = entries * MUL32_PER_MADD[1] / (acc_ms / 1e3) if acc_ms else 0.0
    roofline = {
        "kernel": "k_accumulate<Fm> (G1 bucket accumulation)",
        "bound": "valu",
        "achieved": achieved / 1e12,
        "peak": PEAK_MUL32_PER_S / 1e12,
        "unit": "Tmul32/s",
        "frac": achieved / PEAK_MUL32_PER_S,
        "traffic": ACCUMULATE_TRAFFIC_BYTES_1_1M if args.workload == "large" else None,
        "traffic_unit": "bytes per launch (measured, rocprofv3 --pmc)",
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
        # The sort's achieved HBM rate, on the algorithmic bytes of the reference model:
        # 32 n + 16 m W + 12 W 2^(c*-1) with c* the reference window on the m general scalars.
        one = (pow(2, 256, R_MOD)).to_bytes(32, "little")
        rows = sc.reshape(-1, 32)
        zero_or_one = int((~rows.any(axis=1)).sum() +
                          (rows == np.frombuffer(one, np.uint8)).all(axis=1).sum())
        m = n - zero_or_one
        c_ref = reference_window(m)
        w_ref = -(-255 // c_ref)
        sort_bytes = 32 * n + 16 * m * w_ref + 12 * w_ref * (1 << (c_ref - 1))
        sort_ms = msm["phase_ms"]["sort_ms"]
        msm["sort_hbm"] = {"algorithmic_bytes": sort_bytes, "general_scalars": m,
                           "reference_c": c_ref, "ms": sort_ms,
                           "GB_per_s": sort_bytes / (sort_ms / 1e3) / 1e9,
                           "frac_of_8TB_per_s": sort_bytes / (sort_ms / 1e3) / 8e12}
        log(f"sort: {sort_bytes / 1e6:.0f} MB algorithmic in {sort_ms:.3f} ms -> "
            f"{msm['sort_hbm']['GB_per_s']:.0f} GB/s")

    msm_g2 = None
    if not args.no_msm:
        # G2 MSM at 2^18 (bases k_i G2 from the GPU fixed-base setup, same scalar recipe).
        n2 = 1 << 18
        bases2 = dev.fixed_base(2, scalars(4343, n2, 0))
        sc2 = scalars(2121, n2, 300)
        d_b2, d_s2 = DeviceArray(bases2), DeviceArray(sc2)
        dev.msm_device(2, d_b2.addr, d_s2.addr, n2)
        times2, stats2 = [], []
        for _ in range(args.msm_reps):
            synchronize()
            ts = time.perf_counter()
            dev.msm_device(2, d_b2.addr, d_s2.addr, n2)
            times2.append(time.perf_counter() - ts)
            stats2.append(dev.msm_stats())
        med2 = float(np.median(times2))
        msm_g2 = {
            "n": n2, "points_per_s": n2 / med2, "ms_median": med2 * 1e3, "c": stats2[-1]["c"],
            "entries": stats2[-1]["entries"],
            "phase_ms": {k: float(np.median([st[k] for st in stats2]))
                         for k in ("sort_ms", "accumulate_ms", "merge_ms", "reduce_ms", "total_ms")},
        }
        log(f"MSM G2 2^18: {med2 * 1e3:.2f} ms -> {n2 / med2 / 1e6:.1f} M points/s")

    ntt = None
    if not args.no_msm:
        # NTT over 2^21 resident Fr: one forward transform (64 B per element algorithmic: one
        # read and one write) and the prover's coset pipeline (two transforms), timed around the
        # synchronised library call.
        from snarkrs_amd.devmem import synchronize as _sync
        log_n = 21
        d_x = DeviceArray(scalars(77, 1 << log_n, 0))
        ntt = {}
        for kind, name, passes in ((0, "forward", 1), (2, "coset_pipeline", 2)):
            dev.ntt_device(d_x.addr, log_n, kind)
            ts = []
            for _ in range(10):
                _sync()
                t0 = time.perf_counter()
                dev.ntt_device(d_x.addr, log_n, kind)
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            b = 64 * (1 << log_n) * passes
            ntt[name] = {"log_n": log_n, "ms": t * 1e3, "algorithmic_bytes": b,
                         "GB_per_s": b / t / 1e9, "frac_of_8TB_per_s": b / t / 8e12}
            log(f"NTT 2^{log_n} {name}: {t * 1e3:.3f} ms -> {b / t / 1e9:.0f} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_reps} oracle proves on the host cores ...")
        cpu = cpu_baseline(key, wit, proof, args.cpu_reps)
        log(f"CPU baseline: {cpu['seconds']:.2f} s per proof (median) on {cpu['cores']} threads, "
            f"bit-identical: {cpu['proof_bit_identical_to_gpu']}")

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
                         "in HBM, witness in host memory (packed into page-locked staging and "
                         "uploaded inside every timed proof)"),
            "constraints": wl.n_constraints,
            "domain": key.domain_size,
            "n_vars": key.n_vars,
            "witness": "host",
            "msm_work": "constant" if args.constant_work else "variable",
        },
        "latency_ms": {"p50": percentile(lat, 50) * 1e3, "p99": percentile(lat, 99) * 1e3,
                       "min": min(lat) * 1e3, "max": max(lat) * 1e3},
        "device_witness": device_witness,
        # Per-stage spans from HIP events, summed per proof. Stages 5-9 run on other streams
        # beside stages 0-4, so the spans overlap: they do not add up to the proof's time.
        "stage_spans_us_per_proof_overlapping": {k: v / args.steps
                                                 for k, v in stage.as_dict().items()},
        "proof_verifies": True,
        "verify_ms": verify_ms,
        "msm_g1": msm,
        "msm_g2": msm_g2,
        "ntt": ntt,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
