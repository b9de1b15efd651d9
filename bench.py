#!/usr/bin/env python3
"""Benchmark of the k-mer counting path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2] [--partition both|prefix|learned]

A step is one pass of the whole device pipeline (extraction, partition,
Bloom split, repeat counting, record grouping) over the resident synthetic
input of the workload; the sequence is generated on the host and copied to HBM
before timing.  By default both partition methods run on the same input: the
fixed 256-prefix split, then the learned CDF partition (its network trained on
the GPU from the first 1% of the reads); `value` is the learned path's
k-mers/s over the wall clock of the timed steps, and `learned_over_prefix`
compares the two.  Workload 2 is the at-scale plain path: a 100 Mbp genome at
45% GC with 10% of it in 40 families x 50 copies of 5 kbp repeats, 20M reads
of 150 bp, 0.5% substitutions, 0.1% N, k=31 (about 2.3 billion k-mers).

Prints one JSON line (see the keys below).  The device time of each kernel
comes from HIP events the engine records on its own stream around each launch
(pipeline.phase_ms); roofline.achieved is the dominant kernel's algorithmic
bytes over that time.
cpu_baseline times the reference `Tree` binary (oracle/_ref, compiled from the
reference sources) on a bounded sample of the same workload on the host cores;
where that binary is absent it times the oracle's C restatement instead.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import kmertreecounter_amd as K  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "k-mers counted/s at k=31, learned vs prefix partition; % HBM peak"

WORKLOADS = {
    1: ("2 Mbp uniform genome, 400k x 150 bp reads, 1% substitutions, k=31", 31),
    2: ("100 Mbp genome at 45% GC with 10% in 50-copy 5 kbp repeat families, 20M x 150 bp reads, "
        "0.5% substitutions, 0.1% N, Phred 20-40, k=31", 31),
    3: ("100 Mbp genome at 65% GC plus 5% period 1-6 microsatellite tracts, 20M x 150 bp reads, "
        "0.5% substitutions, k=31", 31),
    4: ("200 Mbp uniform genome, 30x of 150 bp reads plus 10x lognormal long reads (median 8 kbp) "
        "with 5% sub/ins/del errors, k=31", 31),
    5: ("200 genomes x 5 Mbp, GC 35-65%, lognormal abundance, 20 Gbp of 150 bp reads, 0.5% errors, k=63", 63),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(st, seq_bytes: int) -> dict:
    """Bytes each kernel must move at minimum (its inputs read once, outputs written once)."""
    W = 8 * st.key_words
    Kk, F, R = st.total_kmers, st.firsts, st.records
    RB = W + 8
    sub = 1 if st.fine_buckets > 256 else 0
    return {
        "extract_count": seq_bytes,                    # read the sequence
        "extract_scatter": seq_bytes + Kk * W,          # read the sequence, write each key once
        "sub_count": sub * Kk * W,                      # read the keys
        "sub_scatter": sub * 2 * Kk * W,                # read + write the keys
        "count": Kk * W + F * W + R * RB,               # read the keys, write firsts and records
        "group_records": 3 * R * RB,                    # count read, scatter read + write
    }


def phase_ms(st) -> dict:
    return {
        "extract_count": st.ms_extract_count, "extract_scatter": st.ms_extract_scatter,
        "sub_count": st.ms_sub_count, "sub_scatter": st.ms_sub_scatter,
        "count": st.ms_count, "group_records": st.ms_compact,
    }


PHASES = ("extract_count", "extract_scatter", "sub_count", "sub_scatter", "count", "group_records")


def kernel_names(nw: int) -> dict:
    return {
        "extract_count": f"k_extract<{nw}, false>", "extract_scatter": f"k_extract<{nw}, true>",
        "sub_count": f"k_sub<{nw}, false>", "sub_scatter": f"k_sub<{nw}, true>", "count": f"k_count<{nw}>",
        "group_records": f"k_spill_emit<{nw}> + k_seg_scan + k_compact_firsts<{nw}> + k_compact_recs<{nw}> + k_root_scan",
    }


def traffic_from_profiles(phase: str, workload: int, partition: str) -> float | None:
    """HBM bytes per launch of a phase from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by its reading for k_copy16), if present."""
    p = ROOT / "profiles" / f"traffic_w{workload}_{partition}.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())["bytes_per_launch"].get(phase)
    except Exception:
        return None


def cpu_baseline_reference(workload: int, k: int, threads: int) -> dict | None:
    tree = ROOT / "oracle" / "_ref" / "Tree"
    if not tree.exists():
        return None
    scale = {1: 1.0, 2: 5.0, 3: 5.0, 4: 25.0, 5: 50.0}[workload]
    fq = K.generate(workload, scale=scale, seed=1)
    seq = K.generate(workload, scale=scale, seed=1, fmt="seq")
    expect = count_kmers_host(seq, k)
    del seq
    base = "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None
    with tempfile.TemporaryDirectory(dir=base) as td:
        td = Path(td)
        (td / "in.fq").write_bytes(fq)
        del fq
        (td / "tmp").mkdir()
        for attempt in range(3):
            t0 = time.time()
            r = subprocess.run([str(tree), str(td / "in.fq"), str(k), str(threads), "32"], cwd=td,
                               capture_output=True, text=True, timeout=900)
            wall = time.time() - t0
            m = re.search(r"Total read k-mer count: (\d+)", r.stdout)
            got = int(m.group(1)) if m else -1
            if r.returncode == 0 and got == expect:
                break
            log(f"reference run {attempt} read {got} of {expect} k-mers; retrying")
        else:
            return None
        us = {key: int(v) for key, v in re.findall(r"(Init|Mid|Final|Total) elapsed us: (\d+)", r.stdout)}
    count_s = (us.get("Mid", 0) + us.get("Final", 0)) / 1e6
    return {
        "value": expect / count_s if count_s > 0 else None,
        "unit": "k-mers/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"workload {workload} at 1/{scale:g} scale ({expect} k-mers), reference Tree k={k} with "
                   f"{threads} threads, counting phase (Mid+Final) {count_s:.3f} s; Stage 0 (Init) "
                   f"{us.get('Init', 0) / 1e6:.3f} s, Total {us.get('Total', 0) / 1e6:.3f} s, wall {wall:.2f} s"),
    }


def count_kmers_host(seq: bytes, k: int) -> int:
    """Closed-form K (windows of k valid bases) via the oracle, for the baseline check."""
    O = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    O.orc_count_kmers.restype = ctypes.c_uint64
    O.orc_count_kmers.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32]
    return int(O.orc_count_kmers(seq, len(seq), k))


def cpu_baseline_port(workload: int, k: int) -> dict:
    """Single-thread oracle (extract + radix sort + run-length) on a bounded sample."""
    O = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    O.orc_extract.restype = ctypes.c_uint64
    O.orc_extract.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
    O.orc_count_sorted.restype = ctypes.c_uint64
    O.orc_count_sorted.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p]
    import numpy as np
    seq = K.generate(workload, scale=20.0, seed=1, fmt="seq")
    nw = K.words_for_k(k)
    keys = np.zeros(len(seq) * nw, dtype=np.uint64)
    t0 = time.time()
    n = O.orc_extract(seq, len(seq), k, keys.ctypes.data, len(seq))
    counts = np.zeros(max(n, 1), dtype=np.uint32)
    O.orc_count_sorted(keys.ctypes.data, n, nw, counts.ctypes.data)
    dt = time.time() - t0
    return {"value": n / dt, "unit": "k-mers/s", "cores": 1, "kind": "port",
            "sample": f"workload {workload} at 1/20 scale ({n} k-mers), oracle extract + sort + run-length, {dt:.2f} s"}


def run_partition(seq: bytes, k: int, partition: str, steps: int, warmup: int, counter: str, graph: bool,
                  autotune: bool = False) -> dict:
    """Resident run of one partition method: plan (learned only), copy, warm up, time `steps` passes."""
    part = K.KC_PARTITION_LEARNED if partition == "learned" else K.KC_PARTITION_PREFIX
    cnt = K.KC_COUNTER_BLOOM if counter == "bloom" else K.KC_COUNTER_EXACT
    eng = K.Engine(k, partition=part, expected_bytes=len(seq), counter=cnt, graph=graph)
    plan_ms = 0.0
    if part == K.KC_PARTITION_LEARNED:
        # Stage 0 sample: the first 1% of the reads; the network trains on the GPU
        t0 = time.time()
        eng.sample_block(seq[: len(seq) // 100])
        eng.plan()
        plan_ms = (time.time() - t0) * 1e3
    t0 = time.time()
    eng.push_block(seq)
    eng.finish()  # H2D copy + first full pass (also the first warmup)
    tuned = None
    if autotune:  # untimed: time the candidates once, keep the fastest
        t_t = time.perf_counter()
        r = eng.autotune()
        tuned = {"sub_bits": r.sub_bits, "counter": "exact" if r.counter == K.KC_COUNTER_EXACT else "bloom",
                 "heavy_keys": r.heavy_keys, "candidates": r.candidates, "best_ms": r.best_ms,
                 "tune_ms": (time.perf_counter() - t_t) * 1e3}
        log(f"[{partition}] autotune: {tuned}")
    st = eng.stats()
    log(f"[{partition}] resident in {time.time() - t0:.1f} s; K={st.total_kmers} firsts={st.firsts} "
        f"records={st.records} spilled={st.spilled_keys} fine_buckets={st.fine_buckets} "
        f"max/mean={st.max_bucket_keys / max(st.mean_bucket_keys, 1):.2f} extra_passes={st.extra_passes} "
        f"rounds={st.rounds}")
    for _ in range(max(0, warmup - 1)):
        eng.count_resident()
    per_phase = {name: 0.0 for name in PHASES}
    dev = []
    t_wall = time.perf_counter()
    for _ in range(steps):
        eng.count_resident()  # synchronous: returns after the device pipeline and its readback
        st = eng.stats()
        dev.append(st.ms_total)
        for name, v in phase_ms(st).items():
            per_phase[name] += v / steps
    wall_ms = (time.perf_counter() - t_wall) * 1e3 / steps
    log(f"[{partition}] {steps} steps: {wall_ms:.3f} ms/step wall, device " + ", ".join(f"{v:.3f}" for v in dev))
    # histogram_tool's precise histogram straight from the device tables (outside the timed steps)
    t_h = time.perf_counter()
    eng.histogram(1, 10000)
    hist_ms = (time.perf_counter() - t_h) * 1e3
    r = {
        "partition": partition,
        "ms_per_step": wall_ms,
        "device_ms_per_step": sum(dev) / len(dev),
        "kmers_per_s": st.total_kmers / (wall_ms * 1e-3),
        "stats": st,
        "per_phase": per_phase,
        "plan_ms": plan_ms,
        "max_over_mean_bucket": st.max_bucket_keys / max(st.mean_bucket_keys, 1e-9),
        "histogram_ms": hist_ms,
        "autotune": tuned,
    }
    if part == K.KC_PARTITION_LEARNED:
        r["learned"] = {
            "plan_ms": plan_ms,
            "train_keys": st.learned_train_keys,
            "split_buckets": st.learned_split_buckets,
            "train_ms": st.learned_train_ms,
            "loss_first": st.learned_loss_first,
            "loss_last": st.learned_loss_last,
            "sample_max_over_mean": {"learned": st.learned_sample_max_over_mean,
                                     "prefix": st.prefix_sample_max_over_mean},
            "sample_distinct_max_over_mean": {"learned": st.learned_sample_distinct_max_over_mean,
                                              "prefix": st.prefix_sample_distinct_max_over_mean},
        }
    eng.close()
    return r


def summary(r: dict) -> dict:
    st = r["stats"]
    return {
        "kmers_per_s": r["kmers_per_s"], "ms_per_step": r["ms_per_step"],
        "device_ms_per_step": r["device_ms_per_step"], "fine_buckets": st.fine_buckets,
        "max_over_mean_bucket": r["max_over_mean_bucket"], "extra_passes": st.extra_passes,
        "spilled_keys": st.spilled_keys, "heavy_buckets": st.heavy_buckets, "rounds": st.rounds,
        "phase_ms": r["per_phase"],
        "histogram_ms": r["histogram_ms"], "autotune": r["autotune"],
    }


def ingest_report(workload: int, scale: float, seed: int, k: int, threads: int) -> dict:
    """End to end from a FASTQ file (kmc_tree_run, histogram only, no output
    files): the host reader's blocks against raw chunks parsed on the device."""
    import shutil
    import tempfile
    fq = K.generate(workload, scale=scale, seed=seed, fmt="fastq", threads=threads)
    work = Path(tempfile.mkdtemp(prefix="kmc_ingest_"))
    try:
        path = work / "in.fq"
        path.write_bytes(fq)
        nbytes = len(fq)
        del fq
        res = {"fastq_bytes": nbytes}
        for name, host in (("host_reader", True), ("device_parser", False)):
            out = work / name
            out.mkdir()
            st, wall_ms = K.tree_run(str(path), str(out), k, histogram_only=True,
                                     histogram_path=str(out / "hist.tsv"), host_parse=host)
            res[name] = {"wall_s": wall_ms / 1e3, "device_count_ms": st.ms_total, "h2d_ms": st.ms_h2d,
                         "kmers": st.total_kmers}
            log(f"[ingest {name}] {nbytes / 1e9:.2f} GB FASTQ: wall {wall_ms / 1e3:.2f} s, "
                f"device counting {st.ms_total:.1f} ms, K={st.total_kmers}")
        return res
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=2, choices=sorted(WORKLOADS))
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--partition", choices=["prefix", "learned", "both"], default="both")
    ap.add_argument("--counter", choices=["bloom", "exact"], default="bloom")
    ap.add_argument("--autotune", action="store_true",
                    help="time the candidate fan-outs, counters and heavy-bucket thresholds first (untimed)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the pipeline as a captured HIP graph (per-kernel times are then not recorded)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ingest", action="store_true",
                    help="also time kmc_tree_run end to end from a FASTQ file, host reader vs device parser")
    ap.add_argument("--calibration-copy-gib", type=float, default=0.0,
                    help="after the run, copy this many GiB with k_copy16 (a plain copy whose "
                         "FETCH_SIZE/WRITE_SIZE readings scale those of the pipeline kernels)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (the k-mer exchange between "
              "GPUs is a design note only); run with --gpus 1", file=sys.stderr)
        return 2

    desc, k = WORKLOADS[args.workload]
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    seq = K.generate(args.workload, scale=args.scale, seed=args.seed, fmt="seq", threads=threads)
    seq_bytes = len(seq)
    log(f"generated {seq_bytes / 1e9:.2f} GB of sequence in {time.time() - t0:.1f} s")
    parts = ["prefix", "learned"] if args.partition == "both" else [args.partition]
    runs = {p: run_partition(seq, k, p, args.steps, args.warmup, args.counter, args.graph, args.autotune)
            for p in parts}
    del seq
    main_run = runs["learned"] if "learned" in runs else runs["prefix"]
    st = main_run["stats"]
    per_phase = main_run["per_phase"]
    names = kernel_names(st.key_words)

    alg = algorithmic_bytes(st, seq_bytes + (-(seq_bytes) % (512 * 16)))
    dominant = max(per_phase, key=per_phase.get)
    # a graph replay has no per-kernel times: no kernel roofline then (no
    # dominant kernel, no achieved rate, no traffic)
    timed = per_phase[dominant] > 0
    achieved = alg[dominant] / (per_phase[dominant] * 1e-3) / 1e9 if timed else None
    total_alg = sum(alg.values())
    traffic = traffic_from_profiles(dominant, args.workload, main_run["partition"]) if timed else None

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_reference(args.workload, k, max(6, threads))
        except Exception as e:  # a missing or failing reference binary falls back to the port
            log(f"reference baseline failed: {e}")
        if cpu is None:
            cpu = cpu_baseline_port(args.workload, k)

    out = {
        "metric": METRIC,
        "value": main_run["kmers_per_s"],
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if k <= 32 else "u128",
        "data": "synthetic (deterministic generator, seed %d)" % args.seed,
        "config": {
            "workload": desc + (f" at 1/{args.scale:g} scale" if args.scale != 1 else ""),
            "partition": main_run["partition"],
            "counter": args.counter,
            "graph": args.graph,
            "k": k,
            "kmers": st.total_kmers,
            "distinct_first": st.firsts,
            "root_records": st.records,
            "sequence_bytes": seq_bytes,
            "fine_buckets": st.fine_buckets,
            "max_over_mean_bucket": main_run["max_over_mean_bucket"],
            "spilled_keys": st.spilled_keys,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": names[dominant] if timed else None,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
            "traffic": traffic,
        },
        "pipeline": {
            "hbm_frac": total_alg / (main_run["device_ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes": total_alg,
            "device_ms_per_step": main_run["device_ms_per_step"],
            "phase_ms": per_phase,
            "phase_gbs": {n: (alg[n] / (per_phase[n] * 1e-3) / 1e9 if per_phase[n] > 0 else None) for n in alg},
            "kernels": names,
        },
        "partitions": {p: summary(r) for p, r in runs.items()},
        "cpu_baseline": cpu,
    }
    if args.calibration_copy_gib > 0:
        n = int(args.calibration_copy_gib * (1 << 30)) // 16 * 16
        ms = K.debug_copy(n)  # k_copy16: reads n bytes, writes n bytes
        out["calibration_copy"] = {"bytes": n, "ms": ms, "gbs": 2 * n / (ms * 1e-3) / 1e9}
    if args.ingest:
        out["ingest"] = ingest_report(args.workload, args.scale, args.seed, k, threads)
    if "learned" in runs:
        out["learned"] = runs["learned"]["learned"]
    if len(runs) == 2:
        out["learned_over_prefix"] = runs["learned"]["kmers_per_s"] / runs["prefix"]["kmers_per_s"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
