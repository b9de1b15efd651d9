#!/usr/bin/env python3
"""Benchmark of the k-mer counting path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workloads 3,2] [--partition both|prefix|learned]

A step is one pass of the whole device pipeline (extraction, partition,
Bloom split, repeat counting, record grouping) over the resident synthetic
input of a workload; the sequence is generated on the host and copied to HBM
before timing.  Each workload runs both partition methods on the same input:
the fixed 256-prefix split, then the learned CDF partition (its network
trained on the GPU from the first 1% of the reads).  The first workload
(default 3: the skewed 100 Mbp genome at 65% GC with 5% microsatellite
tracts, 20M reads of 150 bp, k=31, where the learned partition is compared
with the prefix one) gives `value` (the learned path's k-mers/s over the wall
clock of the timed steps), `learned_over_prefix` and the roofline; workload 2
(100 Mbp at 45% GC with 50-copy repeat families, 0.1% N, k=31) runs in its
stated mode, a captured HIP graph after the autotuner.  Every run checks, at
full size, that the device histogram holds the closed-form K:
sum f * n(f) + the k-mers above max_freq == K.

Prints one JSON line (see the keys below).  The device time of each kernel
comes from HIP events the engine records on its own stream around each launch
(pipeline.phase_ms); roofline.achieved is the dominant kernel's algorithmic
bytes over that time.
cpu_baseline times the reference `Tree` binary (oracle/_ref, compiled from the
reference sources) on a bounded sample of the first workload on the host
cores, median of three runs; where that binary is absent it times the oracle's
C restatement instead.
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
LEARNED_MAX_FINE = 0  # --learned-max-fine: fine buckets of one root, at most (0: the engine's default)
LEARNED_STEPS = 0  # --learned-steps: training steps of the network (0: the engine's default)
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
        "heavy": 0,                                     # chunks of heavy buckets: their keys are in "count"
        "group_records": 2 * F * W + 3 * R * RB,        # firsts read + write; records: root read, read + write
    }


def phase_ms(st) -> dict:
    return {
        "extract_count": st.ms_extract_count, "extract_scatter": st.ms_extract_scatter,
        "sub_count": st.ms_sub_count, "sub_scatter": st.ms_sub_scatter,
        "count": st.ms_count, "heavy": st.ms_heavy, "group_records": st.ms_compact,
    }


PHASES = ("extract_count", "extract_scatter", "sub_count", "sub_scatter", "count", "heavy", "group_records")


COUNT_KERNEL = {"bloom": "k_count", "exact": "k_count_exact", "sort": "k_count_sort"}


def kernel_names(nw: int, counter: str = "bloom") -> dict:
    return {
        "extract_count": f"k_extract<{nw}, false>", "extract_scatter": f"k_extract<{nw}, true>",
        "sub_count": f"k_sub_count<{nw}>", "sub_scatter": f"k_sub<{nw}>",
        "count": f"{COUNT_KERNEL[counter]}<{nw}>",
        "heavy": f"k_heavy<{nw}>",
        "group_records": f"k_spill_emit<{nw}> + k_seg_part + k_seg_apply + k_compact_firsts<{nw}> + k_compact_recs<{nw}> + k_root_scan",
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


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_reference(workload: int, k: int, threads: int, runs: int = 3) -> dict | None:
    """The reference Tree + histogram_tool precise on a bounded sample of the
    workload, median of `runs` runs whose k-mer count is the closed-form K."""
    import statistics
    tree = ROOT / "oracle" / "_ref" / "Tree"
    tool = ROOT / "oracle" / "_ref" / "histogram_tool"
    if not tree.exists():
        return None
    scale = {1: 1.0, 2: 5.0, 3: 5.0, 4: 25.0, 5: 50.0}[workload]
    fq = K.generate(workload, scale=scale, seed=1)
    seq = K.generate(workload, scale=scale, seed=1, fmt="seq")
    expect = count_kmers_host(seq, k)
    del seq
    base = "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None
    got_runs = []
    with tempfile.TemporaryDirectory(dir=base) as td:
        td = Path(td)
        (td / "in.fq").write_bytes(fq)
        del fq
        tmp = td / "tmp"
        tmp.mkdir()
        for attempt in range(runs + 3):
            if len(got_runs) == runs:
                break
            for f in tmp.iterdir():
                f.unlink()
            t0 = time.time()
            r = subprocess.run([str(tree), str(td / "in.fq"), str(k), str(threads), "32"], cwd=td,
                               capture_output=True, text=True, timeout=900)
            wall = time.time() - t0
            m = re.search(r"Total read k-mer count: (\d+)", r.stdout)
            got = int(m.group(1)) if m else -1
            if r.returncode != 0 or got != expect:
                log(f"reference run {attempt} read {got} of {expect} k-mers; discarded")
                continue
            us = {key: int(v) for key, v in re.findall(r"(Init|Mid|Final|Total) elapsed us: (\d+)", r.stdout)}
            hist_s = None
            if tool.exists():
                t1 = time.time()
                hr = subprocess.run([str(tool), "precise", str(tmp), str(k), str(threads), "32", str(td / "hist.tsv"),
                                     "1", "10000"], capture_output=True, text=True, timeout=900)
                hist_s = time.time() - t1 if hr.returncode == 0 else None
            got_runs.append({"init_s": us.get("Init", 0) / 1e6, "count_s": (us.get("Mid", 0) + us.get("Final", 0)) / 1e6,
                             "total_s": us.get("Total", 0) / 1e6, "wall_s": wall, "histogram_s": hist_s})
    if not got_runs:
        return None
    med = {key: statistics.median(r[key] for r in got_runs if r[key] is not None)
           for key in ("init_s", "count_s", "total_s", "wall_s")}
    hs = [r["histogram_s"] for r in got_runs if r["histogram_s"] is not None]
    med["histogram_s"] = statistics.median(hs) if hs else None
    return {
        "value": expect / med["count_s"] if med["count_s"] > 0 else None,
        "unit": "k-mers/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"workload {workload} at 1/{scale:g} scale ({expect} k-mers), reference Tree k={k} with "
                   f"{threads} threads (the GPU box's CPU share; nproc {os.cpu_count()}), median of "
                   f"{len(got_runs)} runs: counting phase (Mid+Final) {med['count_s']:.3f} s, Stage 0 (Init) "
                   f"{med['init_s']:.3f} s, Total {med['total_s']:.3f} s, histogram_tool precise "
                   + (f"{med['histogram_s']:.3f} s" if med["histogram_s"] is not None else "n/a")),
        "cpu_model": cpu_model(),
        "runs": len(got_runs),
        "median_s": med,
        "kmers": expect,
        "total_kmers_per_s": expect / med["total_s"] if med["total_s"] > 0 else None,
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


def closed_form_kmers(seq: bytes, k: int, chunk: int = 1 << 28) -> int:
    """K = sum over maximal runs of valid bases (A C G T U, either case) of
    max(0, r - k + 1), computed with numpy in chunks: the count every window
    of k valid bases is counted once gives, independent of the device."""
    import numpy as np
    lut = np.zeros(256, dtype=np.int8)
    lut[np.frombuffer(b"ACGTUacgtu", dtype=np.uint8)] = 1
    arr = np.frombuffer(seq, dtype=np.uint8)
    total, carry = 0, 0  # carry: the valid run that reaches the end of the previous chunk
    for s0 in range(0, arr.size, chunk):
        v = lut[arr[s0: s0 + chunk]]
        d = np.diff(v, prepend=np.int8(0), append=np.int8(0))
        starts = np.flatnonzero(d == 1)
        ends = np.flatnonzero(d == -1)
        lens = (ends - starts).astype(np.int64)
        if lens.size and starts[0] == 0:
            lens[0] += carry
        elif carry:
            total += max(0, carry - k + 1)
        carry = 0
        if lens.size and ends[-1] == v.size:
            carry = int(lens[-1])
            lens = lens[:-1]
        total += int(np.maximum(lens - k + 1, 0).sum())
    return total + max(0, carry - k + 1)


def run_partition(seq: bytes, k: int, partition: str, steps: int, warmup: int, counter: str, graph: bool,
                  autotune: bool = False, closed_k: int | None = None, bounds_io: dict | None = None) -> dict:
    """Resident run of one partition method: plan (learned only), copy, warm up, time `steps` passes.
    bounds_io: {"load": path} routes by saved learned boundaries instead of training
    (for profilers that cannot follow the training), {"save": path} keeps them."""
    import numpy as np
    part = K.KC_PARTITION_LEARNED if partition == "learned" else K.KC_PARTITION_PREFIX
    cnt = {"bloom": K.KC_COUNTER_BLOOM, "exact": K.KC_COUNTER_EXACT, "sort": K.KC_COUNTER_SORT}[counter]
    eng = K.Engine(k, partition=part, expected_bytes=len(seq), counter=cnt, graph=graph,
                   learned_max_fine=LEARNED_MAX_FINE, learned_steps=LEARNED_STEPS)
    plan_ms = 0.0
    bounds_io = bounds_io or {}
    bounds = None
    if part == K.KC_PARTITION_LEARNED and bounds_io.get("load"):
        bounds = np.fromfile(bounds_io["load"], dtype=np.uint64)
        eng.set_partition_bounds(bounds)
        log(f"[{partition}] routing by the boundaries in {bounds_io['load']} (no training)")
    elif part == K.KC_PARTITION_LEARNED:
        # Stage 0 sample: the first 1% of the reads; the network trains on the GPU
        t0 = time.time()
        eng.sample_block(seq[: len(seq) // 100])
        eng.plan()
        plan_ms = (time.time() - t0) * 1e3
        bounds = eng.partition_bounds()
        if bounds is not None and bounds_io.get("save"):
            bounds.tofile(bounds_io["save"])
    t0 = time.time()
    eng.push_block(seq)
    eng.finish()  # H2D copy + first full pass (also the first warmup)
    tuned = None
    if autotune:  # untimed: time the candidates once, keep the fastest
        t_t = time.perf_counter()
        r = eng.autotune()
        tuned = {"sub_bits": r.sub_bits, "counter": {K.KC_COUNTER_EXACT: "exact", K.KC_COUNTER_SORT: "sort"}.get(r.counter, "bloom"),
                 "heavy_keys": r.heavy_keys, "learned_max_fine": r.learned_max_fine,
                 "candidates": r.candidates, "skipped": r.skipped, "best_ms": r.best_ms,
                 "tune_ms": (time.perf_counter() - t_t) * 1e3}
        log(f"[{partition}] autotune: {tuned}")
    st = eng.stats()
    log(f"[{partition}] resident in {time.time() - t0:.1f} s; K={st.total_kmers} firsts={st.firsts} "
        f"records={st.records} spilled={st.spilled_keys} fine_buckets={st.fine_buckets} "
        f"max/mean={st.max_bucket_keys / max(st.mean_bucket_keys, 1):.2f} extra_passes={st.extra_passes} "
        f"rounds={st.rounds} 4096-group max/mean={st.bucket4096_max_over_mean:.2f}")
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
    if graph:
        # a graph replay records no per-kernel times: two eager passes (not timed
        # for the value) give the phase split of the same configuration
        eng2 = None
        per_phase = {name: 0.0 for name in PHASES}
        cfg_eager = dict(partition=part, expected_bytes=len(seq), counter=eng_counter(eng, cnt, tuned), graph=False,
                         learned_max_fine=LEARNED_MAX_FINE)
        if tuned:
            cfg_eager.update(sub_bits=tuned["sub_bits"] or K.KC_SUB_BITS_NONE, heavy_keys=tuned["heavy_keys"],
                             learned_max_fine=tuned["learned_max_fine"])
        eng2 = K.Engine(k, **cfg_eager)
        if bounds is not None:  # the same partition, without training again
            eng2.set_partition_bounds(bounds)
        eng2.push_block(seq)
        eng2.finish()
        for _ in range(2):
            eng2.count_resident()
            for name, v in phase_ms(eng2.stats()).items():
                per_phase[name] += v / 2
        eng2.close()
    # histogram_tool's precise histogram straight from the device tables (outside
    # the timed steps), and the full-size property check sum f * n(f) + tail == K
    t_h = time.perf_counter()
    hist = eng.histogram(1, 10000)
    hist_ms = (time.perf_counter() - t_h) * 1e3
    sh = eng.stats()
    fk = int((np.arange(1, hist.size + 1, dtype=np.int64) * hist).sum())
    check = {"hist_sum_f_n": fk, "tail_kmers": sh.hist_tail_kmers, "tail_keys": sh.hist_tail_keys,
             "device_kmers": st.total_kmers}
    if closed_k is not None:
        check["closed_form_kmers"] = closed_k
        check["ok"] = fk + sh.hist_tail_kmers == closed_k == st.total_kmers
        log(f"[{partition}] check: sum f*n(f) {fk} + tail {sh.hist_tail_kmers} vs closed-form K {closed_k}: "
            f"{'ok' if check['ok'] else 'MISMATCH'}")
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
        "check": check,
    }
    if part == K.KC_PARTITION_LEARNED:
        r["learned"] = {
            "bounds_from": bounds_io.get("load"),  # set: no training in this process
            "fallback": st.learned_fallback,
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


def eng_counter(eng, cnt: int, tuned: dict | None) -> int:
    if tuned:
        return {"exact": K.KC_COUNTER_EXACT, "sort": K.KC_COUNTER_SORT}.get(tuned["counter"], K.KC_COUNTER_BLOOM)
    return cnt


def summary(r: dict) -> dict:
    st = r["stats"]
    return {
        "kmers_per_s": r["kmers_per_s"], "ms_per_step": r["ms_per_step"],
        "device_ms_per_step": r["device_ms_per_step"], "fine_buckets": st.fine_buckets,
        "max_over_mean_bucket": r["max_over_mean_bucket"], "extra_passes": st.extra_passes,
        "spilled_keys": st.spilled_keys, "heavy_buckets": st.heavy_buckets, "rounds": st.rounds,
        "phase_ms": r["per_phase"],
        "histogram_ms": r["histogram_ms"], "autotune": r["autotune"], "plan_ms": r["plan_ms"],
        "bucket4096_max_over_mean": st.bucket4096_max_over_mean, "check": r["check"],
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


# the mode each workload's configuration states: C2 and C4 name the hipGraph
# pipeline with autotuned tiles, C5 the autotuned counter; C1 and C3 none
STATED_MODE = {1: (False, False), 2: (True, True), 3: (False, False), 4: (True, True), 5: (False, True)}


def contract_bytes(st, seq_bytes: int) -> dict:
    """Bytes of the path's algorithmic contract (independent of this design's
    passes): read the sequence, write each key once, one 8-byte Bloom word read
    and written per key, read the repeated keys, write the records and the
    first-occurrence keys."""
    W = 8 * st.key_words
    Kk, F, Dr = st.total_kmers, st.firsts, st.records_unfiltered
    Krep = Kk - F
    base = seq_bytes + Kk * W + Krep * W + Dr * (W + 4) + F * W
    return {"with_bloom_words": base + 16 * Kk, "without_bloom_words": base}


def run_workload(w: int, args, threads: int, primary: bool) -> dict:
    desc, k = WORKLOADS[w]
    graph, autotune = STATED_MODE[w]
    if args.graph:
        graph = True
    if args.autotune:
        autotune = True
    if args.eager:
        graph, autotune = False, False
    t0 = time.time()
    seq = K.generate(w, scale=args.scale, seed=args.seed, fmt="seq", threads=threads)
    seq_bytes = len(seq)
    log(f"[workload {w}] generated {seq_bytes / 1e9:.2f} GB of sequence in {time.time() - t0:.1f} s")
    t0 = time.time()
    closed_k = closed_form_kmers(seq, k) if not args.no_check else None
    if closed_k is not None:
        log(f"[workload {w}] closed-form K = {closed_k} ({time.time() - t0:.1f} s)")
    parts = ["prefix", "learned"] if args.partition == "both" else [args.partition]
    bounds_io = {"save": args.save_bounds, "load": args.load_bounds}
    runs = {p: run_partition(seq, k, p, args.steps, args.warmup, args.counter, graph, autotune, closed_k, bounds_io)
            for p in parts}
    del seq
    main_run = runs["learned"] if "learned" in runs else runs["prefix"]
    st = main_run["stats"]
    per_phase = main_run["per_phase"]
    padded = seq_bytes + (-(seq_bytes) % (512 * 16))
    alg = algorithmic_bytes(st, padded)
    contract = contract_bytes(st, seq_bytes)
    dev_s = main_run["device_ms_per_step"] * 1e-3
    res = {
        "workload": desc + (f" at 1/{args.scale:g} scale" if args.scale != 1 else ""),
        "k": k,
        "graph": graph,
        "autotune": autotune,
        "kmers": st.total_kmers,
        "sequence_bytes": seq_bytes,
        "value_partition": main_run["partition"],
        "kmers_per_s": main_run["kmers_per_s"],
        "partitions": {p: summary(r) for p, r in runs.items()},
        "pipeline": {
            "device_ms_per_step": main_run["device_ms_per_step"],
            "phase_ms": per_phase,
            "kernel_bytes": alg,
            "kernel_bytes_frac": sum(alg.values()) / dev_s / 1e9 / HBM_PEAK_GBS,
            "contract_bytes": contract,
            "contract_frac": {n: v / dev_s / 1e9 / HBM_PEAK_GBS for n, v in contract.items()},
            "phase_gbs": {n: (alg[n] / (per_phase[n] * 1e-3) / 1e9 if per_phase[n] > 0 else None) for n in alg},
        },
        "checks_ok": all(r["check"].get("ok", True) for r in runs.values()),
    }
    if len(runs) == 2:
        res["learned_over_prefix"] = runs["learned"]["kmers_per_s"] / runs["prefix"]["kmers_per_s"]
        res["bucket4096_max_over_mean"] = {p: runs[p]["stats"].bucket4096_max_over_mean for p in runs}
    if "learned" in runs:
        res["learned"] = runs["learned"]["learned"]
    if primary:
        names = kernel_names(st.key_words, (main_run["autotune"] or {}).get("counter", args.counter))
        dominant = max(per_phase, key=per_phase.get)
        timed = per_phase[dominant] > 0
        achieved = alg[dominant] / (per_phase[dominant] * 1e-3) / 1e9 if timed else None
        traffic = traffic_from_profiles(dominant, w, main_run["partition"]) if timed else None
        res["roofline"] = {
            "bound": "hbm",
            "kernel": names[dominant] if timed else None,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved is not None else None,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg[dominant],
            "launch_ms": per_phase[dominant],
        }
        res["kernels"] = names
        res["_main"] = main_run
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workloads", default="3,2",
                    help="comma-separated workloads; the first gives value and roofline (default 3,2)")
    ap.add_argument("--workload", type=int, choices=sorted(WORKLOADS), help="one workload only")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--partition", choices=["prefix", "learned", "both"], default="both")
    ap.add_argument("--counter", choices=["bloom", "exact", "sort"], default="bloom")
    ap.add_argument("--autotune", action="store_true", help="autotune every workload (untimed, before the steps)")
    ap.add_argument("--graph", action="store_true", help="replay every workload as a captured HIP graph")
    ap.add_argument("--eager", action="store_true", help="no graph and no autotuner for any workload")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--learned-steps", type=int, default=0,
                    help="learned partition: training steps of the network (0: the engine's default)")
    ap.add_argument("--learned-max-fine", type=int, default=0,
                    help="learned partition: fine buckets of one root bucket, at most (0: the engine's default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the closed-form K check")
    ap.add_argument("--ingest", action="store_true",
                    help="also time kmc_tree_run end to end from a FASTQ file, host reader vs device parser")
    ap.add_argument("--save-bounds", help="write the learned partition's 4097 boundaries (uint64) to this file")
    ap.add_argument("--load-bounds", help="route the learned partition by boundaries from --save-bounds, no training")
    ap.add_argument("--calibration-copy-gib", type=float, default=0.0,
                    help="after the run, copy this many GiB with k_copy16 (a plain copy whose "
                         "FETCH_SIZE/WRITE_SIZE readings scale those of the pipeline kernels)")
    args = ap.parse_args()
    global LEARNED_MAX_FINE, LEARNED_STEPS
    LEARNED_MAX_FINE = args.learned_max_fine
    LEARNED_STEPS = args.learned_steps

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (the k-mer exchange between "
              "GPUs is a design note only); run with --gpus 1", file=sys.stderr)
        return 2

    wl = [args.workload] if args.workload else [int(x) for x in args.workloads.split(",") if x.strip()]
    # the GPU box's CPU share is 16 cores, whatever nproc shows there
    threads = min(16, os.cpu_count() or 1)
    results = {}
    for i, w in enumerate(wl):
        results[w] = run_workload(w, args, threads, primary=(i == 0))
    prim = results[wl[0]]
    main_run = prim.pop("_main")
    st = main_run["stats"]
    k = prim["k"]

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_reference(wl[0], k, max(6, threads))
        except Exception as e:  # a missing or failing reference binary falls back to the port
            log(f"reference baseline failed: {e}")
        if cpu is None:
            cpu = cpu_baseline_port(wl[0], k)

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
            "workload": prim["workload"],
            "partition": main_run["partition"],
            "counter": args.counter,
            "graph": prim["graph"],
            "autotune": prim["autotune"],
            "k": k,
            "kmers": st.total_kmers,
            "distinct_first": st.firsts,
            "root_records": st.records,
            "sequence_bytes": prim["sequence_bytes"],
            "fine_buckets": st.fine_buckets,
            "max_over_mean_bucket": main_run["max_over_mean_bucket"],
            "spilled_keys": st.spilled_keys,
        },
        "roofline": prim["roofline"],
        "learned_over_prefix": prim.get("learned_over_prefix"),
        "bucket4096_max_over_mean": prim.get("bucket4096_max_over_mean"),
        "pipeline": prim["pipeline"],
        "kernels": prim["kernels"],
        "checks_ok": all(r["checks_ok"] for r in results.values()),
        "workloads": {str(w): {key: v for key, v in r.items() if key not in ("roofline", "kernels")}
                      for w, r in results.items()},
        "cpu_baseline": cpu,
    }
    if args.calibration_copy_gib > 0:
        n = int(args.calibration_copy_gib * (1 << 30)) // 16 * 16
        ms = K.debug_copy(n)  # k_copy16: reads n bytes, writes n bytes
        out["calibration_copy"] = {"bytes": n, "ms": ms, "gbs": 2 * n / (ms * 1e-3) / 1e9}
    if args.ingest:
        out["ingest"] = ingest_report(wl[0], args.scale, args.seed, k, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
