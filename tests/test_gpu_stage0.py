"""Stage 0 on the GPU: kc_plan's prefix histogram, LPT loads and Bloom sizing
against the oracle (src/FastqPrefixCounter.h:31-82, src/main.cpp:38-108,
:175-184), and kmc_tree_run's whole Stage 0 against the plan the reference Tree
printed for the same file (tests/golden/stage0_runs.json)."""
from __future__ import annotations

import gzip
import json

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C

pytestmark = pytest.mark.gpu


def plan_of(stats, classifiers: int):
    counts = np.array(stats.prefix_counts, dtype=np.uint64)
    owners = np.array(stats.prefix_owners, dtype=np.int64)
    assert (owners < classifiers).all()
    loads = np.zeros(classifiers, dtype=np.uint64)
    np.add.at(loads, owners, counts)
    return counts, np.sort(loads), np.array(stats.bloom_capacity, dtype=np.uint64)


@pytest.mark.parametrize("partition", [K.KC_PARTITION_PREFIX, K.KC_PARTITION_LEARNED], ids=["prefix", "learned"])
@pytest.mark.parametrize("workload,scale,seed,k", [(1, 10, 11, 31), (3, 200, 3, 31), (1, 10, 15, 63), (2, 400, 8, 100)])
def test_plan_matches_oracle(workload, scale, seed, k, partition):
    fq = K.generate(workload, scale=scale, seed=seed)
    blocks = K.fastq_blocks(fq, k)
    sample = blocks[: max(1, len(blocks) // 2)]
    keys = C.extract_blocks(sample, k)
    classifiers = 3
    est = 3 * 10**9  # large enough that some capacities exceed the 2^15-word floor
    exp_counts, exp_avg, exp_loads, exp_caps = C.stage0_plan(keys, classifiers, est)
    with K.Engine(k, partition=partition, classifier_num=classifiers, est_file_size=est) as e:
        for b in sample:
            e.sample_block(b)
        e.plan()
        s = e.stats()
    assert s.stage0_sample_kmers == keys.shape[0]
    counts, loads, caps = plan_of(s, classifiers)
    np.testing.assert_array_equal(counts, exp_counts)
    assert int(counts.sum()) // 256 == exp_avg
    np.testing.assert_array_equal(loads, exp_loads)
    np.testing.assert_array_equal(caps, exp_caps)
    assert caps.max() > 1 << 15


def stage0_runs() -> dict:
    return json.loads((C.GOLDEN / "stage0_runs.json").read_text())


@pytest.mark.parametrize("name", ["c1_s10_k31_t8", "c1_s40_k31_t8", "c1_s10_k31_t12_gz", "c1_s10_k63_t16",
                                  "c3_s50_k31_t8_gz"])
def test_tree_run_stage0_matches_the_reference_tree(tmp_path, name):
    # kmc_tree_run samples the file as FastqPreReader does and plans from it:
    # the loads, average and largest Bloom capacity the reference Tree printed,
    # and every clamped prefix count of the oracle's Stage-0 chain
    fx = stage0_runs()[name]
    fq = K.generate(fx["workload"], scale=fx["scale"], seed=fx["seed"])
    data = gzip.compress(fq, compresslevel=6, mtime=0) if fx["gzip"] else fq
    assert len(data) == fx["file_size"]
    path = tmp_path / ("in.fq.gz" if fx["gzip"] else "in.fq")
    path.write_bytes(data)
    s, _ = K.tree_run(str(path), str(tmp_path), fx["k"], n_thread=fx["n_thread"], histogram_only=True,
                      histogram_path=str(tmp_path / "hist.tsv"))
    ncls = C.classifiers_for_threads(fx["n_thread"])
    counts, loads, caps = plan_of(s, ncls)
    assert int(counts.sum()) // 256 == fx["average"]
    assert loads.tolist() == fx["loads"]
    assert int(caps.max()) == fx["max_bloom_capacity"]
    keys = C.extract_blocks(C.pre_read_blocks(fq, fx["k"], len(data), fx["gzip"]), fx["k"])
    exp_counts, _, _, exp_caps = C.stage0_plan(keys, ncls, len(data) * (4 if fx["gzip"] else 1))
    np.testing.assert_array_equal(counts, exp_counts)
    np.testing.assert_array_equal(caps, exp_caps)
    # the counting itself still covers the whole file
    seq = K.generate(fx["workload"], scale=fx["scale"], seed=fx["seed"], fmt="seq")
    assert s.total_kmers == int(C.oracle().orc_count_kmers(seq, len(seq), fx["k"]))
