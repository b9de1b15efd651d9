"""GPU count histogram (kc_histogram) against the oracle's restatement of
histogram_tool (precise and approximate), the reference runs' histograms and
the file format of tool/histogram_tool.cpp:425-442."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C
from tests.test_gpu_parity import microsatellite_reads

pytestmark = pytest.mark.gpu


def oracle_precise(firsts, recs, nw, minf, maxf):
    O = C.oracle()
    low = np.ascontiguousarray(firsts.reshape(-1, nw))
    rb = recs.view(np.uint8) if len(recs) else np.zeros(1, np.uint8)
    h = np.zeros(maxf - minf + 1, dtype=np.int64)
    O.orc_histogram_precise(rb.ctypes.data, len(recs), low.ctypes.data, low.shape[0], nw, minf, maxf, h.ctypes.data, 0)
    return h


def oracle_approximate(recs, nw, minf, maxf):
    O = C.oracle()
    rb = recs.view(np.uint8) if len(recs) else np.zeros(1, np.uint8)
    h = np.zeros(maxf - minf + 1, dtype=np.int64)
    O.orc_histogram_approximate(rb.ctypes.data, len(recs), nw, minf, maxf, h.ctypes.data)
    return h


INPUTS = {
    "uniform_k31": (lambda: K.generate(1, scale=20, seed=5, fmt="seq"), 31),
    "metagenome_k63": (lambda: K.generate(5, scale=4000, seed=6, fmt="seq"), 63),
    "microsatellite_k31": (lambda: microsatellite_reads(100_000, 7), 31),
}


@pytest.mark.parametrize("counter", [K.KC_COUNTER_BLOOM, K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT],
                         ids=["bloom", "exact", "sort"])
@pytest.mark.parametrize("name", sorted(INPUTS))
def test_histogram_matches_oracle(name, counter):
    make, k = INPUTS[name]
    seq = make()
    nw = C.nwords(k)
    keys, counts = C.count_table(C.extract(seq, k))
    with K.Engine(k, counter=counter, heavy_keys=50_000) as e:
        e.push_block(seq)
        e.finish()
        firsts, recs = e.firsts(), e.records()
        for minf, maxf in ((1, 10000), (3, 50), (1, 2), (2, 2)):
            got = e.histogram(minf, maxf)
            np.testing.assert_array_equal(got, oracle_precise(firsts, recs, nw, minf, maxf))
            np.testing.assert_array_equal(e.histogram(minf, maxf, precise=False),
                                          oracle_approximate(recs, nw, minf, maxf))
        # with min_freq 1 the precise histogram is the merged table's
        np.testing.assert_array_equal(e.histogram(1, 10000), C.histogram(counts, 1, 10000))
        assert int((e.histogram(1, int(counts.max())) * np.arange(1, int(counts.max()) + 1)).sum()) == int(counts.sum())


@pytest.mark.parametrize("name", [n for n, v in sorted(C.reference_runs().items()) if v["K"] <= 13_000_000])
def test_histogram_matches_reference_runs(name):
    fx = C.reference_runs()[name]
    k = fx["k"]
    fq = K.generate(fx["workload"], scale=fx["scale"], seed=fx["seed"])
    with K.Engine(k) as e:
        for b in K.fastq_blocks(fq, k):
            e.push_block(b)
        e.finish()
        h = e.histogram(1, fx["max_freq"])
    got = {str(i + 1): int(v) for i, v in enumerate(h) if v}
    assert got == fx["histogram_nonzero"]


def test_tree_run_writes_the_histogram_file(tmp_path):
    fq = K.generate(1, scale=50, seed=3)
    path = tmp_path / "in.fq"
    path.write_bytes(fq)
    seq = K.generate(1, scale=50, seed=3, fmt="seq")
    _, counts = C.count_table(C.extract(seq, 31))
    out = tmp_path / "hist.tsv"
    (tmp_path / "tmp").mkdir()
    K.tree_run(str(path), str(tmp_path / "tmp"), 31, histogram_path=str(out), histogram_only=True, hist_min=1,
               hist_max=300)
    lines = out.read_text().splitlines()
    assert len(lines) == 300
    assert [int(x.split("\t")[0]) for x in lines] == list(range(1, 301))
    got = np.array([int(x.split("\t")[1]) for x in lines])
    np.testing.assert_array_equal(got, C.histogram(counts, 1, 300))
    assert not (tmp_path / "tmp" / "low.bin").exists()  # histogram only: no files


@pytest.mark.parametrize("counter", [K.KC_COUNTER_BLOOM, K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT],
                         ids=["bloom", "exact", "sort"])
@pytest.mark.parametrize("rounds", [False, True], ids=["one_round", "rounds"])
def test_histogram_tail_totals(counter, rounds):
    # keys whose merged count is above max_freq are left out of the histogram
    # (as histogram_tool drops them) and reported as tail totals, so that
    # sum f * hist[f] + tail k-mers == K and sum hist[f] + tail keys == D
    seq = microsatellite_reads(100_000, 9)
    keys, counts = C.count_table(C.extract(seq, 31))
    kw = dict(max_round_keys=5_000_000) if rounds else {}
    with K.Engine(31, counter=counter, **kw) as e:
        e.push_block(seq)
        e.finish()
        if rounds:
            assert e.stats().rounds >= 2
        for maxf in (1, 30, 1000):
            h = e.histogram(1, maxf)
            s = e.stats()
            over = counts > maxf
            assert s.hist_tail_keys == int(over.sum())
            assert s.hist_tail_kmers == int(counts[over].sum(dtype=np.uint64))
            f = np.arange(1, maxf + 1, dtype=np.int64)
            assert int((f * h).sum()) + s.hist_tail_kmers == int(counts.sum(dtype=np.uint64)) == s.total_kmers
            assert int(h.sum()) + s.hist_tail_keys == keys.shape[0]
        # approximate mode: the records only, their counts above max_freq in the tail
        recs = e.records()
        h = e.histogram(1, 30, precise=False)
        s = e.stats()
        over = recs["count"] > 30
        assert s.hist_tail_keys == int(over.sum())
        assert s.hist_tail_kmers == int(recs["count"][over].sum(dtype=np.uint64))
