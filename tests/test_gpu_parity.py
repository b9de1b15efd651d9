"""GPU parity: the HIP engine against the oracle and against the reference's
own outputs (tests/golden/reference_runs.json), through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C  # noqa: E402

pytestmark = pytest.mark.gpu


def run_engine(seq_blocks, k, **kw):
    with K.Engine(k, **kw) as e:
        for b in seq_blocks:
            e.push_block(b)
        e.finish()
        return e.firsts(), e.records(), e.stats()


def as_rows(keys: np.ndarray) -> np.ndarray:
    """(n, N) u64 keys -> 1-D structured array that sorts/searches lexicographically."""
    nw = keys.shape[1]
    dt = np.dtype([(f"w{i}", "<u8") for i in range(nw)])
    return np.ascontiguousarray(keys).view(dt).ravel()


COUNTERS = pytest.mark.parametrize("counter", [K.KC_COUNTER_EXACT, K.KC_COUNTER_BLOOM, K.KC_COUNTER_SORT],
                                   ids=["exact", "bloom", "sort"])


def check_outputs(firsts, recs, stats, k, exp_keys, exp_counts, exact=False):
    nw = C.nwords(k)
    assert stats.total_kmers == int(exp_counts.sum(dtype=np.uint64))
    keys, counts = C.merge_outputs(firsts, recs, nw)
    assert keys.shape == exp_keys.shape
    np.testing.assert_array_equal(keys, exp_keys)
    np.testing.assert_array_equal(counts, exp_counts)
    assert stats.firsts == firsts.shape[0]
    assert stats.records == len(recs)
    frows_all = as_rows(firsts.reshape(-1, nw).copy())
    assert np.unique(frows_all).size == frows_all.size  # a key is a first occurrence at most once
    if exact:
        # exact counter: every key is a first occurrence unless it was spilled, and a
        # spilled key's record holds its whole count
        rk = as_rows(recs["key"].reshape(-1, nw)) if len(recs) else np.zeros(0, frows_all.dtype)
        srt = np.sort(frows_all)
        pos = np.searchsorted(srt, rk)
        in_f = (pos < srt.size) & (srt[np.minimum(pos, max(srt.size - 1, 0))] == rk) if srt.size else np.zeros(rk.size, bool)
        assert stats.firsts + int((~in_f).sum()) == exp_keys.shape[0]
    if len(recs):
        assert (recs["pad"] == 0).all()
        assert (recs["count"] >= 1).all()
        # Bloom split (src/FastqClassifier.h:241-349): a key reported first has its
        # remaining c-1 occurrences in its record; a key whose first occurrence was a
        # false positive is absent from the firsts and its record holds all c
        rk = recs["key"].reshape(-1, nw)
        exp_rows = as_rows(exp_keys)
        idx = np.searchsorted(exp_rows, as_rows(rk))
        c = exp_counts[idx]
        frows = np.sort(as_rows(firsts.reshape(-1, nw).copy()))
        pos = np.searchsorted(frows, as_rows(rk))
        in_firsts = (pos < frows.size) & (frows[np.minimum(pos, max(frows.size - 1, 0))] == as_rows(rk)) if frows.size else np.zeros(len(rk), bool)
        np.testing.assert_array_equal(recs["count"], np.where(in_firsts, c - 1, c))


@COUNTERS
@pytest.mark.parametrize("k", [31, 17, 32, 33, 63, 5])
def test_small_random_stream(k, counter):
    rng = np.random.default_rng(k)
    alphabet = np.frombuffer(b"ACGTacgtuUN\n", dtype=np.uint8)
    p = np.array([.22, .22, .22, .22, .02, .02, .02, .02, .01, .01, .01, .01])
    p /= p.sum()
    seq = bytes(rng.choice(alphabet, size=300_000, p=p))
    keys, counts = C.count_table(C.extract(seq, k))
    f, r, s = run_engine([seq], k, counter=counter)
    check_outputs(f, r, s, k, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)


@pytest.mark.parametrize("k", [1, 2, 5, 15, 16, 17, 31, 32, 33, 64, 65, 128])
def test_every_byte_value(k):
    # runs of bases (both cases, U) broken by every byte value 0..255, each
    # one at several offsets inside the 16-byte words the kernels encode:
    # only A C G T U in either case are bases (src/FastqParser.h:194-236),
    # including bytes one bit away from a letter ('@', '`', 'V', 0xC1, 0xE1)
    rng = np.random.default_rng(500 + k)
    n = 240_000
    seq = rng.choice(np.frombuffer(b"ACGTacgtUu", np.uint8), size=n)
    pos = rng.choice(n, size=256 * 24, replace=False)
    seq[pos] = np.tile(np.arange(256, dtype=np.uint8), 24)
    seq = seq.tobytes()
    keys, counts = C.count_table(C.extract(seq, k))
    f, r, s = run_engine([seq], k)
    check_outputs(f, r, s, k, keys, counts)


def test_empty_and_degenerate_inputs():
    for blocks in ([], [b"\n"], [b"NNNNNNNNNN\n"], [b"ACGT\n"], [b"A" * 31 + b"\n"]):
        f, r, s = run_engine(blocks, 31)
        seq = b"".join(blocks)
        exp = C.extract(seq, 31)
        assert s.total_kmers == exp.shape[0]
        if exp.shape[0]:
            keys, counts = C.count_table(exp)
            check_outputs(f, r, s, 31, keys, counts)


@pytest.mark.parametrize("sub_bits", [K.KC_SUB_BITS_NONE, 3, 8, 9, 10])
def test_fanout_does_not_change_table(sub_bits):
    seq = K.generate(1, scale=20, seed=3, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    f, r, s = run_engine([seq], 31, sub_bits=sub_bits)
    assert s.fine_buckets == 256 << (0 if sub_bits == K.KC_SUB_BITS_NONE else sub_bits)
    check_outputs(f, r, s, 31, keys, counts)


@pytest.mark.parametrize("sub_bits", [K.KC_SUB_BITS_NONE, 8, 10])
def test_bucket_statistics(sub_bits):
    # the largest fine bucket and the mean, against the keys' prefix buckets
    # (with the prefix partition a fine bucket is the key's top 8 + sub_bits
    # bits), over up to 2^18 buckets: the output scan reduces them per chunk
    seq = K.generate(1, scale=20, seed=5, fmt="seq")
    words = C.extract(seq, 31)[:, 0]
    sb = 0 if sub_bits == K.KC_SUB_BITS_NONE else sub_bits
    sizes = np.bincount((words >> np.uint64(56 - sb)).astype(np.int64), minlength=256 << sb)
    f, r, s = run_engine([seq], 31, sub_bits=sub_bits)
    assert s.fine_buckets == sizes.size
    assert s.max_bucket_keys == int(sizes.max())
    assert s.mean_bucket_keys == pytest.approx(words.size / sizes.size)


@pytest.mark.parametrize("k", [31, 63, 100])
@pytest.mark.parametrize("sub_bits", [9, 10])
def test_wide_fanout_with_learned_routing(k, sub_bits):
    # 512 or 1024 fine buckets per root (prefix), or learned roots of up to
    # 1024: u16 bucket indexes in the sub-partition kernel, both partitions
    seq = K.generate(2, scale=200, seed=5 + k, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, k))
    for part in (K.KC_PARTITION_PREFIX, K.KC_PARTITION_LEARNED):
        # learned roots of up to 1024 fine buckets (the wide sub-partition
        # variant next to the narrow one in the same pass)
        with K.Engine(k, partition=part, sub_bits=sub_bits, learned_min_skew_pct=1, learned_max_fine=1024) as e:
            if part == K.KC_PARTITION_LEARNED:
                e.sample_block(seq[: len(seq) // 5])
                e.plan()
            e.push_block(seq)
            e.finish()
            f, r, s = e.firsts(), e.records(), e.stats()
            if part == K.KC_PARTITION_PREFIX:
                assert s.fine_buckets == 256 << sub_bits
            else:  # each root's share of 256 << sub_bits, rounded, 1..1024 per root (heavy roots capped)
                assert 256 <= s.fine_buckets <= (256 << sub_bits) + 256
            check_outputs(f, r, s, k, keys, counts)
            for rid in (0, 99, 255):
                rr = e.root(rid)
                if len(rr):
                    assert (rr["key"][:, 0] >> np.uint64(56) == rid).all()


def test_block_carry_path_matches():
    # the host reader's 32 KiB blocks with k-1 carries count exactly like whole lines
    fq = K.generate(4, scale=2000, seed=9)
    blocks = K.fastq_blocks(fq, 31, 4096)
    seq = K.generate(4, scale=2000, seed=9, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    f, r, s = run_engine(blocks, 31)
    check_outputs(f, r, s, 31, keys, counts)


def test_spill_table_path():
    # Bloom counter, no sub-partition and one pass per bucket: ~190k keys (~11k
    # distinct repeats) per bucket, more than the LDS table holds, so the global
    # spill table must keep the counts exact
    seq = K.generate(1, scale=1, seed=1, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    f, r, s = run_engine([seq], 31, sub_bits=K.KC_SUB_BITS_NONE, pass_keys=1 << 30, counter=K.KC_COUNTER_BLOOM)
    assert s.fine_buckets == 256
    assert s.spilled_keys > 0
    check_outputs(f, r, s, 31, keys, counts)


@pytest.mark.parametrize("pass_keys", [2000, 50000])
def test_multi_pass_buckets(pass_keys):
    # Bloom counter: buckets split into hash-range passes give the same table
    seq = K.generate(1, scale=10, seed=21, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    f, r, s = run_engine([seq], 31, sub_bits=2, pass_keys=pass_keys, counter=K.KC_COUNTER_BLOOM)
    check_outputs(f, r, s, 31, keys, counts)


@pytest.mark.parametrize("k", [31, 63])
@pytest.mark.parametrize("table_slots,pass_keys", [(512, 64), (64, 0), (1024, 0), (0, 100)])
def test_exact_counter_narrowed_passes(k, table_slots, pass_keys):
    # exact counter with small tables: passes narrow their hash range as the
    # table fills (and defer a key whose probe finds no slot); keys left out
    # of a pass's final range are counted again by the next pass
    seq = K.generate(1, scale=40, seed=22, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, k))
    # 256 buckets of ~4.7k keys (~1.4k distinct)
    f, r, s = run_engine([seq], k, sub_bits=K.KC_SUB_BITS_NONE, table_slots=table_slots, pass_keys=pass_keys,
                         counter=K.KC_COUNTER_EXACT)
    assert s.extra_passes > 0
    check_outputs(f, r, s, k, keys, counts, exact=True)


@COUNTERS
@pytest.mark.parametrize("heavy_keys", [1000, 9500])
def test_heavy_buckets_counted_in_chunks(counter, heavy_keys):
    # buckets above heavy_keys keys are listed in chunks and counted by the whole
    # GPU into the global table; 256 buckets of ~9.4k keys: 1000 hands over every
    # bucket (~9 chunks of 1024 keys merged across workgroups), 9500 the larger
    # half (both paths in one run)
    seq = K.generate(1, scale=20, seed=8, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    f, r, s = run_engine([seq], 31, sub_bits=K.KC_SUB_BITS_NONE, heavy_keys=heavy_keys, counter=counter)
    assert 0 < s.heavy_buckets <= 256
    if heavy_keys == 9500:
        assert s.heavy_buckets < 256
    check_outputs(f, r, s, 31, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)


@pytest.mark.parametrize("k", [31, 63, 100])
@pytest.mark.parametrize("stage", [64, 1000])
def test_sort_counter_small_stages(k, stage):
    # a stage of 64 or 1000 keys: passes narrow their hash range when they
    # overflow, and a hash value with more occurrences than the stage (a key
    # repeated that often) is counted in the global spill table
    seq = microsatellite_reads(20_000, 30 + k, tract_frac=0.3)
    keys, counts = C.count_table(C.extract(seq, k))
    assert counts.max() > 1000
    f, r, s = run_engine([seq], k, counter=K.KC_COUNTER_SORT, table_slots=stage, sub_bits=2)
    assert s.extra_passes > 0
    assert s.spilled_keys > 0
    check_outputs(f, r, s, k, keys, counts, exact=True)
    # the split without spills: every distinct key is a first occurrence
    seq = K.generate(1, scale=40, seed=3 + k, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, k))
    f, r, s = run_engine([seq], k, counter=K.KC_COUNTER_SORT, table_slots=stage, sub_bits=2)
    assert s.spilled_keys == 0 and s.firsts == keys.shape[0]
    check_outputs(f, r, s, k, keys, counts, exact=True)


def microsatellite_reads(n_reads, seed, tract_frac=0.3):
    """150 bp reads of a random genome whose tracts repeat units of period 1-6."""
    rng = np.random.default_rng(seed)
    g = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=2_000_000)
    for j in range(0, g.size, 10_000):
        if rng.random() < tract_frac:
            p = int(rng.integers(1, 7))
            unit = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=p)
            g[j:j + 5000] = np.resize(unit, 5000)
    starts = rng.integers(0, g.size - 150, size=n_reads)
    out = np.empty((n_reads, 151), dtype=np.uint8)
    out[:, :150] = g[starts[:, None] + np.arange(150)[None, :]]
    out[:, 150] = ord("\n")
    return out.tobytes()


@COUNTERS
@pytest.mark.parametrize("k", [31, 63])
def test_microsatellite_heavy_hitters(counter, k):
    # a few keys repeated hundreds of thousands of times, in runs: the wave-level
    # run counting and the chunked heavy-bucket path must keep every count exact
    seq = microsatellite_reads(200_000, 11 + k)
    keys, counts = C.count_table(C.extract(seq, k))
    assert counts.max() > 100_000
    f, r, s = run_engine([seq], k, heavy_keys=50_000, counter=counter)
    assert s.heavy_buckets > 0
    check_outputs(f, r, s, k, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)
    f, r, s = run_engine([seq], k, counter=counter)   # default threshold: one workgroup per bucket
    check_outputs(f, r, s, k, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)


@COUNTERS
@pytest.mark.parametrize("name", sorted(C.reference_runs().keys()))
def test_reference_fixture(name, counter):
    fx = C.reference_runs()[name]
    k = fx["k"]
    fq = K.generate(fx["workload"], scale=fx["scale"], seed=fx["seed"])
    blocks = K.fastq_blocks(fq, k)
    f, r, s = run_engine(blocks, k, counter=counter)
    assert s.total_kmers == fx["K"]
    keys, counts = C.merge_outputs(f, r, C.nwords(k))
    assert keys.shape[0] == fx["D"]
    assert C.table_sha(keys, counts) == fx["sha256"]
    h = C.histogram(counts, 1, fx["max_freq"])
    got = {str(i + 1): int(v) for i, v in enumerate(h) if v}
    assert got == fx["histogram_nonzero"]


def check_filtered(firsts, recs, stats, k, exp_keys, exp_counts, minc, maxc):
    """The exact record set of append_export_record (src/NewKmerTree.h:1159-1168):
    a key of count c whose first occurrence went to low.bin has the repeat count
    r = c - 1, any other key r = c; it has a root record, of count r, exactly
    when max(1, min) <= r <= max.  The firsts are input keys, each at most once,
    and the unfiltered record count is the number of keys with r >= 1."""
    nw = C.nwords(k)
    maxc = maxc or 0xFFFFFFFF
    assert stats.total_kmers == int(exp_counts.sum(dtype=np.uint64))
    erows = as_rows(exp_keys)
    frows = as_rows(firsts.reshape(-1, nw).copy())
    assert np.unique(frows).size == frows.size
    pos = np.searchsorted(erows, frows)
    assert (pos < erows.size).all() and (erows[np.minimum(pos, erows.size - 1)] == frows).all()
    in_f = np.zeros(erows.size, dtype=bool)
    in_f[pos] = True
    r = exp_counts.astype(np.int64) - in_f
    assert stats.firsts == frows.size
    assert stats.records_unfiltered == int((r >= 1).sum())
    keep = (r >= max(1, minc)) & (r <= maxc)
    assert stats.records == len(recs) == int(keep.sum())
    rk = as_rows(recs["key"].reshape(-1, nw)) if len(recs) else erows[:0]
    order = np.argsort(rk, kind="stable")
    np.testing.assert_array_equal(rk[order], erows[keep])
    np.testing.assert_array_equal(recs["count"][order].astype(np.int64), r[keep])
    if len(recs):
        assert (recs["pad"] == 0).all()


# paths of the counting pipeline the record filter must hold on
FILTER_PATHS = {
    "buckets": dict(),
    "passes": dict(sub_bits=2, pass_keys=2000),
    "heavy": dict(sub_bits=K.KC_SUB_BITS_NONE, heavy_keys=1000),
    "rounds": dict(max_round_keys=1_500_000),
}


@COUNTERS
@pytest.mark.parametrize("path", sorted(FILTER_PATHS))
@pytest.mark.parametrize("minc,maxc", [(3, 9), (1, 1), (2, 0)], ids=["3-9", "1-1", "2-max"])
def test_min_max_filter_exact_set(counter, path, minc, maxc):
    seq = K.generate(1, scale=20, seed=4, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    with K.Engine(31, min_count=minc, max_count=maxc, counter=counter, **FILTER_PATHS[path]) as e:
        e.push_block(seq)
        e.finish()
        s = e.stats()
        if path == "heavy":
            assert s.heavy_buckets > 0
        if path == "rounds":
            assert s.rounds >= 2
        check_filtered(e.firsts(), e.records(), s, 31, keys, counts, minc, maxc)


def test_min_max_filter_spill_path():
    # Bloom counter, one pass per 190k-key bucket: keys that find no LDS slot
    # live in the global spill table, whose records go through the same filter
    seq = K.generate(1, scale=1, seed=1, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    with K.Engine(31, min_count=2, max_count=40, sub_bits=K.KC_SUB_BITS_NONE, pass_keys=1 << 30,
                  counter=K.KC_COUNTER_BLOOM) as e:
        e.push_block(seq)
        e.finish()
        s = e.stats()
        assert s.spilled_keys > 0
        check_filtered(e.firsts(), e.records(), s, 31, keys, counts, 2, 40)


@COUNTERS
def test_min_max_filter_long_keys(counter):
    seq = K.generate(1, scale=40, seed=9, fmt="seq")
    for k in (63, 100):
        keys, counts = C.count_table(C.extract(seq, k))
        with K.Engine(k, min_count=2, max_count=20, counter=counter) as e:
            e.push_block(seq)
            e.finish()
            check_filtered(e.firsts(), e.records(), e.stats(), k, keys, counts, 2, 20)


def test_root_files_hold_their_prefix():
    seq = K.generate(2, scale=400, seed=6, fmt="seq")
    with K.Engine(31) as e:
        e.push_block(seq)
        e.finish()
        for rid in (0, 7, 128, 255):
            recs = e.root(rid)
            if len(recs):
                assert (recs["key"][:, 0] >> np.uint64(56) == rid).all()


@pytest.mark.parametrize("outputs", ["device", "host", "split"])
@pytest.mark.parametrize("k", [31, 63])
@pytest.mark.parametrize("partition", [K.KC_PARTITION_PREFIX, K.KC_PARTITION_LEARNED], ids=["prefix", "learned"])
def test_bucket_range_rounds(monkeypatch, k, partition, outputs):
    # a key capacity below the input's positions: the engine counts contiguous
    # root-bucket ranges in separate rounds and keeps each round's outputs in
    # device memory (device), host memory (host), or device memory until a
    # small arena is full (split); tables, root files and histograms must not change
    seq = K.generate(1, scale=10, seed=14, fmt="seq")
    nw = C.nwords(k)
    keys, counts = C.count_table(C.extract(seq, k))
    if outputs == "split":  # room for about half of the outputs
        out_bytes = len(keys) * 8 * nw + int((counts > 1).sum()) * (8 * nw + 8)
        monkeypatch.setenv("KMC_MAX_ARENA_BYTES", str(out_bytes // 2))
    with K.Engine(k, partition=partition, max_round_keys=1_500_000, learned_min_skew_pct=1,
                  round_outputs_host=outputs == "host") as e:
        if partition == K.KC_PARTITION_LEARNED:
            e.sample_block(seq[: len(seq) // 20])
            e.plan()
        e.push_block(seq)
        e.finish()
        f, r, s = e.firsts(), e.records(), e.stats()
        assert s.rounds >= max(2, -(-s.total_kmers // 1_500_000))
        # the fewest rounds, evened out: the key buffers hold the largest range only
        assert s.round_key_cap <= 1_500_000
        assert s.rounds * s.round_key_cap >= s.total_kmers
        if outputs == "device":
            assert s.rounds_in_device == s.rounds
        elif outputs == "host":
            assert s.rounds_in_device == 0
        else:
            assert 0 < s.rounds_in_device < s.rounds
        check_outputs(f, r, s, k, keys, counts)
        for rid in (0, 17, 128, 255):
            rr = e.root(rid)
            if len(rr):
                assert (rr["key"][:, 0] >> np.uint64(56) == rid).all()
        np.testing.assert_array_equal(e.histogram(1, 1000), C.histogram(counts, 1, 1000))
        # a second pass over the resident input gives the same tables
        e.count_resident()
        f2, r2 = e.firsts(), e.records()
        k2, c2 = C.merge_outputs(f2, r2, nw)
        np.testing.assert_array_equal(k2, keys)
        np.testing.assert_array_equal(c2, counts)


def test_rounds_reject_a_root_bucket_larger_than_a_round():
    seq = K.generate(1, scale=10, seed=14, fmt="seq")
    # ~19k keys per root bucket on average, far above 4096
    with K.Engine(31, max_round_keys=4096) as e:
        e.push_block(seq)
        with pytest.raises(K.KmcError):
            e.finish()


@pytest.mark.parametrize("partition", [K.KC_PARTITION_PREFIX, K.KC_PARTITION_LEARNED], ids=["prefix", "learned"])
@pytest.mark.parametrize("sub_bits", [8, 9])
def test_kernels_launched_in_pieces(monkeypatch, partition, sub_bits):
    # inputs of 64 GiB and more are launched in pieces of at most 2^31 threads;
    # KMC_MAX_GRID_TILES makes the pieces 13 tiles here (not a multiple of the
    # 8 XCDs), so the count and scatter passes of extraction and
    # sub-partition each run in hundreds of launches
    seq = K.generate(2, scale=400, seed=23, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    monkeypatch.setenv("KMC_MAX_GRID_TILES", "13")
    # learned: roots of both sub-partition variants (up to 1024 fine buckets)
    with K.Engine(31, partition=partition, sub_bits=sub_bits, learned_min_skew_pct=1, learned_max_fine=1024) as e:
        if partition == K.KC_PARTITION_LEARNED:
            e.sample_block(seq[: len(seq) // 10])
            e.plan()
        e.push_block(seq)
        e.finish()
        check_outputs(e.firsts(), e.records(), e.stats(), 31, keys, counts)


def test_autotune_keeps_the_tables_exact():
    seq = K.generate(2, scale=400, seed=15, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    with K.Engine(31) as e:
        e.push_block(seq)
        e.finish()
        r = e.autotune()
        assert r.candidates >= 4
        assert r.counter in (K.KC_COUNTER_BLOOM, K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT)
        assert r.candidates >= 6  # three counters, two heavy thresholds (and a second fan-out on larger inputs)
        assert r.best_ms > 0
        f, rr, s = e.firsts(), e.records(), e.stats()
        check_outputs(f, rr, s, 31, keys, counts)
        e.count_resident()  # the tuned configuration stays in the handle
        k2, c2 = C.merge_outputs(e.firsts(), e.records(), 1)
        np.testing.assert_array_equal(k2, keys)
        np.testing.assert_array_equal(c2, counts)
