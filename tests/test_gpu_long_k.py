"""GPU parity for k = 64 and k = 65..128 (N = 2 full words, N = 4 words;
src/main.cpp:579-595) against the oracle, through the C ABI.

Keys of these lengths can hold all-ones words, which the engine's
open-addressing tables cannot store as such (all ones marks a free word); the
inputs here are built so that such keys occur and are counted."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C
from tests.test_gpu_histogram import oracle_approximate, oracle_precise
from tests.test_gpu_parity import COUNTERS, check_outputs, microsatellite_reads, run_engine

pytestmark = pytest.mark.gpu

UMAX = np.uint64(2**64 - 1)
LONG_K = [64, 65, 95, 96, 97, 100, 127, 128]


def noisy_stream(n: int, seed: int) -> bytes:
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGTacgtuUN\n", dtype=np.uint8)
    p = np.array([.24, .24, .24, .24, .005, .005, .005, .005, .002, .002, .002, .002])
    p /= p.sum()
    return bytes(rng.choice(alphabet, size=n, p=p))


def all_ones_reads(k: int, copies: int) -> bytes:
    """Reads whose windows pack to keys with all-ones words: runs of T (code 3)
    of 32 bases or more next to runs of A, in both orders and at every phase."""
    out = []
    for t in (32, 40, 64, 96, 128):
        for a in (1, 16, 33, 64):
            out.append(b"A" * a + b"T" * t + b"A" * a)
            out.append(b"T" * t + b"A" * max(a, k))
            out.append(b"A" * max(a, k) + b"T" * t + b"C" + b"T" * t)
    return b"\n".join(out * copies) + b"\n"


@COUNTERS
@pytest.mark.parametrize("k", LONG_K)
def test_long_k_random_stream(k, counter):
    seq = noisy_stream(400_000, k) + all_ones_reads(k, 3)
    keys, counts = C.count_table(C.extract(seq, k))
    assert (keys == UMAX).any(), "the input must reach all-ones key words"
    f, r, s = run_engine([seq], k, counter=counter)
    assert s.key_words == C.nwords(k)
    check_outputs(f, r, s, k, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)


def test_k64_key_with_an_all_ones_first_word():
    # 32 T then 32 A packs to (~0, 0): the one k <= 64 key the table stores
    # under a substitute
    seq = (b"T" * 32 + b"A" * 32 + b"\n") * 50 + (b"A" * 32 + b"T" * 32 + b"\n") * 7 + noisy_stream(50_000, 3)
    keys, counts = C.count_table(C.extract(seq, 64))
    assert ((keys[:, 0] == UMAX) & (keys[:, 1] == 0)).any()
    for counter in (K.KC_COUNTER_BLOOM, K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT):
        f, r, s = run_engine([seq], 64, counter=counter)
        check_outputs(f, r, s, 64, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)


@pytest.mark.parametrize("k", [96, 127])
def test_long_k_block_carry_path(k):
    # each reader block is parsed from a cleared state (src/FastqParser.h:193);
    # at k = 96 the key's last word holds the 32 bases before the window, so a
    # window at a block start keys differently from the same window mid-line
    fq = K.generate(4, scale=4000, seed=k)
    blocks = K.fastq_blocks(fq, k, 4096)
    keys, counts = C.count_table(C.extract_blocks(blocks, k))
    f, r, s = run_engine(blocks, k)
    check_outputs(f, r, s, k, keys, counts)
    if k == 127:  # away from k = 96 the blocks count like whole lines
        seq = K.generate(4, scale=4000, seed=k, fmt="seq")
        k2, c2 = C.count_table(C.extract(seq, k))
        np.testing.assert_array_equal(k2, keys)
        np.testing.assert_array_equal(c2, counts)


@COUNTERS
@pytest.mark.parametrize("k", [65, 100])
def test_long_k_spill_passes_and_heavy_chunks(k, counter):
    # one bucket per root prefix, small passes, and heavy buckets handed to the
    # whole GPU: every table path at four words per key
    seq = microsatellite_reads(60_000, k)
    keys, counts = C.count_table(C.extract(seq, k))
    f, r, s = run_engine([seq], k, sub_bits=K.KC_SUB_BITS_NONE, pass_keys=3000, heavy_keys=20_000, counter=counter)
    assert s.heavy_buckets > 0
    # the heavy chunks are timed apart from the counting kernel
    assert s.ms_heavy > 0 and s.ms_count > 0
    assert s.ms_count + s.ms_heavy + s.ms_compact <= s.ms_total
    check_outputs(f, r, s, k, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)
    f, r, s = run_engine([seq], k, sub_bits=K.KC_SUB_BITS_NONE, pass_keys=1 << 30, counter=K.KC_COUNTER_BLOOM)
    assert s.spilled_keys > 0
    check_outputs(f, r, s, k, keys, counts)


@pytest.mark.parametrize("k", [64, 96, 128])
def test_long_k_histogram(k):
    seq = microsatellite_reads(30_000, 5 + k) + all_ones_reads(k, 2)
    nw = C.nwords(k)
    keys, counts = C.count_table(C.extract(seq, k))
    with K.Engine(k, heavy_keys=20_000) as e:
        e.push_block(seq)
        e.finish()
        firsts, recs = e.firsts(), e.records()
        for minf, maxf in ((1, 5000), (3, 40)):
            np.testing.assert_array_equal(e.histogram(minf, maxf), oracle_precise(firsts, recs, nw, minf, maxf))
            np.testing.assert_array_equal(e.histogram(minf, maxf, precise=False),
                                          oracle_approximate(recs, nw, minf, maxf))
        np.testing.assert_array_equal(e.histogram(1, 5000), C.histogram(counts, 1, 5000))


@pytest.mark.parametrize("partition", [K.KC_PARTITION_PREFIX, K.KC_PARTITION_LEARNED], ids=["prefix", "learned"])
def test_long_k_partitions_and_rounds(partition):
    k = 100
    seq = K.generate(1, scale=20, seed=31, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, k))
    with K.Engine(k, partition=partition, max_round_keys=400_000, learned_min_skew_pct=1) as e:
        if partition == K.KC_PARTITION_LEARNED:
            e.sample_block(seq[: len(seq) // 10])
            e.plan()
        e.push_block(seq)
        e.finish()
        f, r, s = e.firsts(), e.records(), e.stats()
        assert s.rounds >= 2
        check_outputs(f, r, s, k, keys, counts)
        for rid in (0, 100, 255):
            rr = e.root(rid)
            if len(rr):
                assert (rr["key"][:, 0] >> np.uint64(56) == rid).all()


def test_long_k_tree_run_files(tmp_path):
    k = 100
    fq = K.generate(1, scale=200, seed=12)
    path = tmp_path / "in.fq"
    path.write_bytes(fq)
    seq = K.generate(1, scale=200, seed=12, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, k))
    tmp = tmp_path / "tmp"
    tmp.mkdir()
    K.tree_run(str(path), str(tmp), k)
    nw = C.nwords(k)
    recs = np.concatenate([np.fromfile(tmp / f"root_{i}.bin", dtype=K.record_dtype(nw)) for i in range(256)])
    low = (tmp / "low.bin").read_bytes()
    pb = C.oracle().orc_packed_bytes(k)
    assert pb == 25 and len(low) % pb == 0
    firsts = np.zeros(len(low) // pb * nw, dtype=np.uint64)
    C.oracle().orc_unpack_low(low, len(low) // pb, k, nw, firsts.ctypes.data)
    got_k, got_c = C.merge_outputs(firsts, recs, nw)
    np.testing.assert_array_equal(got_k, keys)
    np.testing.assert_array_equal(got_c, counts)
