"""FASTQ parsed on the device (kc_push_fastq) against a restatement of the
reference reader's line state machine (src/FastqReader.h:140-308), against the
oracle's k-mer tables of the reference reader's blocks, and against the
reference runs' fixtures."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C
from tests.test_gpu_parity import check_filtered, check_outputs

pytestmark = pytest.mark.gpu


def sequence_lines(fq: bytes) -> bytes:
    """Header, sequence, plus, quality: the bytes of every line with index 1
    mod 4, each with its '\\n' (src/FastqReader.h:293-308, :367-388)."""
    parts = fq.split(b"\n")
    out = []
    for i, part in enumerate(parts):
        if i % 4 == 1:
            out.append(part if i == len(parts) - 1 else part + b"\n")
    return b"".join(out)


def push_in_chunks(e: K.Engine, fq: bytes, cuts) -> None:
    prev = 0
    for c in list(cuts) + [len(fq)]:
        if c > prev:
            e.push_fastq(fq[prev:c])
            prev = c


def test_sequence_lines_restatement():
    fq = b"@a\nACGT\n+\nIIII\n@b\nGG\n+\nII\n@c\nTT"
    assert sequence_lines(fq) == b"ACGT\nGG\nTT"
    assert sequence_lines(b"") == b""


@pytest.mark.parametrize("chunking", ["whole", "2MiB", "random", "tiny"])
def test_resident_sequence_equals_the_reader_lines(chunking):
    fq = K.generate(2, scale=400, seed=11)
    rng = np.random.default_rng(3)
    if chunking == "whole":
        cuts = []
    elif chunking == "2MiB":
        cuts = range(2 << 20, len(fq), 2 << 20)
    elif chunking == "random":
        cuts = np.sort(rng.integers(0, len(fq), size=200))
    else:
        # cuts right before and after newlines, and 1-byte chunks
        nl = np.flatnonzero(np.frombuffer(fq, np.uint8) == 10)[:400]
        cuts = np.unique(np.concatenate([nl, nl + 1, nl + 2, np.arange(100, 140)]))
    with K.Engine(31) as e:
        push_in_chunks(e, fq, cuts)
        e.finish()
        assert e.resident() == sequence_lines(fq)


def test_odd_records():
    # CRLF line ends, empty sequence lines, a last record without a final
    # newline and one cut inside its sequence line
    recs = [b"@r0\r\nACGTACGTAC\r\n+\r\nIIIIIIIIII\r\n", b"@r1\n\n+\n\n", b"@r2\nNNNNACGTTTGCA\n+\nIIIIIIIIIIIII\n",
            b"@r3 long header with spaces and \t tabs\n" + b"ACGT" * 40 + b"\n+r3\n" + b"I" * 160 + b"\n", b"@r4\nACGTTGCAAC"]
    fq = b"".join(recs)
    for cuts in ([], [len(fq) - 3], list(range(1, len(fq)))):
        with K.Engine(5) as e:
            push_in_chunks(e, fq, cuts)
            e.finish()
            assert e.resident() == sequence_lines(fq)
            keys, counts = C.count_table(C.extract(sequence_lines(fq), 5))
            f, r, s = e.firsts(), e.records(), e.stats()
            check_outputs(f, r, s, 5, keys, counts)


@pytest.mark.parametrize("k", [17, 31, 63, 100])
def test_tables_equal_the_reference_reader_blocks(k):
    # the reference reader's 32 KiB blocks with k-1 carries (oracle restatement)
    # give the same k-mers as whole lines for every k but 96
    fq = K.generate(4, scale=2000, seed=k)
    keys, counts = C.count_table(C.extract_blocks(K.fastq_blocks(fq, k), k))
    with K.Engine(k) as e:
        push_in_chunks(e, fq, range(1 << 20, len(fq), 1 << 20))
        e.finish()
        check_outputs(e.firsts(), e.records(), e.stats(), k, keys, counts)


@pytest.mark.parametrize("name", [n for n, v in sorted(C.reference_runs().items()) if v["K"] <= 13_000_000])
def test_reference_fixture_through_device_parsing(name):
    fx = C.reference_runs()[name]
    k = fx["k"]
    fq = K.generate(fx["workload"], scale=fx["scale"], seed=fx["seed"])
    with K.Engine(k) as e:
        push_in_chunks(e, fq, range(2 << 20, len(fq), 2 << 20))
        e.finish()
        s = e.stats()
        keys, counts = C.merge_outputs(e.firsts(), e.records(), C.nwords(k))
    assert s.total_kmers == fx["K"]
    assert keys.shape[0] == fx["D"]
    assert C.table_sha(keys, counts) == fx["sha256"]


def test_chunks_above_the_staging_buffer():
    # 64 MiB staging: a 155 MB FASTQ is parsed in three device chunks
    fq = K.generate(2, scale=40, seed=4)
    assert len(fq) > (128 << 20)
    seq = K.generate(2, scale=40, seed=4, fmt="seq")
    with K.Engine(31) as e:
        push_in_chunks(e, fq, range(3 << 20, len(fq), 3 << 20))
        e.finish()
        assert e.resident() == seq
        assert e.stats().total_kmers == C.oracle().orc_count_kmers(seq, len(seq), 31)


def test_k96_and_mixed_pushes_are_refused():
    with K.Engine(96) as e:
        with pytest.raises(K.KmcError):
            e.push_fastq(b"@r\n" + b"A" * 100 + b"\n+\n" + b"I" * 100 + b"\n")
    rec = b"@r\nACGTACGTACGTACGTACGTACGTACGTACGTAC\n+\n" + b"I" * 34 + b"\n"
    with K.Engine(31) as e:
        e.push_fastq(rec[:20])  # stops inside the sequence line
        with pytest.raises(K.KmcError):
            e.push_block(b"ACGT\n")
    with K.Engine(31) as e:
        e.push_fastq(rec)  # a whole record: blocks may follow
        e.push_block(b"TTTTGGGGCCCCAAAATTTTGGGGCCCCAAAAT\n")
        e.finish()
        assert e.resident() == sequence_lines(rec) + b"TTTTGGGGCCCCAAAATTTTGGGGCCCCAAAAT\n"


@pytest.mark.parametrize("gz", [False, True], ids=["plain", "gzip"])
def test_tree_run_parses_on_the_device_like_the_host_reader(tmp_path, gz):
    import gzip
    fq = K.generate(4, scale=2000, seed=21)
    path = tmp_path / ("in.fq.gz" if gz else "in.fq")
    path.write_bytes(gzip.compress(fq, compresslevel=1) if gz else fq)
    tables = []
    for host in (True, False):
        out = tmp_path / ("host" if host else "dev")
        out.mkdir()
        s, _ = K.tree_run(str(path), str(out), 31, host_parse=host)
        nw = 1
        recs = np.concatenate([np.fromfile(out / f"root_{i}.bin", dtype=K.record_dtype(nw)) for i in range(256)])
        low = (out / "low.bin").read_bytes()
        firsts = np.zeros(len(low) // 8, dtype=np.uint64)
        C.oracle().orc_unpack_low(low, len(low) // 8, 31, nw, firsts.ctypes.data)
        tables.append(C.merge_outputs(firsts, recs, nw))
        assert s.total_kmers == int(tables[-1][1].sum())
    keys, counts = C.count_table(C.extract_blocks(K.fastq_blocks(fq, 31), 31))
    for got_k, got_c in tables:
        np.testing.assert_array_equal(got_k, keys)
        np.testing.assert_array_equal(got_c, counts)


@pytest.mark.parametrize("counter", [K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT], ids=["exact", "sort"])
def test_tree_run_files_with_the_other_counters(tmp_path, counter):
    # kmc_tree_run with the exact or the sort counter (KMC_COUNTER=exact|sort
    # on the kmc_tree command line): low.bin and the root files merge to the
    # oracle's table, as with the Bloom counter
    fq = K.generate(4, scale=2000, seed=23)
    path = tmp_path / "in.fq"
    path.write_bytes(fq)
    out = tmp_path / "tmp"
    out.mkdir()
    s, _ = K.tree_run(str(path), str(out), 31, counter=counter)
    recs = np.concatenate([np.fromfile(out / f"root_{i}.bin", dtype=K.record_dtype(1)) for i in range(256)])
    low = (out / "low.bin").read_bytes()
    firsts = np.zeros(len(low) // 8, dtype=np.uint64)
    C.oracle().orc_unpack_low(low, len(low) // 8, 31, 1, firsts.ctypes.data)
    got_k, got_c = C.merge_outputs(firsts, recs, 1)
    keys, counts = C.count_table(C.extract_blocks(K.fastq_blocks(fq, 31), 31))
    np.testing.assert_array_equal(got_k, keys)
    np.testing.assert_array_equal(got_c, counts)
    assert s.total_kmers == int(counts.sum())


@pytest.mark.parametrize("counter", [K.KC_COUNTER_BLOOM, K.KC_COUNTER_EXACT, K.KC_COUNTER_SORT],
                         ids=["bloom", "exact", "sort"])
def test_device_parsing_with_rounds_graph_and_filters(counter):
    # a parsed input counted in bucket-range rounds, then replayed as a graph,
    # with the record filter on: the merged table stays the oracle's
    fq = K.generate(2, scale=200, seed=17)
    keys, counts = C.count_table(C.extract(sequence_lines(fq), 31))
    with K.Engine(31, counter=counter, max_round_keys=600_000) as e:
        push_in_chunks(e, fq, range(1 << 20, len(fq), 1 << 20))
        e.finish()
        s = e.stats()
        assert s.rounds >= 2
        check_outputs(e.firsts(), e.records(), s, 31, keys, counts, exact=counter != K.KC_COUNTER_BLOOM)
    with K.Engine(31, counter=counter, graph=True, min_count=2, max_count=50) as e:
        push_in_chunks(e, fq, range(1 << 20, len(fq), 1 << 20))
        e.finish()
        for _ in range(2):
            e.count_resident()
            r, s = e.records(), e.stats()
            assert len(r) > 0
            check_filtered(e.firsts(), r, s, 31, keys, counts, 2, 50)
