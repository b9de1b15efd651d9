"""CPU tests of the host side of the drop-in (no GPU calls): the library loads
and exports what include/*.h declares, the FASTQ chunker matches the oracle's
restatement of FastqReader, low.bin packing matches the oracle, the generator is
deterministic, and the kmc_tree command line rejects what Tree rejects."""
from __future__ import annotations

import ctypes
import gzip
import hashlib
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"^KC_API\s+[\w\s\*]+?\b(k\w+)\(", h.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = K.lib()
    names = declared_symbols()
    assert {"kc_create", "kc_push_block", "kc_finish", "kc_emit_root", "kmc_tree_run", "kcs_generate"} <= names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(K.EXPORTED) <= names


def test_words_for_k_boundary():
    L = K.lib()
    # src/main.cpp:579-595: N = 1 up to k 32, 2 up to 64, 4 up to 128
    ks = (0, 1, 32, 33, 63, 64, 65, 96, 128, 129)
    assert [L.kc_words_for_k(k) for k in ks] == [0, 1, 1, 2, 2, 2, 4, 4, 4, 0]
    assert [K.words_for_k(k) for k in ks[1:-1]] == [1, 1, 2, 2, 2, 4, 4, 4]
    for bad in (0, 129):
        with pytest.raises(ValueError):
            K.words_for_k(bad)


def oracle_blocks(fq: bytes, k: int, bs: int) -> list[bytes]:
    O = C.oracle()
    cap = len(fq) * 2 + 1024
    out = ctypes.create_string_buffer(cap)
    lens = np.zeros(len(fq) // 16 + 1024, dtype=np.uint64)
    nb, ol = ctypes.c_uint64(), ctypes.c_uint64()
    rc = O.orc_fastq_blocks(fq, len(fq), k, bs, out, cap, lens.ctypes.data, lens.size, ctypes.byref(nb),
                            ctypes.byref(ol))
    assert rc == 0
    res, off = [], 0
    raw = out.raw
    for n in lens[: nb.value]:
        res.append(raw[off: off + int(n)])
        off += int(n)
    return res


@pytest.mark.parametrize("k,bs", [(31, 32 * 1024), (31, 1000), (63, 300), (5, 200), (1, 129), (127, 4096)])
def test_chunker_matches_oracle_reader(k, bs):
    fq = K.generate(4, scale=4000, seed=2)  # short and long reads: lines longer than blocks
    assert K.fastq_blocks(fq, k, bs) == oracle_blocks(fq, k, bs)


def test_chunker_edge_records():
    cases = [
        b"",
        b"@r0\nACGT",                                # no trailing newline
        b"@r0\nACGTNNacgt\n+\nIIIIIIIIII\n@r1\n\n+\n\n",  # empty sequence line
        b"@r0\r\nACGTACGTACGT\r\n+\r\nIIIIIIIIIIII\r\n",  # CRLF: '\r' stays in the block
        b"@r0\n" + b"ACGTTGCA" * 5000 + b"\n+\n" + b"I" * 40000 + b"\n",  # one line over many blocks
    ]
    for fq in cases:
        for k, bs in ((31, 130), (3, 129), (63, 1000)):
            assert K.fastq_blocks(fq, k, bs) == oracle_blocks(fq, k, bs)


def test_block_split_preserves_kmer_multiset():
    fq = K.generate(1, scale=400, seed=5)
    seq = K.generate(1, scale=400, seed=5, fmt="seq")
    whole = C.count_table(C.extract(seq, 31))
    for bs in (129, 777, 32 * 1024):
        blocks = K.fastq_blocks(fq, 31, bs)
        keys = np.concatenate([C.extract(b, 31) for b in blocks])
        got = C.count_table(keys)
        np.testing.assert_array_equal(got[0], whole[0])
        np.testing.assert_array_equal(got[1], whole[1])


def test_file_reader_plain_and_gzip(tmp_path):
    fq = K.generate(2, scale=2000, seed=8)
    (tmp_path / "a.fq").write_bytes(fq)
    (tmp_path / "a.fq.gz").write_bytes(gzip.compress(fq))
    L = K.lib()
    for name in ("a.fq", "a.fq.gz"):
        blocks = []

        @K.BLOCK_CB
        def cb(_u, d, n):
            blocks.append(ctypes.string_at(d, n))
            return 0

        assert L.kmc_fastq_blocks_file(str(tmp_path / name).encode(), 31, 32 * 1024, 0, cb, None) == 0
        assert blocks == K.fastq_blocks(fq, 31)


@pytest.mark.parametrize("case", ["small", "parts", "gzip", "k63"])
def test_stage0_sampler_matches_oracle(tmp_path, case):
    # the host Stage-0 sampler against the oracle's restatement of
    # FastqPreReader::pre_read: overlapping 2 MiB reads below 6 MiB, three
    # record-aligned parts above, a gzip stream read up to its compressed size
    scale, k = {"small": (40.0, 31), "parts": (10.0, 31), "gzip": (10.0, 31), "k63": (7.0, 63)}[case]
    fq = K.generate(1, scale=scale, seed=31)
    path = tmp_path / "in.fq"
    if case == "gzip":
        data = gzip.compress(fq, compresslevel=6, mtime=0)
        path = tmp_path / "in.fq.gz"
        path.write_bytes(data)
    else:
        data = fq
        path.write_bytes(fq)
    blocks, est = K.stage0_sample_file(path, k)
    want = C.pre_read_blocks(fq, k, len(data), case == "gzip")
    assert blocks == want
    assert est == len(data) * (4 if case == "gzip" else 1)


@pytest.mark.parametrize("k", [17, 31, 32, 33, 63, 64, 65, 96, 100, 128])
def test_pack_low_matches_oracle(k):
    rng = np.random.default_rng(k)
    nw = C.nwords(k)
    keys = C.extract(bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=5000)), k)
    got = K.pack_low(keys, k)
    O = C.oracle()
    buf = (ctypes.c_uint8 * (O.orc_packed_bytes(k) * keys.shape[0]))()
    kk = np.ascontiguousarray(keys)
    O.orc_pack_low(kk.ctypes.data, keys.shape[0], k, nw, buf)
    assert got == bytes(buf)


def test_generator_deterministic_across_threads():
    a = K.generate(3, scale=500, seed=4, threads=1)
    b = K.generate(3, scale=500, seed=4, threads=7)
    assert hashlib.sha256(a).digest() == hashlib.sha256(b).digest()
    assert K.generate(3, scale=500, seed=5, threads=2) != a
    # sequence-only form equals the FASTQ's sequence lines
    s = K.generate(3, scale=500, seed=4, fmt="seq")
    assert s == b"".join(line + b"\n" for line in a.split(b"\n")[1::4])


def test_generator_shapes():
    fq = K.generate(1, scale=100, seed=1)
    lines = fq.split(b"\n")
    assert lines[0] == b"@r0" and lines[2] == b"+" and lines[3] == b"I" * 150
    assert len(lines[1]) == 150 and set(lines[1]) <= set(b"ACGT")
    q = K.generate(2, scale=2000, seed=1).split(b"\n")[3::4]
    allq = b"".join(q)
    assert min(allq) >= 33 + 20 and max(allq) <= 33 + 40


def test_tree_cli_rejects_bad_arguments(tmp_path):
    exe = K.TREE_PATH
    if not exe.exists():
        pytest.skip("kmc_tree not built")
    bad = [[], ["x.fq", "31", "4", "8"], ["x.fq", "31", "8", "8", "1"], ["x.fq", "31", "8", "8", "1024", "5", "2"],
           ["x.fq", "0", "8", "8"], ["x.fq", "129", "8", "8"]]
    for args in bad:
        r = subprocess.run([str(exe), *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode != 0, args


def test_bench_closed_form_kmers_matches_oracle():
    # bench.py's full-size check counts K with numpy in chunks; it must give the
    # oracle's K, also when runs cross chunk boundaries
    import bench
    for w, sc in ((2, 400), (3, 800), (4, 2000)):
        seq = K.generate(w, scale=sc, seed=3, fmt="seq")
        for k in (1, 31, 63):
            want = int(C.oracle().orc_count_kmers(seq, len(seq), k))
            assert bench.closed_form_kmers(seq, k, chunk=1 << 15) == want
            assert bench.closed_form_kmers(seq, k) == want
