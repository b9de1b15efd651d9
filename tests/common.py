"""Shared test helpers: the oracle (test infrastructure, oracle/liboracle.so),
fixture access and merged-table utilities."""
from __future__ import annotations

import ctypes
import hashlib
import json
import subprocess
from functools import lru_cache
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_HARNESS = ROOT / "oracle" / "_ref" / "libref_harness.so"

u64, u32, vp, cp = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_char_p
UMAX = 2**64 - 1


@lru_cache(maxsize=1)
def oracle() -> ctypes.CDLL:
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True, capture_output=True)
    L = ctypes.CDLL(str(ORACLE_SO))
    L.orc_words_for_k.restype = u32
    L.orc_words_for_k.argtypes = [u32]
    L.orc_fastq_blocks.argtypes = [cp, u64, u32, u64, vp, u64, vp, u64, vp, vp]
    L.orc_extract.restype = u64
    L.orc_extract.argtypes = [cp, u64, u32, vp, u64]
    L.orc_count_kmers.restype = u64
    L.orc_count_kmers.argtypes = [cp, u64, u32]
    L.orc_extract_blocks.restype = u64
    L.orc_extract_blocks.argtypes = [cp, vp, u64, u32, vp, u64]
    L.orc_sort_keys.argtypes = [vp, u64, u32]
    L.orc_count_sorted.restype = u64
    L.orc_count_sorted.argtypes = [vp, u64, u32, vp]
    L.orc_xxh3_128.argtypes = [vp, u32, vp, vp]
    L.orc_bloom_probe.argtypes = [vp, u32, u64, vp, vp]
    L.orc_bloom_classify.restype = u64
    L.orc_bloom_classify.argtypes = [vp, u64, u32, u64, vp, vp]
    L.orc_packed_bytes.restype = u64
    L.orc_packed_bytes.argtypes = [u32]
    L.orc_pack_low.argtypes = [vp, u64, u32, u32, vp]
    L.orc_unpack_low.argtypes = [vp, u64, u32, u32, vp]
    L.orc_record_bytes.restype = u64
    L.orc_record_bytes.argtypes = [u32]
    L.orc_merge_outputs.restype = u64
    L.orc_merge_outputs.argtypes = [vp, u64, vp, u64, u32, vp, vp]
    L.orc_histogram_precise.argtypes = [vp, u64, vp, u64, u32, u32, u32, vp, ctypes.c_int]
    L.orc_histogram_from_counts.argtypes = [vp, u64, u32, u32, vp]
    L.orc_histogram_approximate.argtypes = [vp, u64, u32, u32, u32, vp]
    L.orc_prefix_counts.argtypes = [vp, u64, u32, vp]
    L.orc_clamp_prefix_counts.restype = u64
    L.orc_clamp_prefix_counts.argtypes = [vp]
    L.orc_lpt.argtypes = [vp, u32, vp, vp]
    L.orc_bloom_capacity.argtypes = [vp, u64, vp]
    L.orc_pre_read.argtypes = [cp, u64, u64, ctypes.c_int, u32, u64, vp, u64, vp, u64, vp, vp]
    return L


@lru_cache(maxsize=1)
def ref_harness():
    """Reference implementation shim (oracle/_ref), or None where it was not built."""
    if not REF_HARNESS.exists():
        return None
    L = ctypes.CDLL(str(REF_HARNESS))
    L.ref_extract.restype = u64
    L.ref_extract.argtypes = [cp, u64, u32, vp, u64]
    L.ref_xxh3_128.argtypes = [vp, u64, vp, vp]
    L.ref_bloom_classify.restype = u64
    L.ref_bloom_classify.argtypes = [vp, u64, u32, u64, vp, vp]
    return L


def nwords(k: int) -> int:
    return 1 if k <= 32 else (2 if k <= 64 else 4)


def extract(seq: bytes, k: int) -> np.ndarray:
    """Oracle canonical keys of a byte stream, shape (n, N)."""
    O = oracle()
    nw = nwords(k)
    cap = max(1, len(seq))
    out = np.zeros(cap * nw, dtype=np.uint64)
    n = O.orc_extract(seq, len(seq), k, out.ctypes.data, cap)
    assert n != UMAX
    return out[: n * nw].reshape(n, nw)


def extract_blocks(blocks: list[bytes], k: int) -> np.ndarray:
    """Oracle keys of reader blocks, each parsed from a cleared state, shape (n, N)."""
    O = oracle()
    nw = nwords(k)
    data = b"".join(blocks)
    lens = np.array([len(b) for b in blocks] or [0], dtype=np.uint64)
    cap = max(1, len(data))
    out = np.zeros(cap * nw, dtype=np.uint64)
    n = O.orc_extract_blocks(data, lens.ctypes.data, len(blocks), k, out.ctypes.data, cap)
    assert n != UMAX
    return out[: n * nw].reshape(n, nw)


def pre_read_blocks(fq: bytes, k: int, file_size: int | None = None, gz: bool = False,
                    block_size: int = 32 * 1024) -> list[bytes]:
    """Oracle Stage-0 sample blocks of a FASTQ file (FastqPreReader::pre_read);
    fq is the decompressed content, file_size the size on disk."""
    O = oracle()
    fs = len(fq) if file_size is None else file_size
    cap = 2 * len(fq) + (1 << 20)
    out = ctypes.create_string_buffer(cap)
    lens = np.zeros(cap // 64 + 16, dtype=np.uint64)
    nb, ol = u64(), u64()
    rc = O.orc_pre_read(fq, len(fq), fs, 1 if gz else 0, k, block_size, out, cap, lens.ctypes.data, lens.size,
                        ctypes.byref(nb), ctypes.byref(ol))
    assert rc == 0
    raw = out.raw[: ol.value]
    blocks, at = [], 0
    for n in lens[: nb.value]:
        blocks.append(raw[at: at + int(n)])
        at += int(n)
    return blocks


def stage0_plan(keys: np.ndarray, classifiers: int, est_file_size: int):
    """Oracle Stage-0 plan of sampled keys: clamped prefix counts, their
    average, sorted LPT loads and Bloom capacities (src/main.cpp:38-108,175-184)."""
    O = oracle()
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.zeros(256, dtype=np.uint64)
    O.orc_prefix_counts(keys.ctypes.data, keys.shape[0], keys.shape[1] if keys.ndim == 2 else 1, counts.ctypes.data)
    avg = int(O.orc_clamp_prefix_counts(counts.ctypes.data))
    owners = np.zeros(256, dtype=np.uint8)
    loads = np.zeros(classifiers, dtype=np.uint64)
    O.orc_lpt(counts.ctypes.data, classifiers, owners.ctypes.data, loads.ctypes.data)
    caps = np.zeros(256, dtype=np.uint64)
    O.orc_bloom_capacity(counts.ctypes.data, est_file_size, caps.ctypes.data)
    return counts, avg, np.sort(loads), caps


def classifiers_for_threads(n_thread: int) -> int:
    """classifier_num of the reference's thread split (src/main.cpp:114-117)."""
    parsers = max(1, n_thread // 8)
    budget = n_thread - 2 - parsers
    return max(1, int(budget / 2.0))


def count_table(keys: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Oracle exact table: sorted distinct keys and their counts."""
    O = oracle()
    keys = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    n, nw = keys.shape
    counts = np.zeros(max(n, 1), dtype=np.uint32)
    d = O.orc_count_sorted(keys.ctypes.data, n, nw, counts.ctypes.data)
    assert d != UMAX
    return keys[:d].copy(), counts[:d].copy()


def merge_outputs(firsts: np.ndarray, recs: np.ndarray, nw: int) -> tuple[np.ndarray, np.ndarray]:
    """Merged (key, count) table of engine outputs; raises if a key repeats
    within the firsts or within the records (the file contract forbids it)."""
    O = oracle()
    firsts = np.ascontiguousarray(firsts, dtype=np.uint64)
    nf = firsts.size // nw
    rb = np.ascontiguousarray(recs).view(np.uint8) if len(recs) else np.zeros(0, dtype=np.uint8)
    nr = len(recs)
    keys = np.zeros(max(nf + nr, 1) * nw, dtype=np.uint64)
    counts = np.zeros(max(nf + nr, 1), dtype=np.uint32)
    d = O.orc_merge_outputs(firsts.ctypes.data, nf, rb.ctypes.data if nr else None, nr, nw,
                            keys.ctypes.data, counts.ctypes.data)
    assert d != UMAX, "a key appears twice in low.bin or in the root records"
    return keys[: d * nw].reshape(d, nw), counts[:d]


def table_sha(keys: np.ndarray, counts: np.ndarray) -> str:
    nw = keys.shape[1]
    dt = np.dtype([("k", "<u8", (nw,)), ("c", "<u4")])
    rec = np.zeros(keys.shape[0], dtype=dt)
    rec["k"] = keys
    rec["c"] = counts
    return hashlib.sha256(rec.tobytes()).hexdigest()


def histogram(counts: np.ndarray, minf: int, maxf: int) -> np.ndarray:
    O = oracle()
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    h = np.zeros(maxf - minf + 1, dtype=np.int64)
    O.orc_histogram_from_counts(counts.ctypes.data, counts.size, minf, maxf, h.ctypes.data)
    return h


@lru_cache(maxsize=1)
def reference_runs() -> dict:
    return json.loads((GOLDEN / "reference_runs.json").read_text())
