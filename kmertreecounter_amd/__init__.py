"""kmertreecounter_amd -- MI355X k-mer counting engine (Python mirror of the C ABI).

The product is the native library ``libkmc.so`` (HIP kernels for gfx950 plus
the C++ host layer) declared by ``include/kmc.h`` / ``include/kmc_host.h``.
This module binds it with ctypes for tests, the benchmark and scripting.  It
has no compute path of its own: when the library is missing, or no GPU is
visible, the calls fail loudly.

Reference seams mirrored here (KmerTreeCounter tree_v4):
  Engine.push_block  <- FastqReader -> parser ring (src/FastqReader.h:367-388)
  Engine.finish      <- pipeline joins + final drain (src/main.cpp:242-264)
  Engine.firsts      <- low.bin contents (src/ExportWriter.h:400-550)
  Engine.root        <- root_{id}.bin contents (src/FinalDrainWriter.h:31-91)
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libkmc.so"
TREE_PATH = PKG / "kmc_tree"

KC_PARTITION_PREFIX = 0
KC_PARTITION_LEARNED = 1
KC_SUB_BITS_NONE = 0xFFFFFFFF
KC_COUNTER_BLOOM = 0
KC_COUNTER_EXACT = 1
KC_COUNTER_SORT = 2
LEARNED_BOUNDS = 4097  # boundaries of a learned partition (4096 buckets)
KC_HISTOGRAM_PRECISE = 0
KC_HISTOGRAM_APPROXIMATE = 1


class KmcError(RuntimeError):
    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_uint32), ("partition", ctypes.c_uint32),
        ("min_count", ctypes.c_uint32), ("max_count", ctypes.c_uint32),
        ("expected_bytes", ctypes.c_uint64), ("sub_bits", ctypes.c_uint32),
        ("device", ctypes.c_uint32), ("classifier_num", ctypes.c_uint32),
        ("pass_keys", ctypes.c_uint32), ("counter", ctypes.c_uint32), ("table_slots", ctypes.c_uint32),
        ("heavy_keys", ctypes.c_uint32), ("learned_min_skew_pct", ctypes.c_uint32), ("max_round_keys", ctypes.c_uint64),
        ("graph", ctypes.c_uint32), ("round_outputs", ctypes.c_uint32), ("est_file_size", ctypes.c_uint64),
        ("learned_max_fine", ctypes.c_uint32), ("learned_steps", ctypes.c_uint32),
    ]


class NoallocResult(ctypes.Structure):
    _fields_ = [("total_kmers", ctypes.c_uint64), ("firsts", ctypes.c_uint64), ("records", ctypes.c_uint64),
                ("records_unfiltered", ctypes.c_uint64), ("spilled_keys", ctypes.c_uint64),
                ("root_off", (ctypes.c_uint64 * 257) * 2)]


class TuneResult(ctypes.Structure):
    _fields_ = [("sub_bits", ctypes.c_uint32), ("counter", ctypes.c_uint32), ("heavy_keys", ctypes.c_uint32),
                ("candidates", ctypes.c_uint32), ("best_ms", ctypes.c_double), ("learned_max_fine", ctypes.c_uint32),
                ("skipped", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_uint32), ("key_words", ctypes.c_uint32),
        ("total_kmers", ctypes.c_uint64), ("firsts", ctypes.c_uint64),
        ("records", ctypes.c_uint64), ("records_unfiltered", ctypes.c_uint64),
        ("spilled_keys", ctypes.c_uint64), ("sequence_bytes", ctypes.c_uint64),
        ("fine_buckets", ctypes.c_uint64), ("max_bucket_keys", ctypes.c_uint64),
        ("mean_bucket_keys", ctypes.c_double),
        ("ms_extract_count", ctypes.c_double), ("ms_extract_scatter", ctypes.c_double),
        ("ms_sub_count", ctypes.c_double), ("ms_sub_scatter", ctypes.c_double),
        ("ms_count", ctypes.c_double), ("ms_compact", ctypes.c_double),
        ("ms_total", ctypes.c_double), ("ms_h2d", ctypes.c_double),
        ("stage0_sample_kmers", ctypes.c_uint64),
        ("prefix_counts", ctypes.c_uint64 * 256), ("bloom_capacity", ctypes.c_uint64 * 256),
        ("prefix_owners", ctypes.c_uint8 * 256),
        ("learned_train_keys", ctypes.c_uint64), ("learned_train_ms", ctypes.c_double),
        ("learned_loss_first", ctypes.c_double), ("learned_loss_last", ctypes.c_double),
        ("learned_sample_max_over_mean", ctypes.c_double), ("prefix_sample_max_over_mean", ctypes.c_double),
        ("learned_sample_distinct_max_over_mean", ctypes.c_double),
        ("prefix_sample_distinct_max_over_mean", ctypes.c_double), ("extra_passes", ctypes.c_uint64),
        ("heavy_buckets", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
        ("learned_split_buckets", ctypes.c_uint64), ("bucket4096_max_over_mean", ctypes.c_double),
        ("hist_tail_keys", ctypes.c_uint64), ("hist_tail_kmers", ctypes.c_uint64),
        ("learned_fallback", ctypes.c_uint64), ("ms_heavy", ctypes.c_double),
        ("rounds_in_device", ctypes.c_uint64), ("round_key_cap", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class SynthSpec(ctypes.Structure):
    _fields_ = [
        ("config", ctypes.c_uint32), ("threads", ctypes.c_uint32), ("seed", ctypes.c_uint64),
        ("scale", ctypes.c_double), ("format", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
        ("genome_len", ctypes.c_uint64), ("n_reads", ctypes.c_uint64),
    ]


class TreeOpts(ctypes.Structure):
    _fields_ = [
        ("fastq_path", ctypes.c_char_p), ("tmp_dir", ctypes.c_char_p), ("k", ctypes.c_uint32),
        ("n_thread", ctypes.c_uint32), ("min_count", ctypes.c_uint32), ("max_count", ctypes.c_uint32),
        ("partition", ctypes.c_uint32), ("verbose", ctypes.c_uint32), ("sample_bytes", ctypes.c_uint64),
        ("counter", ctypes.c_uint32), ("histogram_only", ctypes.c_uint32), ("histogram_path", ctypes.c_char_p),
        ("hist_min", ctypes.c_uint32), ("hist_max", ctypes.c_uint32), ("host_parse", ctypes.c_uint32),
    ]


KEYS_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64)
RECS_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64)
BLOCK_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_char), ctypes.c_uint64)

EXPORTED = [
    "kc_create", "kc_sample_block", "kc_plan", "kc_push_block", "kc_finish", "kc_emit_firsts",
    "kc_emit_root", "kc_stats_get", "kc_destroy", "kc_count_resident", "kc_words_for_k",
    "kc_last_error", "kcs_generate", "kcs_free", "kmc_fastq_blocks",
    "kmc_fastq_blocks_file", "kmc_packed_bytes", "kmc_pack_low", "kmc_write_outputs", "kmc_tree_run",
    "kc_debug_gemm_bf16", "kc_debug_mlp_grad", "kc_debug_learned_fit", "kc_debug_learned_routing", "kc_debug_copy", "kc_histogram",
    "kmc_write_histogram", "kc_autotune", "kc_push_fastq", "kc_debug_resident", "kmc_stage0_sample_file",
    "kc_debug_fault", "kc_device_memory", "kc_count_noalloc", "kc_noalloc_max_bytes", "kmc_write_low_keys",
    "kmc_write_root_records", "kc_get_partition_bounds", "kc_set_partition_bounds",
]

_lib = None

# bytes from a C buffer of any size (ctypes.string_at takes a C int length)
_bytes_from = ctypes.pythonapi.PyBytes_FromStringAndSize
_bytes_from.restype = ctypes.py_object
_bytes_from.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t]


def lib() -> ctypes.CDLL:
    """Load libkmc.so (built in-tree by kmertreecounter_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise KmcError(f"{LIB_PATH} is missing: build it with `python -m kmertreecounter_amd.build`")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, u32, u64, cp = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_char_p
    L.kc_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.kc_sample_block.argtypes = [vp, cp, u64]
    L.kc_plan.argtypes = [vp]
    L.kc_get_partition_bounds.argtypes = [vp, ctypes.POINTER(u64), u32, ctypes.POINTER(u32)]
    L.kc_set_partition_bounds.argtypes = [vp, ctypes.POINTER(u64), u32]
    L.kc_push_block.argtypes = [vp, cp, u64]
    L.kc_push_fastq.argtypes = [vp, cp, u64]
    L.kc_debug_resident.argtypes = [vp, vp, u64, ctypes.POINTER(u64)]
    L.kc_finish.argtypes = [vp]
    L.kc_count_resident.argtypes = [vp]
    L.kc_emit_firsts.argtypes = [vp, KEYS_CB, vp]
    L.kc_emit_root.argtypes = [vp, u32, RECS_CB, vp]
    L.kc_stats_get.argtypes = [vp, ctypes.POINTER(Stats)]
    L.kc_destroy.argtypes = [vp]
    L.kc_words_for_k.argtypes = [u32]
    L.kc_words_for_k.restype = u32
    L.kc_last_error.argtypes = [vp]
    L.kc_last_error.restype = cp
    L.kc_debug_copy.argtypes = [u64, u32, ctypes.POINTER(ctypes.c_float)]
    L.kc_debug_fault.argtypes = [u32, ctypes.c_int64]
    L.kc_debug_fault.restype = ctypes.c_int64
    L.kc_device_memory.argtypes = [u32, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.kc_count_noalloc.argtypes = [ctypes.POINTER(Config), vp, u64, vp, u64, vp, u64, ctypes.POINTER(NoallocResult), vp]
    L.kc_noalloc_max_bytes.argtypes = [u32]
    L.kc_noalloc_max_bytes.restype = u64
    L.kc_histogram.argtypes = [vp, u32, u32, u32, vp]
    L.kc_autotune.argtypes = [vp, vp]
    L.kmc_write_histogram.argtypes = [cp, vp, u32, u32]
    L.kc_debug_gemm_bf16.argtypes = [vp, vp, vp, u32, u32, u32]
    L.kc_debug_mlp_grad.argtypes = [vp, vp, vp, u32, vp, vp]
    L.kc_debug_learned_fit.argtypes = [vp, u64, u32, ctypes.c_float, u32, u64, u32, vp, vp, u32]
    L.kc_debug_learned_routing.argtypes = [vp, u32, vp, vp, u64, vp]
    L.kcs_generate.argtypes = [ctypes.POINTER(SynthSpec), ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(u64),
                               ctypes.POINTER(u64)]
    L.kcs_free.argtypes = [vp]
    L.kmc_fastq_blocks.argtypes = [cp, u64, u32, u64, BLOCK_CB, vp]
    L.kmc_fastq_blocks_file.argtypes = [cp, u32, u64, u64, BLOCK_CB, vp]
    L.kmc_stage0_sample_file.argtypes = [cp, u32, u64, BLOCK_CB, vp, ctypes.POINTER(u64)]
    L.kmc_packed_bytes.argtypes = [u32]
    L.kmc_packed_bytes.restype = u64
    L.kmc_pack_low.argtypes = [vp, u64, u32, vp]
    L.kmc_write_outputs.argtypes = [vp, cp]
    L.kmc_tree_run.argtypes = [ctypes.POINTER(TreeOpts), ctypes.POINTER(Stats), ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def words_for_k(k: int) -> int:
    """Key words N of k (src/main.cpp:579-595): 1 up to 32, 2 up to 64, 4 up to 128."""
    if not 1 <= k <= 128:
        raise ValueError("k_len must be in [1, 128]")
    return 1 if k <= 32 else (2 if k <= 64 else 4)


def record_dtype(nw: int) -> np.dtype:
    """numpy view of ExportRecord<N> (src/definition.h:159-163)."""
    return np.dtype([("key", "<u8", (nw,)), ("count", "<u4"), ("pad", "<u4")])


# ------------------------------------------------------------------ synthetic data

def generate(config: int, scale: float = 1.0, seed: int = 1, fmt: str = "fastq", threads: int | None = None,
             genome_len: int = 0, n_reads: int = 0) -> bytes:
    """Deterministic synthetic input of workload `config` (1..5); fmt 'fastq' or 'seq'."""
    L = lib()
    spec = SynthSpec(config, threads or min(16, os.cpu_count() or 1), seed, float(scale),
                     0 if fmt == "fastq" else 1, 0, genome_len, n_reads)
    out = ctypes.c_void_p()
    n = ctypes.c_uint64()
    nr = ctypes.c_uint64()
    rc = L.kcs_generate(ctypes.byref(spec), ctypes.byref(out), ctypes.byref(n), ctypes.byref(nr))
    if rc != 0:
        raise KmcError(f"kcs_generate failed: {rc}")
    try:
        return _bytes_from(out, n.value)
    finally:
        L.kcs_free(out)


# ------------------------------------------------------------------ host helpers

def fastq_blocks(fastq: bytes, k: int, block_size: int = 32 * 1024) -> list[bytes]:
    """Sequence-only blocks of a FASTQ buffer, as the reference FastqReader emits them."""
    blocks: list[bytes] = []

    @BLOCK_CB
    def cb(_user, data, n):
        blocks.append(ctypes.string_at(data, n))
        return 0

    rc = lib().kmc_fastq_blocks(fastq, len(fastq), k, block_size, cb, None)
    if rc != 0:
        raise KmcError(f"kmc_fastq_blocks failed: {rc}")
    return blocks


def stage0_sample_file(path: str | os.PathLike, k: int, block_size: int = 32 * 1024) -> tuple[list[bytes], int]:
    """Stage-0 sample blocks of a FASTQ file as the reference's FastqPreReader reads
    them, and the reference's estimated file size (the Bloom sizing input)."""
    blocks: list[bytes] = []

    @BLOCK_CB
    def cb(_user, data, n):
        blocks.append(ctypes.string_at(data, n))
        return 0

    est = ctypes.c_uint64()
    rc = lib().kmc_stage0_sample_file(str(path).encode(), k, block_size, cb, None, ctypes.byref(est))
    if rc != 0:
        raise KmcError(f"kmc_stage0_sample_file failed: {rc}")
    return blocks, est.value


def pack_low(keys: np.ndarray, k: int) -> bytes:
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    nw = words_for_k(k)
    n = keys.size // nw
    out = ctypes.create_string_buffer(int(lib().kmc_packed_bytes(k)) * n or 1)
    lib().kmc_pack_low(keys.ctypes.data, n, k, out)
    return out.raw[: int(lib().kmc_packed_bytes(k)) * n]


# ------------------------------------------------------------------ engine

class Engine:
    """One counting job on one GPU (a kc_engine handle)."""

    def __init__(self, k: int, partition: int = KC_PARTITION_PREFIX, min_count: int = 1, max_count: int = 0,
                 expected_bytes: int = 0, sub_bits: int = 0, device: int = 0, classifier_num: int = 3,
                 pass_keys: int = 0, counter: int = 0, table_slots: int = 0, heavy_keys: int = 0,
                 max_round_keys: int = 0, graph: bool = False, est_file_size: int = 0,
                 learned_min_skew_pct: int = 0, round_outputs_host: bool = False, learned_max_fine: int = 0,
                 learned_steps: int = 0):
        self.k = k
        self.nw = words_for_k(k)
        self._L = lib()
        cfg = Config(k, partition, min_count, max_count, expected_bytes, sub_bits, device, classifier_num, pass_keys,
                     counter, table_slots, heavy_keys, learned_min_skew_pct, max_round_keys, 1 if graph else 0,
                     1 if round_outputs_host else 0,
                     est_file_size, learned_max_fine, learned_steps)
        h = ctypes.c_void_p()
        rc = self._L.kc_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise KmcError(f"kc_create failed with status {rc}", rc)
        self._h = h

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self._L.kc_last_error(self._h)
            raise KmcError(f"{what} failed with status {rc}: {msg.decode() if msg else ''}", rc)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.kc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sample_block(self, data: bytes) -> None:
        self._check(self._L.kc_sample_block(self._h, data, len(data)), "kc_sample_block")

    def plan(self) -> None:
        self._check(self._L.kc_plan(self._h), "kc_plan")

    def partition_bounds(self):
        """The learned partition's 4097 boundaries (numpy uint64), or None when
        the handle routes by prefix."""
        import numpy as np
        out = np.zeros(LEARNED_BOUNDS, dtype=np.uint64)
        n = ctypes.c_uint32()
        self._check(self._L.kc_get_partition_bounds(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                    LEARNED_BOUNDS, ctypes.byref(n)), "kc_get_partition_bounds")
        return out if n.value else None

    def set_partition_bounds(self, bounds) -> None:
        """Route by these learned boundaries without training (kc_set_partition_bounds)."""
        import numpy as np
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        self._check(self._L.kc_set_partition_bounds(self._h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                    b.shape[0]), "kc_set_partition_bounds")

    def push_block(self, data: bytes) -> None:
        self._check(self._L.kc_push_block(self._h, data, len(data)), "kc_push_block")

    def push_fastq(self, data: bytes) -> None:
        """Raw FASTQ bytes, parsed on the device (any cut points between calls)."""
        self._check(self._L.kc_push_fastq(self._h, data, len(data)), "kc_push_fastq")

    def resident(self) -> bytes:
        """The resident sequence bytes after finish() (test hook)."""
        n = ctypes.c_uint64()
        self._check(self._L.kc_debug_resident(self._h, None, 0, ctypes.byref(n)), "kc_debug_resident")
        buf = ctypes.create_string_buffer(max(1, n.value))
        self._check(self._L.kc_debug_resident(self._h, buf, n.value, ctypes.byref(n)), "kc_debug_resident")
        return buf.raw[: n.value]

    def finish(self) -> None:
        self._check(self._L.kc_finish(self._h), "kc_finish")

    def count_resident(self) -> None:
        self._check(self._L.kc_count_resident(self._h), "kc_count_resident")

    def autotune(self) -> TuneResult:
        """Time the candidate configurations on the resident input and keep the fastest."""
        r = TuneResult()
        self._check(self._L.kc_autotune(self._h, ctypes.byref(r)), "kc_autotune")
        return r

    def histogram(self, min_freq: int = 1, max_freq: int = 10000, precise: bool = True) -> np.ndarray:
        """histogram_tool's count histogram (precise or approximate) of the counted
        results: element f - min_freq holds the number of k-mers of count f."""
        out = np.zeros(max_freq - min_freq + 1, dtype=np.int64)
        mode = KC_HISTOGRAM_PRECISE if precise else KC_HISTOGRAM_APPROXIMATE
        self._check(self._L.kc_histogram(self._h, mode, min_freq, max_freq, out.ctypes.data), "kc_histogram")
        return out

    def stats(self) -> Stats:
        s = Stats()
        self._check(self._L.kc_stats_get(self._h, ctypes.byref(s)), "kc_stats_get")
        return s

    def firsts(self) -> np.ndarray:
        """First-occurrence keys (the low.bin key set), shape (n, N)."""
        parts: list[np.ndarray] = []
        nw = self.nw

        @KEYS_CB
        def cb(_user, keys, n):
            parts.append(np.ctypeslib.as_array(keys, shape=(n * nw,)).copy())
            return 0

        self._check(self._L.kc_emit_firsts(self._h, cb, None), "kc_emit_firsts")
        flat = np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint64)
        return flat.reshape(-1, nw)

    def root(self, root_id: int) -> np.ndarray:
        """Root records of one root id as a structured ExportRecord<N> array."""
        dt = record_dtype(self.nw)
        parts: list[bytes] = []

        @RECS_CB
        def cb(_user, recs, n):
            parts.append(ctypes.string_at(recs, n * dt.itemsize))
            return 0

        self._check(self._L.kc_emit_root(self._h, root_id, cb, None), "kc_emit_root")
        return np.frombuffer(b"".join(parts), dtype=dt)

    def records(self) -> np.ndarray:
        return np.concatenate([self.root(r) for r in range(256)])

    def write_outputs(self, tmp_dir: str | os.PathLike) -> None:
        self._check(self._L.kmc_write_outputs(self._h, str(tmp_dir).encode()), "kmc_write_outputs")


def count_sequence(seq: bytes, k: int, **kw) -> tuple[np.ndarray, np.ndarray, Stats]:
    """Count a sequence-only buffer ('\\n'-separated lines) on the GPU."""
    with Engine(k, **kw) as e:
        e.push_block(seq)
        e.finish()
        return e.firsts(), e.records(), e.stats()


def noalloc_max_bytes(k: int) -> int:
    """Largest input kc_count_noalloc takes at k."""
    return int(lib().kc_noalloc_max_bytes(k))


def count_noalloc(k: int, d_seq: int, seq_len: int, d_firsts: int, firsts_cap: int, d_records: int,
                  records_cap: int, stream: int | None = None, min_count: int = 1, max_count: int = 0,
                  counter: int = KC_COUNTER_BLOOM, pass_keys: int = 0, table_slots: int = 0,
                  device: int = 0) -> NoallocResult:
    """kc_count_noalloc on device buffers given as integer addresses; returns the
    result struct (raises KmcError with the status on failure)."""
    cfg = Config(k, KC_PARTITION_PREFIX, min_count, max_count, 0, 0, device, 1, pass_keys, counter, table_slots,
                 0, 0, 0, 0, 0, 0)
    res = NoallocResult()
    L = lib()
    rc = L.kc_count_noalloc(ctypes.byref(cfg), ctypes.c_void_p(d_seq), seq_len, ctypes.c_void_p(d_firsts), firsts_cap,
                            ctypes.c_void_p(d_records), records_cap, ctypes.byref(res), ctypes.c_void_p(stream))
    if rc != 0:
        msg = L.kc_last_error(None)
        err = KmcError(f"kc_count_noalloc failed with status {rc}: {msg.decode() if msg else ''}", rc)
        err.result = res
        raise err
    return res


def device_memory(device: int = 0) -> tuple[int, int]:
    """(free, total) bytes of a device's memory."""
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    rc = lib().kc_device_memory(device, ctypes.byref(f), ctypes.byref(t))
    if rc != 0:
        raise KmcError(f"kc_device_memory failed: {rc}", rc)
    return f.value, t.value


FAULT_ALLOC = 1
FAULT_COPY = 2


def debug_fault(kind: int, nth: int) -> int:
    """Make the nth engine allocation (kind 1) or copy (kind 2) from now on fail
    once (nth <= 0: none); returns the calls of that kind since the last call."""
    return int(lib().kc_debug_fault(kind, nth))


def debug_copy(nbytes: int, device: int = 0) -> float:
    """One plain device copy of nbytes (reads and writes nbytes); returns its device ms."""
    ms = ctypes.c_float()
    rc = lib().kc_debug_copy(ctypes.c_uint64(nbytes), ctypes.c_uint32(device), ctypes.byref(ms))
    if rc != 0:
        raise KmcError(f"kc_debug_copy failed: {rc}")
    return ms.value


MLP_NPARAM = 64 * 64 + 64 + 64 * 64 + 64 + 4096 * 64 + 4096


def debug_gemm_bf16(a_bits: np.ndarray, bt_bits: np.ndarray) -> np.ndarray:
    """MFMA product on the GPU: a (M, K) and bt (N, K) as uint16 bf16 bit patterns."""
    a = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b = np.ascontiguousarray(bt_bits, dtype=np.uint16)
    M, K = a.shape
    N = b.shape[0]
    c = np.zeros((M, N), dtype=np.float32)
    rc = lib().kc_debug_gemm_bf16(a.ctypes.data, b.ctypes.data, c.ctypes.data, M, N, K)
    if rc != 0:
        raise KmcError(f"kc_debug_gemm_bf16 failed: {rc}")
    return c


def debug_mlp_grad(params: np.ndarray, keys: np.ndarray, labels: np.ndarray) -> tuple[float, np.ndarray]:
    p = np.ascontiguousarray(params, dtype=np.float32)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    y = np.ascontiguousarray(labels, dtype=np.uint32)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    rc = lib().kc_debug_mlp_grad(p.ctypes.data, k.ctypes.data, y.ctypes.data, k.size, g.ctypes.data, loss.ctypes.data)
    if rc != 0:
        raise KmcError(f"kc_debug_mlp_grad failed: {rc}")
    return float(loss[0]), g


def debug_learned_fit(sorted_keys: np.ndarray, steps: int = 8000, lr: float = 1e-2, batch: int = 2048,
                      seed: int = 1, distinct: bool = True, expect: bool = False) -> tuple[np.ndarray, np.ndarray]:
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    b = np.zeros(4097, dtype=np.uint64)
    cap = steps // 50 + 2
    lc = np.zeros(cap, dtype=np.float32)
    flags = (1 if distinct else 0) | (2 if expect else 0)
    rc = lib().kc_debug_learned_fit(k.ctypes.data, k.size, steps, lr, batch, seed, flags, b.ctypes.data,
                                    lc.ctypes.data, cap)
    if rc != 0:
        raise KmcError(f"kc_debug_learned_fit failed: {rc}")
    return b, lc


def debug_learned_routing(bounds: np.ndarray, nf_target: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Fine buckets of a learned partition from its 4097 boundaries (host only):
    (fine_base[257], fine_bounds, lut[256, 4096]) as kc_debug_learned_routing returns them."""
    b = np.ascontiguousarray(bounds, dtype=np.uint64)
    fb = np.zeros(257, dtype=np.uint32)
    cap = 256 * 1023
    bd = np.zeros(cap, dtype=np.uint64)
    lut = np.zeros((256, 4096), dtype=np.uint16)
    rc = lib().kc_debug_learned_routing(b.ctypes.data, nf_target, fb.ctypes.data, bd.ctypes.data, cap, lut.ctypes.data)
    if rc != 0:
        raise KmcError(f"kc_debug_learned_routing failed: {rc}")
    return fb, bd[: int(fb[256]) - 256].copy(), lut


def write_histogram(path: str, hist: np.ndarray, min_freq: int = 1) -> None:
    """Write a histogram as histogram_tool does: `f<TAB>n` lines for f = min_freq.."""
    h = np.ascontiguousarray(hist, dtype=np.int64)
    rc = lib().kmc_write_histogram(str(path).encode(), h.ctypes.data, min_freq, min_freq + h.size - 1)
    if rc != 0:
        raise KmcError(f"kmc_write_histogram failed: {rc}")


def tree_run(fastq_path: str, tmp_dir: str, k: int, n_thread: int = 8, min_count: int = 1, max_count: int = 0,
             partition: int = KC_PARTITION_PREFIX, verbose: bool = False, sample_bytes: int = 0,
             counter: int = KC_COUNTER_BLOOM, histogram_path: str | None = None, histogram_only: bool = False,
             hist_min: int = 0, hist_max: int = 0, host_parse: bool = False) -> tuple[Stats, float]:
    o = TreeOpts(str(fastq_path).encode(), str(tmp_dir).encode(), k, n_thread, min_count, max_count, partition,
                 1 if verbose else 0, sample_bytes, counter, 1 if histogram_only else 0,
                 str(histogram_path).encode() if histogram_path else None, hist_min, hist_max,
                 1 if host_parse else 0)
    s = Stats()
    wall = ctypes.c_double()
    rc = lib().kmc_tree_run(ctypes.byref(o), ctypes.byref(s), ctypes.byref(wall))
    if rc != 0:
        raise KmcError(f"kmc_tree_run failed: {rc}")
    return s, wall.value
