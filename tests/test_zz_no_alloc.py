"""kc_count_noalloc (the entry point that allocates nothing) against kc_finish
on the same input, through the C ABI.  The caller's device buffers are made
with the HIP runtime libkmc.so itself is linked against."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C
from tests.test_gpu_parity import as_rows, check_filtered

pytestmark = pytest.mark.gpu

KC_ERR_CAPACITY = -6
H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libkmc.so loaded."""

    def __init__(self):
        K.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 not mapped"
        self.L = ctypes.CDLL(path)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def malloc(self, n: int) -> int:
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(256, n)) == 0
        self.live.append(p.value)
        return p.value

    def h2d(self, dst: int, data: bytes) -> None:
        if data:
            assert self.L.hipMemcpy(ctypes.c_void_p(dst), data, len(data), H2D) == 0

    def d2h(self, src: int, n: int) -> bytes:
        buf = ctypes.create_string_buffer(max(1, n))
        if n:
            assert self.L.hipMemcpy(buf, ctypes.c_void_p(src), n, D2H) == 0
        return buf.raw[:n]

    def free_all(self) -> None:
        for p in self.live:
            self.L.hipFree(ctypes.c_void_p(p))
        self.live = []


@pytest.fixture
def hip():
    h = Hip()
    yield h
    h.free_all()


def noalloc(hip, seq: bytes, k: int, caps=None, **kw):
    nw = C.nwords(k)
    rb = 8 * nw + 8
    d_seq = hip.malloc(len(seq))
    hip.h2d(d_seq, seq)
    fcap, rcap = caps or (max(1, len(seq)), max(1, len(seq)))
    d_f = hip.malloc(fcap * 8 * nw)
    d_r = hip.malloc(rcap * rb)
    res = K.count_noalloc(k, d_seq, len(seq), d_f, fcap, d_r, rcap, **kw)
    firsts = np.frombuffer(hip.d2h(d_f, res.firsts * 8 * nw), dtype=np.uint64).reshape(-1, nw)
    recs = np.frombuffer(hip.d2h(d_r, res.records * rb), dtype=K.record_dtype(nw))
    # the root ranges partition the records, each range holding its root's keys
    covered = np.zeros(res.records, dtype=np.int64)
    for part in range(2):
        off = np.array(res.root_off[part][:], dtype=np.int64)
        assert (np.diff(off) >= 0).all()
        for r in range(256):
            lo, hi = off[r], off[r + 1]
            covered[lo:hi] += 1
            if hi > lo:
                assert (recs["key"][lo:hi, 0] >> np.uint64(56) == r).all()
    assert (covered == 1).all()
    return firsts, recs, res


def engine(seq: bytes, k: int, **kw):
    with K.Engine(k, sub_bits=K.KC_SUB_BITS_NONE, **kw) as e:
        e.push_block(seq)
        e.finish()
        return e.firsts(), e.records(), e.stats()


def same_outputs(a, b, nw, exact):
    fa, ra = a[0], a[1]
    fb, rb = b[0], b[1]
    ka, ca = C.merge_outputs(fa, ra, nw)
    kb, cb = C.merge_outputs(fb, rb, nw)
    np.testing.assert_array_equal(ka, kb)
    np.testing.assert_array_equal(ca, cb)
    if exact:  # the exact counter's split is deterministic: same key sets and records
        np.testing.assert_array_equal(np.sort(as_rows(fa.copy())), np.sort(as_rows(fb.copy())))
        oa = np.argsort(as_rows(ra["key"].reshape(-1, nw)))
        ob = np.argsort(as_rows(rb["key"].reshape(-1, nw)))
        np.testing.assert_array_equal(ra[oa], rb[ob])


@pytest.mark.parametrize("counter", [K.KC_COUNTER_EXACT, K.KC_COUNTER_BLOOM], ids=["exact", "bloom"])
@pytest.mark.parametrize("k", [31, 63, 100])
def test_noalloc_matches_finish_twice(hip, k, counter):
    nw = C.nwords(k)
    big = K.generate(2, scale=400, seed=40 + k, fmt="seq")[: K.noalloc_max_bytes(k) - 1000]
    small = K.generate(1, scale=300, seed=41 + k, fmt="seq")
    assert len(small) < len(big) // 2
    for seq, minc, maxc in ((big, 1, 0), (small, 2, 30)):  # two calls in a row, the second smaller
        keys, counts = C.count_table(C.extract(seq, k))
        f, r, res = noalloc(hip, seq, k, counter=counter, min_count=minc, max_count=maxc)
        rf, rr, rs = engine(seq, k, counter=counter, min_count=minc, max_count=maxc)
        assert res.total_kmers == rs.total_kmers == int(counts.sum())
        # the record set of the filter (append_export_record), as kc_finish gives it
        check_filtered(f, r, NoallocStats(res), k, keys, counts, minc, maxc)
        check_filtered(rf, rr, rs, k, keys, counts, minc, maxc)
        if minc == 1 and maxc == 0:
            same_outputs((f, r), (rf, rr), nw, counter == K.KC_COUNTER_EXACT)
        elif counter == K.KC_COUNTER_EXACT:
            np.testing.assert_array_equal(np.sort(as_rows(r["key"].reshape(-1, nw))),
                                          np.sort(as_rows(rr["key"].reshape(-1, nw))))


class NoallocStats:
    """The kc_stats fields check_filtered reads, from a kc_noalloc_result."""

    def __init__(self, res):
        self.total_kmers = res.total_kmers
        self.firsts = res.firsts
        self.records = res.records
        self.records_unfiltered = res.records_unfiltered


def test_noalloc_rejects_what_does_not_fit(hip):
    k = 31
    limit = K.noalloc_max_bytes(k)
    assert limit == 2 << 20 and K.noalloc_max_bytes(63) == 1 << 20 and K.noalloc_max_bytes(100) == 512 << 10
    seq = K.generate(1, scale=10, seed=5, fmt="seq")[: limit + 4096]
    with pytest.raises(K.KmcError) as ei:
        noalloc(hip, seq, k)
    assert ei.value.status == KC_ERR_CAPACITY
    # output buffers too small: the counts still come back
    seq = seq[:200_000]
    with pytest.raises(K.KmcError) as ei:
        noalloc(hip, seq, k, caps=(100, 100))
    assert ei.value.status == KC_ERR_CAPACITY
    assert ei.value.result.total_kmers == int(C.oracle().orc_count_kmers(seq, len(seq), k))
    assert ei.value.result.firsts > 100
    # and the next call with room is exact
    keys, counts = C.count_table(C.extract(seq, k))
    f, r, _ = noalloc(hip, seq, k)
    mk, mc = C.merge_outputs(f, r, 1)
    np.testing.assert_array_equal(mk, keys)
    np.testing.assert_array_equal(mc, counts)
