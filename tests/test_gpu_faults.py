"""Failure handling, forced from the host: the engine's allocations and copies
are made to fail one at a time through kc_debug_fault, as the runtime would
refuse them (no GPU fault is involved).  The reference exits or throws on the
same failures (src/FastqReader.h:78-90, src/ExportWriter.h:595-599,
src/BloomFilter.h:45-48); the engine must return KC_ERR_NOMEM or KC_ERR_HIP,
free what it allocated, and stay usable: a failed kc_finish is followed by a
correct kc_finish on the same handle, and a new handle counts exactly."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C

pytestmark = pytest.mark.gpu

KC_ERR_HIP, KC_ERR_NOMEM = -2, -3


def device_free_bytes() -> int:
    return K.device_memory(0)[0]


MODES = {
    # name: engine options, learned plan, device FASTQ parsing
    "prefix": (dict(), False, False),
    "learned": (dict(partition=K.KC_PARTITION_LEARNED, learned_min_skew_pct=1), True, False),
    "fastq_rounds": (dict(max_round_keys=600_000), False, True),
}


def inputs(mode):
    if MODES[mode][2]:
        fq = K.generate(2, scale=400, seed=19)
        seq = b"".join(K.fastq_blocks(fq, 31, 1 << 30))
        return fq, seq
    seq = K.generate(2, scale=400, seed=19, fmt="seq")
    return None, seq


def run_job(mode, fq, seq, stage):
    """One job; `stage` records how far it got.  Returns the engine (open) and
    its merged table and histogram."""
    opts, learned, parse = MODES[mode]
    stage["at"] = "create"
    e = K.Engine(31, **opts)
    stage["engine"] = e
    if learned:
        stage["at"] = "plan"
        e.sample_block(seq[: len(seq) // 4])
        e.plan()
    stage["at"] = "push"
    if parse:
        for i in range(0, len(fq), 1 << 20):
            e.push_fastq(fq[i: i + (1 << 20)])
    else:
        e.push_block(seq)
    stage["at"] = "finish"
    e.finish()
    stage["at"] = "emit"
    table = C.merge_outputs(e.firsts(), e.records(), 1)
    hist = e.histogram(1, 200)
    return e, table, hist


def assert_exact(table, hist, keys, counts):
    np.testing.assert_array_equal(table[0], keys)
    np.testing.assert_array_equal(table[1], counts)
    np.testing.assert_array_equal(hist, C.histogram(counts, 1, 200))


@pytest.mark.parametrize("kind", [K.FAULT_ALLOC, K.FAULT_COPY], ids=["alloc", "copy"])
@pytest.mark.parametrize("mode", sorted(MODES))
def test_every_failure_point(kind, mode):
    fq, seq = inputs(mode)
    keys, counts = C.count_table(C.extract(seq, 31))
    K.debug_fault(kind, 0)
    e, table, hist = run_job(mode, fq, seq, {})
    e.close()
    assert_exact(table, hist, keys, counts)
    calls = K.debug_fault(kind, 0)
    assert calls > 0
    # every point where it is cheap; the learned plan's training copies are sampled
    points = list(range(1, calls + 1))
    cap = 24 if MODES[mode][1] else 60  # a learned job trains for about a second
    if len(points) > cap:
        head = cap * 2 // 3
        points = points[:head] + points[head::max(1, (len(points) - head) // (cap - head))]
    free0 = device_free_bytes()
    failures = 0
    for nth in points:
        stage: dict = {}
        K.debug_fault(kind, nth)
        err = None
        try:
            e, table, hist = run_job(mode, fq, seq, stage)
        except K.KmcError as x:
            err = x
        finally:
            K.debug_fault(kind, 0)
        e = stage.get("engine")
        if err is None:
            # a failure the engine absorbs (pinned host memory falls back to
            # pageable) must leave the results exact
            assert_exact(table, hist, keys, counts)
        else:
            failures += 1
            assert err.status in (KC_ERR_NOMEM, KC_ERR_HIP), (nth, stage["at"], str(err))
            if e is not None and stage["at"] in ("finish", "emit"):
                # the input was pushed whole: the same handle finishes again
                e.finish()
                assert_exact(C.merge_outputs(e.firsts(), e.records(), 1), e.histogram(1, 200), keys, counts)
        if e is not None:
            e.close()
    assert failures > 0
    # a new handle counts exactly, and nothing stayed allocated
    e, table, hist = run_job(mode, fq, seq, {})
    e.close()
    assert_exact(table, hist, keys, counts)
    assert device_free_bytes() >= free0 - (64 << 20)


def test_fault_hook_rejects_bad_kind():
    assert K.debug_fault(0, 1) == -1
    assert K.debug_fault(3, 1) == -1


def test_autotune_skips_a_candidate_that_does_not_fit():
    # a candidate whose allocation fails (as a spill table that cannot grow on a
    # full device would) is skipped; the others are timed and the kept
    # configuration counts exactly
    seq = K.generate(2, scale=400, seed=19, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    with K.Engine(31) as e:
        e.push_block(seq)
        e.finish()
        K.debug_fault(K.FAULT_ALLOC, 0)
        full = e.autotune()
        calls = K.debug_fault(K.FAULT_ALLOC, 0)
    assert full.skipped == 0 and full.candidates > 1
    assert calls > 0
    skipped_seen = False
    for nth in range(1, calls + 1):
        with K.Engine(31) as e:
            e.push_block(seq)
            e.finish()
            K.debug_fault(K.FAULT_ALLOC, nth)
            try:
                r = e.autotune()
            finally:
                K.debug_fault(K.FAULT_ALLOC, 0)
            assert r.candidates + r.skipped == full.candidates, (nth, r.candidates, r.skipped)
            assert_exact(C.merge_outputs(e.firsts(), e.records(), 1), e.histogram(1, 200), keys, counts)
        if r.skipped:
            skipped_seen = True
            break
    assert skipped_seen


def test_a_refused_device_allocation_leaves_no_error_behind():
    # the runtime refuses a 64 TiB sequence buffer (nothing is allocated); the
    # engine reports KC_ERR_NOMEM, and the runtime's last-error state must not
    # surface as a launch failure in the next job on the same device
    with pytest.raises(K.KmcError) as x:
        K.Engine(31, expected_bytes=64 << 40)
    assert x.value.status == KC_ERR_NOMEM
    seq = K.generate(2, scale=400, seed=19, fmt="seq")
    keys, counts = C.count_table(C.extract(seq, 31))
    with K.Engine(31) as e:
        e.push_block(seq)
        e.finish()
        assert_exact(C.merge_outputs(e.firsts(), e.records(), 1), e.histogram(1, 200), keys, counts)
