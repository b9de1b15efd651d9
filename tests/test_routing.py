"""Fine buckets of the learned partition (host only, no GPU): roots are the
keys' top byte and each root's share of the fine buckets, its bounds and its
lookup table come from the learned boundaries (kc_debug_learned_routing).
Checked against closed forms, an independent numpy restatement of the learned
CDF, and brute-force routing."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K

ROOT_SPAN = 1 << 56


def route(fb: np.ndarray, bd: np.ndarray, lut: np.ndarray, key: int) -> int:
    """The kernels' lookup: the key's cell in its root's table, then the
    bounds inside the cell stepped over."""
    r = key >> 56
    lo, nsub = int(fb[r]), int(fb[r + 1] - fb[r])
    bounds = bd[lo - r: lo - r + nsub - 1]
    f = int(lut[r, (key - (r << 56)) >> 44])
    while f < nsub - 1 and int(bounds[f]) <= key:
        f += 1
    return lo + f


def brute(fb: np.ndarray, bd: np.ndarray, key: int) -> int:
    r = key >> 56
    lo, nsub = int(fb[r]), int(fb[r + 1] - fb[r])
    bounds = bd[lo - r: lo - r + nsub - 1].astype(object)
    return lo + sum(1 for b in bounds if b <= key)


def cdf(b: np.ndarray, x: int) -> float:
    """Piecewise-linear CDF through (b[j], j), b[4096] read as 2^64 (numpy restatement)."""
    bb = [int(v) for v in b[:4096]] + [1 << 64]
    j = int(np.searchsorted(np.array(bb[:4096], dtype=object), x, side="right")) - 1
    lo, hi = bb[j], bb[j + 1]
    return j + ((x - lo) / (hi - lo) if hi > lo else 0.0)


def check_invariants(fb, bd, lut, nf_target):
    nsub = np.diff(fb.astype(np.int64))
    assert fb[0] == 0 and (nsub >= 1).all() and (nsub <= 1024).all()
    assert bd.size == fb[256] - 256
    for r in range(256):
        b = bd[int(fb[r]) - r: int(fb[r + 1]) - r - 1].astype(object)
        if len(b):
            assert all(r << 56 < v < (r + 1) << 56 for v in b), r
            assert all(b[i] <= b[i + 1] for i in range(len(b) - 1)), r
        # every cell holds the bounds at or below its first key
        cells = np.array([(r << 56) + (c << 44) for c in range(0, 4096, 97)], dtype=object)
        want = [sum(1 for v in b if v <= x) for x in cells]
        assert list(lut[r, ::97]) == want, r


def test_uniform_boundaries_give_equal_fan_out_and_closed_form_bounds():
    b = np.array([j << 52 for j in range(4096)] + [(1 << 64) - 1], dtype=np.uint64)
    fb, bd, lut = K.debug_learned_routing(b, 65536)
    assert (np.diff(fb.astype(np.int64)) == 256).all()
    # root r's bounds sit every 2^48 keys: the prefix partition's next 8 bits
    want = np.array([(r << 56) + (s << 48) for r in range(256) for s in range(1, 256)], dtype=np.uint64)
    np.testing.assert_array_equal(bd, want)
    check_invariants(fb, bd, lut, 65536)


@pytest.mark.parametrize("nf_target", [256, 4096, 65536, 262144])
def test_skewed_boundaries(nf_target):
    rng = np.random.default_rng(nf_target)
    # a skewed key distribution: dense near small keys (as canonical keys are)
    u = rng.random(400_000)
    keys = np.sort((u ** 3 * float(2 ** 64 - 1)).astype(np.uint64))
    keys = np.unique(keys)
    n = keys.size
    b = np.empty(4097, dtype=np.uint64)
    b[0] = 0
    b[1:4096] = keys[(np.arange(1, 4096) * n) // 4096]
    b[4096] = np.uint64(2 ** 64 - 1)
    fb, bd, lut = K.debug_learned_routing(b, nf_target)
    check_invariants(fb, bd, lut, nf_target)
    nsub = np.diff(fb.astype(np.int64))
    # each root's share follows the CDF's mass on its key range (restated here)
    f = np.array([cdf(b, r << 56) for r in range(256)] + [4096.0])
    mass = np.diff(f) / 4096.0
    want = np.clip(np.floor(mass * nf_target + 0.5), 1, 1024)
    assert np.abs(nsub - want).max() <= 1
    # routing by table and bounds equals counting the bounds (brute force),
    # at bounds, next to them and at random keys
    probes = [int(v) for v in rng.choice(keys, 300)]
    for v in bd[:: max(1, bd.size // 200)]:
        probes += [int(v) - 1, int(v), int(v) + 1]
    probes += [0, (1 << 64) - 1] + [r << 56 for r in range(0, 256, 17)]
    for x in probes:
        assert route(fb, bd, lut, x) == brute(fb, bd, x), x
    # the fine buckets balance the keys the boundaries came from wherever the
    # fan-out is large enough to (16 fine buckets per root on average, and
    # none capped at 1024)
    if nf_target >= 4096 and (want < 1024).all() and nf_target <= n // 64:
        occ = np.zeros(int(fb[256]), dtype=np.int64)
        sample = np.sort(rng.choice(keys, 200_000, replace=False))
        for r in range(256):
            lo, hi = np.searchsorted(sample, [np.uint64(r << 56), np.uint64(min((r + 1) << 56, 2 ** 64 - 1))])
            if r == 255:
                hi = sample.size
            part = sample[lo:hi]
            rb = bd[int(fb[r]) - r: int(fb[r + 1]) - r - 1]
            occ[int(fb[r]):int(fb[r + 1])] += np.bincount(np.searchsorted(rb, part, side="right"),
                                                           minlength=int(fb[r + 1] - fb[r]))
        mean = occ.sum() / occ.size
        assert occ.max() <= 1.5 * mean + 8 * np.sqrt(mean), (occ.max(), mean)


def test_mass_in_one_root_is_capped():
    # every learned bucket inside root 0: root 0 takes 1024 fine buckets, the others one
    b = np.array([j << 40 for j in range(4096)] + [(1 << 64) - 1], dtype=np.uint64)
    fb, bd, lut = K.debug_learned_routing(b, 65536)
    nsub = np.diff(fb.astype(np.int64))
    assert nsub[0] == 1024 and (nsub[1:] == 1).all()
    check_invariants(fb, bd, lut, 65536)


def test_bad_boundaries_are_refused():
    b = np.array([j << 52 for j in range(4096)] + [(1 << 64) - 1], dtype=np.uint64)
    bad = b.copy()
    bad[10], bad[11] = bad[11], bad[10]
    with pytest.raises(K.KmcError):
        K.debug_learned_routing(bad, 65536)
    bad = b.copy()
    bad[0] = 1
    with pytest.raises(K.KmcError):
        K.debug_learned_routing(bad, 65536)
    with pytest.raises(K.KmcError):
        K.debug_learned_routing(b, 0)
