"""The pipeline captured in a HIP graph and replayed (kc_config.graph): every
replay, with eager runs of another engine and synchronizations between the
replays, must give the table of the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import kmertreecounter_amd as K
from tests import common as C
from tests.test_gpu_parity import check_outputs, microsatellite_reads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["prefix", "learned", "heavy", "spill"])
def test_graph_replays_match_oracle_and_eager(case):
    k = 31
    if case == "heavy":
        seq = microsatellite_reads(100_000, 5)
        kw = dict(heavy_keys=50_000)
    elif case == "spill":
        # one pass per 190k-key bucket: the spill table fills and grows, which
        # captures the graph again
        seq = K.generate(1, scale=1, seed=1, fmt="seq")
        kw = dict(sub_bits=K.KC_SUB_BITS_NONE, pass_keys=1 << 30)
    else:
        seq = K.generate(2, scale=200, seed=12, fmt="seq")
        kw = {}
    keys, counts = C.count_table(C.extract(seq, k))
    part = K.KC_PARTITION_LEARNED if case == "learned" else K.KC_PARTITION_PREFIX
    with K.Engine(k, partition=part, graph=True, learned_min_skew_pct=1, **kw) as g, K.Engine(k, **kw) as eager:
        if part == K.KC_PARTITION_LEARNED:
            g.sample_block(seq[: len(seq) // 20])
            g.plan()
        g.push_block(seq)
        eager.push_block(seq)
        g.finish()
        eager.finish()
        ek, ec = C.merge_outputs(eager.firsts(), eager.records(), 1)
        for replay in range(2 if case == "spill" else 4):
            f, r, s = g.firsts(), g.records(), g.stats()
            check_outputs(f, r, s, k, keys, counts)
            if case == "spill":
                assert s.spilled_keys > 0
            assert s.ms_total > 0
            eager.count_resident()           # eager work between the replays
            ek2, ec2 = C.merge_outputs(eager.firsts(), eager.records(), 1)
            np.testing.assert_array_equal(ek2, ek)
            np.testing.assert_array_equal(ec2, ec)
            g.count_resident()               # graph replay
        f, r, s = g.firsts(), g.records(), g.stats()
        check_outputs(f, r, s, k, keys, counts)
        np.testing.assert_array_equal(g.histogram(1, 500), C.histogram(counts, 1, 500))
