"""The HIP library against the golden vectors written by the reference's own
CPU hash join (tests/golden/join_fixtures.json): the batched execution
method, the operator-level radix join and the no-partitioning join."""
import numpy as np
import pytest

import fij
from gpu_helpers import sync, to_dev
from test_golden import FIXTURES, generate

pytestmark = pytest.mark.gpu

CHUNKED = fij.HistogramAlgorithm.Chunked


def want(case):
    return (case["count"], case["sum_build"], case["sum_probe"])


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
@pytest.mark.parametrize("bits", [(9, 9), (6, 0), (3, 4)])
def test_plan_matches_reference(case, bits):
    rk, rp, sk, sp = generate(case)
    plan = fij.JoinPlan(rk.dtype, len(rk), len(sk), bits[0], bits[1], pass1_chunks=8, pass1_block=256)
    got, _ = plan.execute(to_dev(rk), to_dev(rp), to_dev(sk), to_dev(sp))
    assert got == want(case)


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_operator_radix_join_matches_reference(case):
    """prefix_sum + partition (one chunk) + CudaRadixJoin bucket chaining."""
    rk, rp, sk, sp = generate(case)
    kt = fij.KeyType.I32 if rk.dtype == np.int32 else fij.KeyType.I64
    bits = 5
    rels = []
    for k, p in ((rk, rp), (sk, sp)):
        part = fij.GpuRadixPartitioner(CHUNKED, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixBits(bits), 1, 256)
        offs = fij.PartitionOffsets(CHUNKED, 1, bits, kt)
        rel = fij.PartitionedRelation(len(k), CHUNKED, bits, 1, kt)
        dk, dp = to_dev(k), to_dev(p)
        part.prefix_sum(fij.RadixPass.First, dk, offs)
        part.partition(fij.RadixPass.First, dk, dp, offs, rel)
        rels.append(rel)
    grid, block = 16, 256
    agg = fij.aggregates_array(grid)
    tasks = fij.DeviceArray.zeros(grid + 1, np.uint32)
    j = fij.HipRadixJoin(fij.RadixPass.First, fij.RadixBits(bits), fij.HashingScheme.BucketChaining, grid, block)
    j.join(rels[0], rels[1], None, tasks, agg)
    sync()
    assert fij.sum_aggregates(agg) == want(case)


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_np_join_matches_reference(case):
    rk, rp, sk, sp = generate(case)
    kt = fij.KeyType.I32 if rk.dtype == np.int32 else fij.KeyType.I64
    entries = 2 * (1 << max(0, (len(rk) - 1).bit_length()))
    ht = fij.HashTable(entries, kt)
    hj = fij.HipHashJoin(fij.HashingScheme.LinearProbing, ht, (8, 256), (8, 256))
    hj.build(to_dev(rk), to_dev(rp))
    agg = fij.aggregates_array(8)
    hj.probe_sum(to_dev(sk), to_dev(sp), None, agg)
    sync()
    assert fij.sum_aggregates(agg) == want(case)
