"""GPU parity: radix join and no-partitioning join against the oracle.

Mirrors sql-ops/tests/test_cuda_radix_join.rs (gpu_verify_join_aggregate
:44-185 and its 15 cases) and the CUDA tests of
sql-ops/src/join/no_partitioning_join.rs:816-1104: the reference's closed
form SUM(probe.value) = N(N+1)/2, plus count and SUM(build.value) compared
with the oracle on the same inputs.
"""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import sync, to_dev

pytestmark = pytest.mark.gpu

CHUNKED = fij.HistogramAlgorithm.Chunked


def partition_one_pass(keys, pays, bits, kt):
    """One-pass, single-chunk partitioning as in gpu_verify_join_aggregate."""
    part = fij.GpuRadixPartitioner(CHUNKED, fij.RadixPartitionAlgorithm.NC, fij.RadixBits(bits), 1, 128)
    offs = fij.PartitionOffsets(CHUNKED, 1, bits, kt)
    rel = fij.PartitionedRelation(len(keys), CHUNKED, bits, 1, kt)
    dk, dp = to_dev(keys), to_dev(pays)
    part.prefix_sum(fij.RadixPass.First, dk, offs)
    part.partition(fij.RadixPass.First, dk, dp, offs, rel)
    return rel


def verify_join(build_n, probe_n, scheme, bits, grid, block, dtype=np.int32):
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    rk = orc.gen_perm(build_n, 3, dtype, base=0)       # gen_primary_key: 0..N-1 shuffled
    sk = orc.gen_perm(probe_n, 4, dtype, base=0)       # FK of an equal-size PK: a permutation
    rp = orc.gen_iota(build_n, 1, dtype)
    sp = orc.gen_iota(probe_n, 1, dtype)
    brel = partition_one_pass(rk, rp, bits, kt)
    prel = partition_one_pass(sk, sp, bits, kt)
    result = fij.DeviceArray.zeros(grid * block, np.int64)
    tasks = fij.DeviceArray.zeros(grid + 1, np.uint32)
    agg = fij.aggregates_array(grid)
    j = fij.HipRadixJoin(fij.RadixPass.First, fij.RadixBits(bits), scheme, grid, block)
    j.join(brel, prel, result, tasks, agg)
    sync()
    assert fij.device_status() == 0
    # sql-ops/tests/test_cuda_radix_join.rs:179-182
    assert int(result.to_host().astype(np.int64).sum(dtype=np.int64)) == probe_n * (probe_n + 1) // 2
    count, sb, sp_sum = fij.sum_aggregates(agg)
    assert sp_sum == probe_n * (probe_n + 1) // 2
    assert (count, sb, sp_sum) == orc.sort_merge_join(rk, rp, sk, sp)
    # task assignment equals the oracle's greedy rule on the same offsets
    ref_tasks = orc.assign_tasks(prel.host_offsets(), prel.padded_len(), orc.pad_len(dtype), bits, grid)
    np.testing.assert_array_equal(tasks.to_host(), ref_tasks)


@pytest.mark.parametrize("bits", [0, 1, 2, 3, 4, 5, 8])
def test_smem_perfect(bits):
    verify_join(6100 * 2 ** bits, 6100 * 2 ** bits, fij.HashingScheme.Perfect, bits, 1, 128)


@pytest.mark.parametrize("bits,grid", [(2, 3), (8, 3), (1, 80), (8, 80)])
def test_task_assignment(bits, grid):
    verify_join(6100 * 2 ** bits, 6100 * 2 ** bits, fij.HashingScheme.Perfect, bits, grid, 128)


@pytest.mark.parametrize("bits,grid", [(0, 1), (1, 1), (8, 1), (8, 80)])
def test_smem_bucket_chaining(bits, grid):
    verify_join(4096 * 2 ** bits, 4096 * 2 ** bits, fij.HashingScheme.BucketChaining, bits, grid, 128)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("scheme", [fij.HashingScheme.Perfect, fij.HashingScheme.BucketChaining])
def test_i64_and_wide_blocks(dtype, scheme):
    verify_join(4096 * 64, 4096 * 64, scheme, 6, 256, 256, dtype)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_chaining_duplicates_and_misses(dtype):
    """Build keys with duplicates, probe keys partly outside the build set:
    every matching pair counts (radix_join.cu:291-299)."""
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    n_r, n_s, bits, grid, block = 200_000, 800_000, 7, 64, 256
    rk = orc.gen_uniform(n_r, 9, 1, 1, 50_000, dtype)
    rp = orc.gen_uniform(n_r, 9, 2, -1000, 1 << 20, dtype)
    sk = orc.gen_uniform(n_s, 9, 3, 1, 60_000, dtype)
    sp = orc.gen_uniform(n_s, 9, 4, -1000, 1 << 20, dtype)
    brel = partition_one_pass(rk, rp, bits, kt)
    prel = partition_one_pass(sk, sp, bits, kt)
    agg = fij.aggregates_array(grid)
    tasks = fij.DeviceArray.zeros(grid + 1, np.uint32)
    j = fij.HipRadixJoin(fij.RadixPass.First, fij.RadixBits(bits), fij.HashingScheme.BucketChaining, grid, block)
    j.join(brel, prel, None, tasks, agg)
    sync()
    assert fij.device_status() == 0
    assert fij.sum_aggregates(agg) == orc.sort_merge_join(rk, rp, sk, sp)


def np_case(dtype, scheme, selective):
    """no_partitioning_join.rs test_cuda!: ROWS = 32 MiB / sizeof(T), HT_LEN = 2 * ROWS,
    grid 16, block 1024, payload = i + 1, excluded key 1 when selective."""
    rows = (32 << 20) // np.dtype(dtype).itemsize // 4
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    rk = orc.gen_perm(rows, 21, dtype, base=0)
    sk = orc.gen_perm(rows, 22, dtype, base=0)
    rp = orc.gen_iota(rows, 1, dtype)
    sp = orc.gen_iota(rows, 1, dtype)
    exclude = 0
    if selective:
        pos = np.nonzero(rk == 1)[0][0]
        rk[pos] = -1
        exclude = int(sp[sk == 1].astype(np.int64).sum())
    ht = fij.HashTable(2 * rows, kt)
    hj = fij.HipHashJoin(scheme, ht, (16, 1024), (16, 1024), is_selective=selective)
    hj.build(to_dev(rk), to_dev(rp))
    result = fij.DeviceArray.zeros(16 * 1024, np.uint64)
    agg = fij.aggregates_array(16)
    hj.probe_sum(to_dev(sk), to_dev(sp), result, agg)
    sync()
    expected = rows * (rows + 1) // 2 - exclude
    assert int(result.to_host().astype(np.int64).sum(dtype=np.int64)) == expected
    got = fij.sum_aggregates(agg)
    assert got[2] == expected
    live = rk != -1
    assert got == orc.sort_merge_join(rk[live], rp[live], sk, sp)
    if scheme == fij.HashingScheme.LinearProbing:
        assert got == orc.np_join_lp(rk, rp, sk, sp, entries=2 * rows)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_np_perfect(dtype):
    np_case(dtype, fij.HashingScheme.Perfect, False)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_np_perfect_selective(dtype):
    np_case(dtype, fij.HashingScheme.Perfect, True)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_np_linear_probing(dtype):
    np_case(dtype, fij.HashingScheme.LinearProbing, False)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_np_linear_probing_duplicates(dtype):
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    rk = orc.gen_uniform(300_000, 5, 1, 1, 100_000, dtype)
    rp = orc.gen_uniform(300_000, 5, 2, -50, 1 << 16, dtype)
    sk = orc.gen_uniform(1_000_000, 5, 3, 1, 120_000, dtype)
    sp = orc.gen_uniform(1_000_000, 5, 4, -50, 1 << 16, dtype)
    entries = 1 << 20
    ht = fij.HashTable(entries, kt)
    hj = fij.HipHashJoin(fij.HashingScheme.LinearProbing, ht, (64, 256), (64, 256))
    hj.build(to_dev(rk), to_dev(rp))
    agg = fij.aggregates_array(64)
    hj.probe_sum(to_dev(sk), to_dev(sp), None, agg)
    sync()
    got = fij.sum_aggregates(agg)
    assert got == orc.sort_merge_join(rk, rp, sk, sp)
    assert got == orc.np_join_lp(rk, rp, sk, sp, entries=entries)


@pytest.mark.parametrize("bits,grid", [(13, 300), (13, 1), (14, 1000)])
def test_task_assignment_wide_fanout(bits, grid):
    """8192 partitions (the largest fanout assigned by pointer doubling) and
    16384 (assigned by one wave, 64 partitions at a time): the same tasks as
    the oracle's greedy rule."""
    kt = fij.KeyType.I32
    n = 1 << 20
    sk = orc.gen_uniform(n, 5, 1, 0, 1 << 24, np.int32)
    sp = orc.gen_iota(n, 0, np.int32)
    prel = partition_one_pass(sk, sp, bits, kt)
    tasks = fij.DeviceArray.zeros(grid + 1, np.uint32)
    fij.assign_tasks(prel, bits, grid, tasks)
    sync()
    assert fij.device_status() == 0
    ref = orc.assign_tasks(prel.host_offsets(), prel.padded_len(), orc.pad_len(np.int32), bits, grid)
    np.testing.assert_array_equal(tasks.to_host(), ref)
