"""GPU parity of the batched radix-join execution method (fij_join_plan_*)
against the oracle: same inputs, identical count and wrapping sums.

Mirrors radix-join/tests/test_gpu_radix_join.rs (closed form N(N+1)/2 for
PK/FK joins, :146-149) and adds duplicate keys, misses, negative payloads,
single-pass plans and the benchmark shapes (2^20 x 2^24 int32 in one pass of 8 bits,
2^27 x 2^27 int64 in 9 + 9 bits at 1/16 size).
"""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import to_dev

pytestmark = pytest.mark.gpu


def run_plan(rk, rp, sk, sp, bits1, bits2, **kw):
    plan = fij.JoinPlan(rk.dtype, len(rk), len(sk), bits1, bits2, **kw)
    got, timing = plan.execute(to_dev(rk), to_dev(rp), to_dev(sk), to_dev(sp))
    assert fij.device_status() == 0
    return got, timing


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [(9, 9), (6, 5), (4, 0), (8, 0), (1, 1)])
def test_plan_pk_fk_closed_form(dtype, bits):
    n = 1 << 20
    rk = orc.gen_perm(n, 1, dtype, 1)
    sk = orc.gen_perm(n, 2, dtype, 1)
    rp = orc.gen_iota(n, 1, dtype)
    sp = orc.gen_iota(n, 1, dtype)
    got, _ = run_plan(rk, rp, sk, sp, *bits)
    assert got == (n, n * (n + 1) // 2, n * (n + 1) // 2)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [(9, 9), (7, 3), (5, 0)])
def test_plan_duplicates_misses(dtype, bits):
    rk = orc.gen_uniform(300_000, 3, 1, 1, 100_000, dtype)
    rp = orc.gen_uniform(300_000, 3, 2, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(1_300_000, 3, 3, 1, 130_000, dtype)
    sp = orc.gen_uniform(1_300_000, 3, 4, -(1 << 20), 1 << 21, dtype)
    got, _ = run_plan(rk, rp, sk, sp, *bits)
    assert got == orc.sort_merge_join(rk, rp, sk, sp)
    assert got == orc.radix_join(rk, rp, sk, sp, bits[0], bits[1], threads=4)


def test_plan_2e20_x_2e24_i32_one_pass_8_bits():
    """CPU-runnable shape: R 2^20 permutation of 1..2^20 (payload = key), S 2^24
    uniform over R's keys (payload = row index), one pass of 8 bits."""
    n_r, n_s = 1 << 20, 1 << 24
    rk = orc.gen_perm(n_r, 42, np.int32, 1)
    rp = rk.copy()
    sk = orc.gen_uniform(n_s, 42, 1, 1, n_r, np.int32)
    sp = orc.gen_iota(n_s, 0, np.int32)
    got, _ = run_plan(rk, rp, sk, sp, 8, 0)
    assert got == orc.radix_join(rk, rp, sk, sp, 8, 0, threads=8)
    assert got[0] == n_s


def test_plan_2e23_i64_9_9_bits():
    """The 2^27 x 2^27 int64 shape at 1/16 size: R = S = 2^23, 9 + 9 bits."""
    n = 1 << 23
    rk = orc.gen_perm(n, 42, np.int64, 1)
    rp = rk.copy()
    sk = orc.gen_uniform(n, 42, 1, 1, n, np.int64)
    sp = orc.gen_iota(n, 0, np.int64)
    got, _ = run_plan(rk, rp, sk, sp, 9, 9)
    assert got == orc.radix_join(rk, rp, sk, sp, 9, 9, threads=8)


def test_plan_repeat_and_chunks():
    """A plan can be executed repeatedly; chunk count and block size change
    nothing in the result."""
    n = 1 << 19
    rk = orc.gen_perm(n, 5, np.int64, 1)
    sk = orc.gen_uniform(2 * n, 5, 1, 1, n, np.int64)
    rp = orc.gen_uniform(n, 5, 2, 0, 1 << 30, np.int64)
    sp = orc.gen_iota(2 * n, 0, np.int64)
    want = orc.sort_merge_join(rk, rp, sk, sp)
    for chunks, block in ((256, 1024), (512, 512), (100, 256)):
        plan = fij.JoinPlan(np.int64, n, 2 * n, 9, 9, pass1_chunks=chunks, pass1_block=block)
        dev = [to_dev(a) for a in (rk, rp, sk, sp)]
        for _ in range(2):
            got, _ = plan.execute(*dev)
            assert got == want


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [(9, 9), (6, 0)])
def test_plan_empty_sides(dtype, bits):
    """An empty build side, an empty probe side and both (zero-length views
    of real buffers) through a plan of 2^16 x 2^18 tuples join nothing; the
    full join through the same plan afterwards equals the oracle's, so the
    empty joins left no state behind."""
    n_r, n_s = 1 << 16, 1 << 18
    rk = orc.gen_perm(n_r, 9, dtype, 1)
    rp = orc.gen_uniform(n_r, 9, 2, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(n_s, 9, 3, 1, n_r + n_r // 4, dtype)
    sp = orc.gen_iota(n_s, 0, dtype)
    d = [to_dev(a) for a in (rk, rp, sk, sp)]
    e = [a.view(0, 0) for a in d]
    plan = fij.JoinPlan(dtype, n_r, n_s, *bits)
    res = fij.DeviceArray(3, np.uint64)
    for cols in ((e[0], e[1], d[2], d[3]), (d[0], d[1], e[2], e[3]), tuple(e)):
        plan.enqueue_sized(*cols, res)
        assert tuple(int(x) for x in res.to_host()) == (0, 0, 0)
    plan.enqueue_sized(*d, res)
    assert tuple(int(x) for x in res.to_host()) == orc.sort_merge_join(rk, rp, sk, sp)
    assert fij.device_status() == 0


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_null_keys_never_join(dtype):
    rk = orc.gen_uniform(200_000, 4, 1, -3, 50_000, dtype)
    sk = orc.gen_uniform(600_000, 4, 2, -3, 50_000, dtype)
    rp = orc.gen_iota(200_000, 1, dtype)
    sp = orc.gen_iota(600_000, 1, dtype)
    want = orc.sort_merge_join(rk, rp, sk, sp)
    got, _ = run_plan(rk, rp, sk, sp, 6, 5)
    assert got == want
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    ht = fij.HashTable(1 << 19, kt)
    hj = fij.HipHashJoin(fij.HashingScheme.LinearProbing, ht, (64, 256), (64, 256))
    hj.build(to_dev(rk), to_dev(rp))
    agg = fij.aggregates_array(64)
    hj.probe_sum(to_dev(sk), to_dev(sp), None, agg)
    fij.synchronize()
    assert fij.sum_aggregates(agg) == want


def test_plan_skewed_probe():
    """One probe key carries a fifth of S: its sub-partition is probed by one
    work item of the join."""
    n = 1 << 18
    rk = orc.gen_perm(n, 9, np.int64, 1)
    rp = rk.copy()
    sk = orc.gen_uniform(4 * n, 9, 1, 1, n, np.int64)
    sk[: (4 * n) // 5] = 777
    sp = orc.gen_iota(4 * n, 0, np.int64)
    got, _ = run_plan(rk, rp, sk, sp, 9, 9)
    assert got == orc.sort_merge_join(rk, rp, sk, sp)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_plan_deferred_many_runs(dtype):
    """A 2^24 probe side over two pass-1 partitions (1 + 12 bits): each of an
    item's 1,024 probe runs (one per pass-2 tile) holds about two tuples, so
    every item is deferred and the block join takes its runs in two batches
    of 512; its 16 build runs form one batch.  Closed form: every probe key
    is one of R's keys, and R's payload is its key."""
    n_r, n_s = 1 << 18, 1 << 24
    rk = orc.gen_perm(n_r, 21, dtype, 1)
    rp = rk.copy()
    sk = orc.gen_uniform(n_s, 21, 1, 1, n_r, dtype)
    sp = orc.gen_iota(n_s, 0, dtype)
    got, _ = run_plan(rk, rp, sk, sp, 1, 12)
    assert got == (n_s, int(sk.astype(np.int64).sum()), n_s * (n_s - 1) // 2)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("log2r,log2s", [(16, 20), (17, 20), (20, 16), (20, 17), (16, 16), (20, 20)])
def test_plan_register_join_shapes(dtype, log2r, log2s):
    """The register join gathers per side what the items need (2, 4 or 10
    tuples per lane from the average item side): build-light, probe-light,
    both light and both full shapes at 6 + 5 bits, each against the
    sort-merge join; the Zipf probe side puts heavy items past a light
    side's capacity, which then go to the deferred kernels."""
    n_r, n_s = 1 << log2r, 1 << log2s
    rk = orc.gen_perm(n_r, 5, dtype, 1)
    rp = orc.gen_uniform(n_r, 5, 2, -(1 << 20), 1 << 21, dtype)
    for probe in ("uniform", "zipf"):
        if probe == "uniform":
            sk = orc.gen_uniform(n_s, 6, 1, 1, n_r + n_r // 4, dtype)  # a fifth of the probes miss
        else:
            sk = orc.gen_zipf_keys(n_s, n_r, 1.0, 6, 1, 5, dtype=dtype)  # ranks over R's keys
        sp = orc.gen_iota(n_s, 3, dtype)
        got, _ = run_plan(rk, rp, sk, sp, 6, 5)
        assert got == orc.sort_merge_join(rk, rp, sk, sp), probe


@pytest.mark.parametrize("skewed", [False, True])
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_single_pass_big_block_tables(dtype, skewed):
    """Single-pass plans whose build partitions need the 128 KiB block tables
    (fij_join_plan_bits: |R| of 1.5 x 2^20 int64 or 3 x 2^20 int32 at 9 bits).
    Skewed: a quarter of the build keys fall into one partition, which no
    block table holds (the wave-table path joins it)."""
    n = (3 << 19) if dtype == np.int64 else (3 << 20)
    assert fij.plan_bits(dtype, n, 4 * n) == (9, 0)
    rk = orc.gen_perm(n, 21, dtype, 1)
    if skewed:
        rk[: n // 4] = (np.arange(n // 4, dtype=np.int64) * 512 + (1 << 30)).astype(dtype)  # all in partition 0
    rp = orc.gen_uniform(n, 21, 1, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(4 * n, 21, 2, 1, n + n // 4, dtype)
    if skewed:
        sk[: n // 2] = rk[: n // 2]
    sp = orc.gen_iota(4 * n, 3, dtype)
    got, _ = run_plan(rk, rp, sk, sp, 9, 0)
    assert got == orc.sort_merge_join(rk, rp, sk, sp)
