"""Batched joins (fij_join_batch_*): many independent joins of uneven sizes
with one submission must give, per join, exactly the result of joining that
pair alone -- checked against the oracle's sort-merge join and the plain
plan -- with and without recording the submission into a HIP graph."""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import to_dev

pytestmark = pytest.mark.gpu


def make_join(i, n_r, factor, dtype, seed=1000):
    rk = orc.gen_perm(n_r, seed + i, dtype, 1)
    rp = orc.gen_uniform(n_r, seed + i, 1, -(1 << 20), 1 << 21, dtype)
    n_s = n_r * factor
    sk = orc.gen_uniform(n_s, seed + i, 2, 1, n_r + n_r // 4, dtype)  # a fifth of the keys miss
    sp = orc.gen_iota(n_s, i, dtype)
    return rk, rp, sk, sp


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_batch_matches_each_join_alone(dtype, use_graph, lanes):
    rng = np.random.default_rng(7)
    sizes = [int(2 ** rng.uniform(6, 18)) for _ in range(40)] + [0, 1, 5000, 12288, 12289, 1 << 17]
    host = [make_join(i, max(n, 0), int(rng.integers(4, 17)), dtype) for i, n in enumerate(sizes)]
    dev = [tuple(to_dev(a) if len(a) else fij.DeviceArray(1, dtype) for a in j) for j in host]
    # empty columns: pass zero-length views of real buffers
    dev = [tuple(d if len(h) else d.view(0, 0) for d, h in zip(dj, hj)) for dj, hj in zip(dev, host)]
    # lanes: the size classes spread over one or four streams
    batch = fij.JoinBatch(dtype, len(host), use_graph=use_graph, lanes=lanes)
    for _ in range(2):  # the second execution reuses the pooled plans
        got, stats = batch.execute(dev)
        for j, (h, g) in enumerate(zip(host, got)):
            assert g == orc.sort_merge_join(*h), f"join {j} ({len(h[0])} x {len(h[2])})"
    assert stats["small_joins"] + stats["partitioned_joins"] == len(host)
    assert stats["small_joins"] > 0 and stats["partitioned_joins"] > 0
    assert fij.device_status() == 0


def test_batch_equals_plain_plan():
    dtype = np.int64
    host = [make_join(i, n, 8, dtype) for i, n in enumerate([3000, 40000, 300000])]
    dev = [tuple(to_dev(a) for a in j) for j in host]
    got, _ = fij.JoinBatch(dtype, 3).execute(dev)
    for h, d, g in zip(host, dev, got):
        small, b1, b2 = fij.autotune(dtype, len(h[0]), len(h[2]))
        plan = fij.JoinPlan(dtype, len(h[0]), len(h[2]), b1, b2)
        want, _ = plan.execute(*d)
        assert g == want


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_single_pass_classes(dtype):
    """Build sides from 2^18 tuples up to what block tables hold after one
    pass run as single-pass classes (fij_join_plan_bits), beside two-pass and
    small joins in the same submission; duplicates and null keys included.
    Batched classes stop where 64 KiB block tables stop holding the build
    partitions."""
    one_pass_max = (1 << 21) if dtype == np.int32 else (1 << 20)
    sizes = [1 << 18, 300001, one_pass_max, one_pass_max + 1, 20000, 3000]
    host = [make_join(i, n, 2 + i % 3, dtype, seed=2000) for i, n in enumerate(sizes)]
    rk = orc.gen_uniform(400000, 77, 1, -3, 150000, dtype)  # duplicates and null keys
    host.append((rk, orc.gen_iota(400000, 1, dtype), orc.gen_uniform(800000, 77, 2, -3, 150000, dtype),
                 orc.gen_iota(800000, 1, dtype)))
    classes = [fij.autotune(dtype, len(h[0]), len(h[2])) for h in host]
    assert [c[2] == 0 for c in classes] == [True, True, True, False, False, False, True]
    dev = [tuple(to_dev(a) for a in j) for j in host]
    got, stats = fij.JoinBatch(dtype, len(host), lanes=2).execute(dev)
    for j, (h, g) in enumerate(zip(host, got)):
        assert g == orc.sort_merge_join(*h), f"join {j} ({len(h[0])} x {len(h[2])})"
    assert fij.device_status() == 0


def test_batch_null_keys_and_duplicates():
    dtype = np.int32
    host = []
    for i, n in enumerate([2000, 50000]):
        rk = orc.gen_uniform(n, 5 + i, 1, -3, n // 2, dtype)
        rp = orc.gen_iota(n, 1, dtype)
        sk = orc.gen_uniform(4 * n, 5 + i, 2, -3, n // 2, dtype)
        sp = orc.gen_iota(4 * n, 1, dtype)
        host.append((rk, rp, sk, sp))
    got, _ = fij.JoinBatch(dtype, 2).execute([tuple(to_dev(a) for a in j) for j in host])
    assert got == [orc.sort_merge_join(*h) for h in host]




def views_at_odd_offsets(host_joins, dtype):
    """All columns in one device buffer, each starting 1-3 elements past the
    previous column's end (aligned to the element size only)."""
    cols = [a for j in host_joins for a in j]
    gaps = [1 + i % 3 for i in range(len(cols))]
    total = sum(len(c) for c in cols) + sum(gaps)
    buf = np.zeros(total, dtype)
    at, starts = 0, []
    for c, g in zip(cols, gaps):
        at += g
        buf[at:at + len(c)] = c
        starts.append(at)
        at += len(c)
    dev = to_dev(buf)
    views = [dev.view(s_, s_ + len(c)) for s_, c in zip(starts, cols)]
    return [tuple(views[4 * i:4 * i + 4]) for i in range(len(host_joins))]


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_batch_unaligned_columns_and_growth(dtype):
    """Partitioned joins of several size classes, columns at odd element
    offsets; a second, larger submission grows the size classes' pooled
    buffers, a third, smaller one reuses them."""
    rng = np.random.default_rng(11)
    small = [make_join(i, int(2 ** rng.uniform(14, 16)), int(rng.integers(4, 9)), dtype, seed=300) for i in range(6)]
    large = [make_join(i, int(2 ** rng.uniform(14, 19)), int(rng.integers(4, 17)), dtype, seed=400) for i in range(20)]
    batch = fij.JoinBatch(dtype, 20)
    for host in (small, large, small[:3]):
        got, stats = batch.execute(views_at_odd_offsets(host, dtype))
        assert stats["partitioned_joins"] == len(host)
        for j, (h, g) in enumerate(zip(host, got)):
            assert g == orc.sort_merge_join(*h), f"join {j} ({len(h[0])} x {len(h[2])})"
    assert stats["size_classes"] >= 1
    assert fij.device_status() == 0


def test_batch_deferred_items():
    """Heavy build and probe keys inside batched partitioned joins: their
    work items exceed the register join and are joined by the deferred
    kernels, whose counts and sums go to each join's own result."""
    dtype = np.int32
    host = []
    # (n, copies of the heavy key on the build side, on the probe side):
    # a probe-heavy item with a short build side (block-table units), one
    # with more build tuples than a block table holds (wave-table units), a
    # heavy probe key against a single build tuple; the last two share a
    # size class, so one batched plan routes two joins' deferred items
    for i, (n, hb, hp) in enumerate([(30000, 800, 40000), (60000, 7000, 5000), (90000, 1, 100000),
                                     (100000, 2000, 20000)]):
        rk = orc.gen_perm(n, 70 + i, dtype, 1)
        rk[:hb] = 7 + i
        rp = orc.gen_iota(n, 1, dtype)
        sk = orc.gen_uniform(8 * n, 70 + i, 2, 1, n, dtype)
        sk[:hp] = 7 + i
        sp = orc.gen_iota(8 * n, 0, dtype)
        host.append((rk, rp, sk, sp))
    got, stats = fij.JoinBatch(dtype, len(host)).execute([tuple(to_dev(a) for a in j) for j in host])
    assert stats["partitioned_joins"] == len(host)
    assert got == [orc.sort_merge_join(*h) for h in host]


def test_plan_enqueue_sized_matches_execute():
    """fij_join_plan_enqueue_sized: joins shorter than the plan, result summed
    into a device address (the form the batch records into its graph)."""
    dtype = np.int64
    plan = fij.JoinPlan(dtype, 1 << 16, 1 << 18, 6, 5)
    res = fij.DeviceArray(3, np.uint64)
    for i, (n_r, f) in enumerate([(1 << 16, 4), (1000, 5), (30000, 8)]):
        h = make_join(i, n_r, f, dtype)
        d = [to_dev(a) for a in h]
        plan.enqueue_sized(*d, res)
        got = tuple(int(x) for x in res.to_host())
        assert got == orc.sort_merge_join(*h)
        # a plan of exactly these lengths (execute reads the plan's lengths)
        want, _ = fij.JoinPlan(dtype, len(h[0]), len(h[2]), 6, 5).execute(*d)
        assert got == want
    with pytest.raises(ValueError):  # shorter columns than the plan's: refused before any launch
        plan.execute(*d)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_every_tuning_candidate_gives_the_same_results(dtype):
    """The autotuner's grid for a size class (radix bits moved by one either
    way and split three ways, ring caps 0 and 4, blocks 512 and 1024, and the
    small path where it qualifies): every configuration joins every pair of
    the class exactly as the oracle does."""
    sys_path_pkg()
    import autotune

    n_r, factor = 6000, 10  # class (13, 16): qualifies for the small path
    host = [make_join(i, n_r, factor, dtype, seed=77) for i in range(6)]
    dev = [tuple(to_dev(a) for a in j) for j in host]
    want = [orc.sort_merge_join(*h) for h in host]
    rule = fij.tuning_lookup(dtype, n_r, n_r * factor)
    cands = autotune.candidates(rule, n_r, n_r * factor) if dtype == np.int32 else [
        c for c in autotune.candidates(rule, n_r, n_r * factor) if not c["small"]]
    assert len(cands) >= 20
    try:
        for cfg in cands:
            fij.tuning_set(dtype, [dict(cfg, build_log2=rule["build_log2"], probe_log2=rule["probe_log2"])])
            got, stats = fij.JoinBatch(dtype, len(dev)).execute(dev)
            assert got == want, cfg
            assert stats["small_joins"] == (len(dev) if cfg["small"] else 0), cfg
    finally:
        fij.tuning_set(dtype, builtin=True)
    assert fij.device_status() == 0


def sys_path_pkg():
    import os
    import sys

    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fast-interconnects_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
