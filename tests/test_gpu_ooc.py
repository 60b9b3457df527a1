"""Out-of-core join (fij_ooc_join) and histogram mixing (fij_cpu_prefix_sum
feeding fij_partition).  The out-of-core join must give the in-memory
join's result whatever part of pass 1 is cached in device memory, however
the partitions are grouped, and whether pass 1 runs on the GPU (relations in
pinned host or device memory) or on host threads."""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import assert_same_partitions, relation_ranges, sync, to_dev

pytestmark = pytest.mark.gpu


def make(nr, ns, dtype, seed, dups=False):
    if dups:  # duplicate and null build keys
        rk = orc.gen_uniform(nr, seed, 1, -2, nr // 2, dtype)
    else:
        rk = orc.gen_perm(nr, seed, dtype, 1)
    rp = orc.gen_uniform(nr, seed, 2, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(ns, seed, 3, 1, nr + nr // 4, dtype)  # a fifth of the probe keys miss
    sp = orc.gen_iota(ns, 0, dtype)
    return rk, rp, sk, sp


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("cache_frac", [0.0, 0.4, 2.0])
@pytest.mark.parametrize("staging", [0, 1 << 16])
def test_gpu_pass1_from_pinned_memory(dtype, cache_frac, staging):
    h = make(1 << 18, 1 << 20, dtype, 5)
    want = orc.sort_merge_join(*h)
    cols = [fij.HostArray.from_numpy(a) for a in h]
    tb = 2 * np.dtype(dtype).itemsize
    cache = int(cache_frac * (len(h[0]) + len(h[2])) * tb)
    got, t = fij.ooc_join(dtype, *cols, 7, 5, cache_bytes=cache, staging_tuples=staging)
    assert got == want
    if cache_frac == 0.0:
        assert t["cached_build_tuples"] == 0 and t["cached_probe_tuples"] == 0
    if cache_frac == 2.0:  # everything fits: padded lengths cached
        pad = 128 // tb
        assert t["cached_build_tuples"] == len(h[0]) + 128 * pad
        assert t["cached_probe_tuples"] == len(h[2]) + 128 * pad
    if cache_frac == 0.4:
        assert 0 < t["cached_probe_tuples"] < len(h[2])
    if staging:
        assert t["groups"] > 8
    assert fij.device_status() == 0


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_gpu_pass1_high_fanout_spill(dtype):
    """Pass 1 at 11 bits (the tile-sorting scatter) writing part of the
    partitions to the device cache and the rest to pinned host memory."""
    h = make(1 << 18, 1 << 20, dtype, 6)
    want = orc.sort_merge_join(*h)
    cols = [fij.HostArray.from_numpy(a) for a in h]
    tb = 2 * np.dtype(dtype).itemsize
    cache = int(0.5 * (len(h[0]) + len(h[2])) * tb)
    got, t = fij.ooc_join(dtype, *cols, 11, 3, cache_bytes=cache, staging_tuples=1 << 17)
    assert got == want
    assert 0 < t["cached_probe_tuples"] < len(h[2])
    assert fij.device_status() == 0


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [(9, 6), (4, 3)])
def test_cpu_pass1(dtype, bits):
    h = make(200_000, 900_001, dtype, 8, dups=True)
    want = orc.sort_merge_join(*h)
    got, t = fij.ooc_join(dtype, *h, *bits, pass1_cpu=True, cpu_threads=4, staging_tuples=1 << 17)
    assert got == want
    assert t["cached_build_tuples"] == 0


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("slices", [1, 3, 8])
@pytest.mark.parametrize("cache_frac", [0.0, 0.5, 2.0])
def test_cpu_pass1_overlapped_slices(dtype, slices, cache_frac):
    """CPU pass 1 overlapped with the GPU (cpu_partitioned_radix_join.rs:
    457-547): the build relation cached in device memory up to cache_bytes
    while the probe relation is partitioned in slices, every slice joined
    while the next is partitioned.  A heavy probe key makes one slice's
    partition longer than the staging groups, so the group buffers grow."""
    h = make(150_000, 1_000_003, dtype, 13, dups=True)
    h[2][::7] = 12345  # a seventh of the probe keys
    want = orc.sort_merge_join(*h)
    tb = 2 * np.dtype(dtype).itemsize
    cache = int(cache_frac * len(h[0]) * tb)
    got, t = fij.ooc_join(dtype, *h, 7, 5, pass1_cpu=True, cpu_threads=4, staging_tuples=1 << 16,
                          cache_bytes=cache, probe_slices=slices)
    assert got == want
    if cache_frac == 2.0:
        assert t["cached_build_tuples"] == len(h[0]) + 128 * (128 // tb)
    elif cache_frac == 0.5:
        assert 0 < t["cached_build_tuples"] < len(h[0])
    else:
        assert t["cached_build_tuples"] == 0
    assert t["groups"] >= slices
    assert fij.device_status() == 0


def test_device_inputs_and_plain_plan():
    dtype = np.int64
    h = make(1 << 20, 1 << 22, dtype, 11)
    d = [to_dev(a) for a in h]
    got, _ = fij.ooc_join(dtype, *d, 9, 5, cache_bytes=24 << 20, staging_tuples=1 << 19)
    want, _ = fij.JoinPlan(dtype, 1 << 20, 1 << 22, 9, 5).execute(*d)
    assert got == want
    sync()


def test_pageable_memory_needs_cpu_pass1():
    h = make(1 << 12, 1 << 14, np.int32, 3)
    with pytest.raises(fij.FijError) as e:
        fij.ooc_join(np.int32, *h, 5, 3)  # numpy arrays: pageable host memory
    assert e.value.status == fij.Status.INVALID_ARGUMENT
    got, _ = fij.ooc_join(np.int32, *h, 5, 3, pass1_cpu=True)
    assert got == orc.sort_merge_join(*h)


def test_empty_relations():
    h = make(1 << 12, 1 << 14, np.int64, 3)
    e = np.zeros(0, np.int64)
    assert fij.ooc_join(np.int64, e, e, h[2], h[3], 5, 3, pass1_cpu=True)[0] == (0, 0, 0)
    assert fij.ooc_join(np.int64, h[0], h[1], e, e, 5, 3, pass1_cpu=True)[0] == (0, 0, 0)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist", [0, 1])
def test_cpu_histogram_with_gpu_partitioner(dtype, hist):
    """Offsets from the host prefix sum, copied into device offsets, drive
    the GPU scatter: both the host-driven and the device-driven relation
    equal the oracle's prefix sum and partition (offsets bit for bit, every
    (chunk, partition) the same tuple multiset, padding untouched)."""
    n, bits, grid = 1_000_003, 9, 40
    keys = orc.gen_uniform(n, 2, 3, 0, 1 << 30, dtype)
    pays = orc.gen_iota(n, 0, dtype)
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    dk, dp = to_dev(keys), to_dev(pays)
    algo = fij.RadixPartitionAlgorithm.SSWWCv2
    if hist == 0:
        ref_off = orc.chunked_prefix_sum(keys, grid, bits)
        chunk_off, chunks = ref_off, grid
    else:
        ref_off, ref_local = orc.contiguous_prefix_sum(keys, grid, bits)
        chunk_off, chunks = ref_local, 1
    for on_cpu in (False, True):
        part = fij.GpuRadixPartitioner(hist, algo, fij.RadixBits(bits), grid, 1024)
        offs = fij.PartitionOffsets(hist, grid, bits, kt)
        rel = fij.PartitionedRelation(n, hist, bits, grid, kt, fill_null=True)
        if on_cpu:
            host = fij.cpu_prefix_sum(hist, keys, bits, grid)
            if hist == 1:
                np.testing.assert_array_equal(host[0], ref_off)
                np.testing.assert_array_equal(host[1], ref_local)
                offs.offsets.copy_from_host(host[0])
                offs.local_offsets.copy_from_host(host[1])
            else:
                np.testing.assert_array_equal(host, ref_off)
                offs.offsets.copy_from_host(host)
            offs.s.data_len = n
        else:
            part.prefix_sum(fij.RadixPass.First, dk, offs)
        part.partition(fij.RadixPass.First, dk, dp, offs, rel)
        sync()
        np.testing.assert_array_equal(rel.host_offsets(), ref_off)
        gk, gv = rel.host_relation()
        ref = orc.partition(keys, pays, grid, bits, 0, chunk_off, rel.s.relation_len)
        ranges = relation_ranges(ref_off, chunks, bits, rel.padded_len(), orc.pad_len(dtype))
        assert_same_partitions(gk, gv, ref["key"], ref["value"], ranges)
