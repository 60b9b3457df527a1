"""GPU parity of the allocation-free prefix sum (fij_prefix_sum_noalloc).

Its offsets must equal fij_prefix_sum's on the same input, bit for bit, and
the oracle's, on two calls in a row through the library's one static scratch
array, the second with a smaller input (fewer keys, a smaller grid and
fanout), so that stale counters or scan state from the first call would
show.  Sizes beyond the scratch are rejected with InvalidArgument.
"""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import sync, to_dev

pytestmark = pytest.mark.gpu

CONTIG = fij.HistogramAlgorithm.Contiguous
CHUNKED = fij.HistogramAlgorithm.Chunked
SSWWC = fij.RadixPartitionAlgorithm.SSWWCv2


def offsets_both_ways(keys, hist, bits, grid, block=1024, ignore=0):
    kt = fij.KeyType.I32 if keys.dtype == np.int32 else fij.KeyType.I64
    rb = fij.RadixBits(bits) if ignore == 0 else fij.RadixBits(ignore, bits)
    rpass = fij.RadixPass.First if ignore == 0 else fij.RadixPass.Second
    dk = to_dev(keys)
    out = []
    for noalloc in (False, True):
        part = fij.GpuRadixPartitioner(hist, SSWWC, rb, grid, block)
        offs = fij.PartitionOffsets(hist, grid, bits, kt)
        if noalloc:
            part.prefix_sum_noalloc(rpass, dk, offs)
        else:
            part.prefix_sum(rpass, dk, offs)
        sync()
        local = offs.host_local_offsets() if hist == CONTIG else None
        out.append((offs.host_offsets(), local, offs.len()))
    return out


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_noalloc_prefix_sum_two_calls(dtype):
    # first call: 2^22 keys, 9 bits over 1024 chunks (2^19 counters, 65 scan
    # tiles); second: 100,003 keys, 5 bits over 96 chunks
    cases = [(1 << 22, 9, 1024, 0, 11), (100_003, 5, 96, 3, 12)]
    for n, bits, grid, ignore, seed in cases:
        print(f"noalloc prefix sum: {np.dtype(dtype).name} n={n} bits={bits} grid={grid} ignore={ignore}",
              flush=True)
        keys = orc.gen_uniform(n, seed, 3, 0, 1 << 30, dtype)
        (ref_off, ref_local, ref_len), (got_off, got_local, got_len) = offsets_both_ways(
            keys, CONTIG, bits, grid, ignore=ignore)
        np.testing.assert_array_equal(got_off, ref_off)
        np.testing.assert_array_equal(got_local, ref_local)
        assert got_len == ref_len == n
        want_off, want_local = orc.contiguous_prefix_sum(keys, grid, bits, ignore)
        np.testing.assert_array_equal(got_off, want_off)
        np.testing.assert_array_equal(got_local, want_local)


def test_noalloc_prefix_sum_chunked():
    keys = orc.gen_uniform(300_000, 5, 3, 0, 1 << 30, np.int64)
    (ref_off, _, _), (got_off, _, _) = offsets_both_ways(keys, CHUNKED, 8, 64)
    np.testing.assert_array_equal(got_off, ref_off)
    np.testing.assert_array_equal(got_off, orc.chunked_prefix_sum(keys, 64, 8, 0))


def test_noalloc_prefix_sum_then_partition():
    """The offsets drive the scatter exactly as fij_prefix_sum's do."""
    n, bits, grid = 1 << 20, 8, 256
    keys = orc.gen_perm(n, 3, np.int64, base=0)
    pays = orc.gen_uniform(n, 3, 5, 0, 10000, np.int64)
    kt = fij.KeyType.I64
    part = fij.GpuRadixPartitioner(CONTIG, SSWWC, fij.RadixBits(bits), grid, 1024)
    offs = fij.PartitionOffsets(CONTIG, grid, bits, kt)
    rel = fij.PartitionedRelation(n, CONTIG, bits, grid, kt, fill_null=True)
    dk, dp = to_dev(keys), to_dev(pays)
    part.prefix_sum_noalloc(fij.RadixPass.First, dk, offs)
    part.partition(fij.RadixPass.First, dk, dp, offs, rel)
    sync()
    gk, gv = rel.host_relation()
    ref_off, _ = orc.contiguous_prefix_sum(keys, grid, bits, 0)
    np.testing.assert_array_equal(rel.host_offsets(), ref_off)
    live = gk != -1
    assert live.sum() == n
    np.testing.assert_array_equal(np.sort(gk[live]), np.sort(keys))


def test_noalloc_rejects_oversize():
    # 2^12 partitions x 512 chunks = 2^21 counters > the 2^20 of the scratch
    keys = to_dev(np.arange(1 << 20, dtype=np.int64))
    part = fij.GpuRadixPartitioner(CONTIG, SSWWC, fij.RadixBits(12), 512, 1024)
    offs = fij.PartitionOffsets(CONTIG, 512, 12, fij.KeyType.I64)
    with pytest.raises(fij.FijError) as e:
        part.prefix_sum_noalloc(fij.RadixPass.First, keys, offs)
    assert e.value.status == fij.Status.INVALID_ARGUMENT
