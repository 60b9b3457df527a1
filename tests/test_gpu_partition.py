"""GPU parity: prefix sums and scatters against the oracle.

Mirrors sql-ops/tests/test_gpu_radix_partition.rs (run_gpu_partitioning
:54-174, two-pass :285-456, copy-with-payload :176-283): offsets must be
bit-identical to the oracle's; every (chunk, partition) must hold the same
tuple multiset as the oracle's partition of the same input; padding stays
untouched.
"""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import sync, assert_membership, assert_same_partitions, relation_ranges, to_dev

pytestmark = pytest.mark.gpu

CHUNKED = fij.HistogramAlgorithm.Chunked
CONTIG = fij.HistogramAlgorithm.Contiguous


def make_input(n, dtype, seed, key_kind="pk"):
    if key_kind == "pk":
        keys = orc.gen_perm(n, seed, dtype, base=0)
    else:
        keys = orc.gen_uniform(n, seed, 3, 0, 32 << 20, dtype)
    pays = orc.gen_uniform(n, seed, 5, 0, 10000, dtype)
    return keys, pays


def unaligned_dev(a, shift):
    """A device view of `a` that starts `shift` elements into its buffer
    (aligned to the element size only)."""
    if shift == 0:
        return to_dev(a)
    buf = to_dev(np.concatenate([np.zeros(shift, a.dtype), a]))
    return buf.view(shift, shift + len(a))


def run_partition(n, dtype, hist, algo, bits, grid, block=128, seed=1, key_kind="pk", ignore=0, shift=(0, 0)):
    keys, pays = make_input(n, dtype, seed, key_kind)
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    rb = fij.RadixBits(bits) if ignore == 0 else fij.RadixBits(ignore, bits)
    rpass = fij.RadixPass.First if ignore == 0 else fij.RadixPass.Second
    part = fij.GpuRadixPartitioner(hist, algo, rb, grid, block)
    offs = fij.PartitionOffsets(hist, grid, bits, kt)
    rel = fij.PartitionedRelation(n, hist, bits, grid, kt, fill_null=True)
    dk, dp = unaligned_dev(keys, shift[0]), unaligned_dev(pays, shift[1])
    part.prefix_sum(rpass, dk, offs)
    part.partition(rpass, dk, dp, offs, rel)
    sync()
    assert fij.device_status() == 0
    # offsets: bit-exact
    F = 1 << bits
    if hist == CHUNKED:
        ref_off = orc.chunked_prefix_sum(keys, grid, bits, ignore)
        chunk_off = ref_off
        chunks = grid
    else:
        ref_off, ref_local = orc.contiguous_prefix_sum(keys, grid, bits, ignore)
        chunk_off = ref_local
        chunks = 1
        np.testing.assert_array_equal(offs.host_local_offsets(), ref_local)
    np.testing.assert_array_equal(rel.host_offsets(), ref_off)
    assert offs.len() == n and rel.len() == n
    # tuples
    ref = orc.partition(keys, pays, grid, bits, ignore, chunk_off, rel.s.relation_len)
    gk, gv = rel.host_relation()
    ranges = relation_ranges(ref_off, chunks, bits, rel.padded_len(), orc.pad_len(dtype))
    assert_same_partitions(gk, gv, ref["key"], ref["value"], ranges)
    assert_membership(gk, ranges, bits, ignore, F)
    # partition_len agrees with the oracle restatement of partitioned_relation.rs:429-459
    h = rel.host_offsets()
    for pid in (0, F // 2, F - 1):
        assert rel.partition_len(pid, h) == orc.partition_len(ref_off, chunks, bits, rel.padded_len(),
                                                              orc.pad_len(dtype), pid)
    return keys, pays, rel


ALGOS = [fij.RadixPartitionAlgorithm.NC, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixPartitionAlgorithm.HSSWWCv4]


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("bits", [2, 10, 12])
def test_partition_reference_grid(dtype, hist, algo, bits):
    """test_gpu_radix_partition.rs: 32 MiB of keys, grid 10, block 128."""
    n = (32 << 20) // np.dtype(dtype).itemsize // 4
    run_partition(n, dtype, hist, algo, bits, grid=10, block=128, key_kind="attr")


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
@pytest.mark.parametrize("algo", [fij.RadixPartitionAlgorithm.NC, fij.RadixPartitionAlgorithm.SSWWCv2])
def test_partition_non_power_two(dtype, hist, algo):
    run_partition(10 ** 6, dtype, hist, algo, 10, grid=10, block=128)


@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
@pytest.mark.parametrize("algo", [fij.RadixPartitionAlgorithm.NC, fij.RadixPartitionAlgorithm.SSWWCv2])
def test_partition_small(hist, algo):
    """gpu_tuple_loss_or_duplicates_small_chunked_i32_2_bits: 100 tuples, grid 4."""
    run_partition(100, np.int32, hist, algo, 2, grid=4, block=128)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [0, 1, 8, 9])
def test_partition_wide_grid(dtype, bits):
    """MI355X-sized launch: 1024 threads per block, 512 chunks."""
    run_partition(3_000_017, dtype, CHUNKED, fij.RadixPartitionAlgorithm.SSWWCv2, bits, grid=512, block=1024)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("steps,extra", [(0, 1), (0, "v-1"), (1, 0), (1, "v+1"), (7, 3), (8, 0), (8, 1), (9, -1),
                                         (16, 5), (17, 0)])
def test_partition_tile_counts(dtype, steps, extra):
    """The buffered scatter's loaders take 512 x V tuples per step (V = 4
    for i32, 2 for i64: one 16-byte load per lane and column), eight steps
    per loop round, and the last len % V tuples apart: one chunk of every
    kind of remainder."""
    v = 16 // np.dtype(dtype).itemsize
    step = 512 * v
    extra = {"v-1": v - 1, "v+1": v + 1}.get(extra, extra)
    n = steps * step + (extra if extra >= 0 else step + extra)
    run_partition(n, dtype, CHUNKED, fij.RadixPartitionAlgorithm.SSWWCv2, 9, grid=1, block=1024)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
@pytest.mark.parametrize("shift", [(1, 3), (2, 0), (0, 1)])
def test_partition_unaligned_columns(dtype, hist, shift):
    """Columns aligned to their element size only (views starting 1-3
    elements into a buffer): the buffered scatter loads 16 bytes per lane
    from any element-aligned address."""
    run_partition(300_007, dtype, hist, fij.RadixPartitionAlgorithm.SSWWCv2, 9, grid=37, block=1024, shift=shift)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
@pytest.mark.parametrize("bits", [10, 12, 14])
@pytest.mark.parametrize("algo", [fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixPartitionAlgorithm.HSSWWCv4])
def test_partition_high_fanout_buffered(dtype, hist, bits, algo):
    """Fanouts whose line rings do not fit LDS: the buffered variants sort
    each tile of the chunk by digit in LDS and write runs (32 chunks)."""
    run_partition(4_000_037, dtype, hist, algo, bits, grid=32, block=1024)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("tiles,extra", [(0, 1), (0, 4095), (1, 0), (1, 1), (2, 7), (3, -1)])
def test_partition_high_fanout_tile_counts(dtype, tiles, extra):
    """One chunk of every kind of remainder of the sorting scatter's tiles
    (8192 tuples at i32, 4096 at i64), at 11 bits."""
    tile = 8192 if dtype == np.int32 else 4096
    n = tiles * tile + (extra if extra >= 0 else tile + extra)
    run_partition(n, dtype, CHUNKED, fij.RadixPartitionAlgorithm.SSWWCv2, 11, grid=1, block=1024)


@pytest.mark.parametrize("hist", [CHUNKED, CONTIG])
def test_partition_empty(hist):
    run_partition(0, np.int64, hist, fij.RadixPartitionAlgorithm.SSWWCv2, 4, grid=3, block=256)


def test_contiguous_many_scan_tiles():
    """F * G = 4096 * 64 counters: the look-back scan spans 32 tiles."""
    run_partition(1 << 21, np.int32, CONTIG, fij.RadixPartitionAlgorithm.NC, 12, grid=64, block=256)


def test_partition_skewed_single_key():
    """All keys equal: one partition receives everything (buffer overflow path)."""
    n = 1 << 20
    keys = np.full(n, 77, np.int64)
    pays = np.arange(n, dtype=np.int64)
    kt = fij.KeyType.I64
    for algo in (fij.RadixPartitionAlgorithm.NC, fij.RadixPartitionAlgorithm.SSWWCv2):
        part = fij.GpuRadixPartitioner(CHUNKED, algo, fij.RadixBits(9), 64, 1024)
        offs = fij.PartitionOffsets(CHUNKED, 64, 9, kt)
        rel = fij.PartitionedRelation(n, CHUNKED, 9, 64, kt, fill_null=True)
        dk, dp = to_dev(keys), to_dev(pays)
        part.prefix_sum(fij.RadixPass.First, dk, offs)
        part.partition(fij.RadixPass.First, dk, dp, offs, rel)
        sync()
        ref_off = orc.chunked_prefix_sum(keys, 64, 9)
        np.testing.assert_array_equal(rel.host_offsets(), ref_off)
        ref = orc.partition(keys, pays, 64, 9, 0, ref_off, rel.s.relation_len)
        gk, gv = rel.host_relation()
        ranges = relation_ranges(ref_off, 64, 9, rel.padded_len(), 8)
        assert_same_partitions(gk, gv, ref["key"], ref["value"], ranges)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("hist1", [CHUNKED, CONTIG])
@pytest.mark.parametrize("bits", [(2, 2), (10, 10), (0, 0)])
def test_two_pass(dtype, hist1, bits):
    """run_gpu_two_pass_partitioning: pass 1, then per partition transform +
    pass 2.  The transform must equal the oracle's concatenation of the GPU's
    own pass-1 pieces exactly (mod.rs:190-218)."""
    b1, b2 = bits
    n = 1 << 20
    grid, block = 10, 128
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    keys, pays = make_input(n, dtype, 11, "attr")
    rb = fij.RadixBits(b1, b2)
    p1 = fij.GpuRadixPartitioner(hist1, fij.RadixPartitionAlgorithm.SSWWCv2, rb, grid, block)
    p2 = fij.GpuRadixPartitioner(CONTIG, fij.RadixPartitionAlgorithm.SSWWCv2, rb, grid, block)
    offs = fij.PartitionOffsets(hist1, grid, b1, kt)
    rel1 = fij.PartitionedRelation(n, hist1, b1, grid, kt, fill_null=True)
    dk, dp = to_dev(keys), to_dev(pays)
    p1.prefix_sum(fij.RadixPass.First, dk, offs)
    sync()
    F1 = 1 << b1
    h0 = offs.host_offsets()
    max_len = max(offs.partition_len(p, h0) for p in range(F1))
    p1.partition(fij.RadixPass.First, dk, dp, offs, rel1)
    sync()
    h1 = rel1.host_offsets()
    src = np.zeros(rel1.s.relation_len, orc.tuple_dtype(dtype))
    gk1, gv1 = rel1.host_relation()
    src["key"], src["value"] = gk1, gv1
    cached_k = fij.DeviceArray(max(max_len, 1), dtype)
    cached_v = fij.DeviceArray(max(max_len, 1), dtype)
    offs2 = fij.PartitionOffsets(CONTIG, grid, b2, kt)
    rel2 = fij.PartitionedRelation(max_len, CONTIG, b2, grid, kt, fill_null=True)
    for pid in list(range(0, F1, max(1, F1 // 8))) + [F1 - 1]:
        plen = rel1.partition_len(pid, h1)
        rel2.resize(plen)
        rel2.fill_null()
        ck, cv = cached_k.view(0, plen), cached_v.view(0, plen)
        p2.prefix_sum_and_transform(fij.RadixPass.Second, pid, rel1, ck, cv, offs2)
        p2.partition(fij.RadixPass.Second, ck, cv, offs2, rel2)
        sync()
        assert fij.device_status() == 0
        tk, tv = orc.transform(src, h1, rel1.num_chunks(), b1, rel1.padded_len(), pid, dtype)
        np.testing.assert_array_equal(ck.to_host(), tk)
        np.testing.assert_array_equal(cv.to_host(), tv)
        off2, loc2 = orc.contiguous_prefix_sum(tk, grid, b2, b1)
        np.testing.assert_array_equal(rel2.host_offsets(), off2)
        ref2 = orc.partition(tk, tv, grid, b2, b1, loc2, rel2.s.relation_len)
        gk, gv = rel2.host_relation()
        ranges = relation_ranges(off2, 1, b2, rel2.padded_len(), orc.pad_len(dtype))
        assert_same_partitions(gk, gv, ref2["key"], ref2["value"], ranges)
        assert_membership(gk, ranges, b2, b1, 1 << b2)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_copy_with_payload(dtype):
    """run_gpu_partitioning_and_copy_with_payload (test_gpu_radix_partition.rs:176-283)."""
    n = 1_000_003
    grid, bits = 10, 10
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    keys, pays = make_input(n, dtype, 5, "attr")
    part = fij.GpuRadixPartitioner(CONTIG, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixBits(bits), grid, 256)
    offs = fij.PartitionOffsets(CONTIG, grid, bits, kt)
    rel = fij.PartitionedRelation(n, CONTIG, bits, grid, kt, fill_null=True)
    dk, dp = to_dev(keys), to_dev(pays)
    ck, cv = fij.DeviceArray(n, dtype), fij.DeviceArray(n, dtype)
    part.prefix_sum_and_copy_with_payload(fij.RadixPass.First, dk, dp, ck, cv, offs)
    part.partition(fij.RadixPass.First, ck, cv, offs, rel)
    sync()
    np.testing.assert_array_equal(ck.to_host(), keys)
    np.testing.assert_array_equal(cv.to_host(), pays)
    off, loc = orc.contiguous_prefix_sum(keys, grid, bits)
    np.testing.assert_array_equal(rel.host_offsets(), off)
    ref = orc.partition(keys, pays, grid, bits, 0, loc, rel.s.relation_len)
    gk, gv = rel.host_relation()
    assert_same_partitions(gk, gv, ref["key"], ref["value"], relation_ranges(off, 1, bits, rel.padded_len(),
                                                                            orc.pad_len(dtype)))


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_generators_match_oracle(dtype):
    n = 1_234_567
    kt = fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64
    a = fij.DeviceArray(n, dtype)
    fij.gen_permutation(a, 42, 1)
    np.testing.assert_array_equal(a.to_host(), orc.gen_perm(n, 42, dtype, 1))
    fij.gen_uniform(a, 42, 1, 1, n)
    np.testing.assert_array_equal(a.to_host(), orc.gen_uniform(n, 42, 1, 1, n, dtype))
    fij.gen_iota(a, 7)
    np.testing.assert_array_equal(a.to_host(), orc.gen_iota(n, 7, dtype))
