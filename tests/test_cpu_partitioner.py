"""Host partitioner (fij_cpu_prefix_sum, fij_cpu_radix_partition): the GPU
prefix sums' offsets computed on host threads, and pass 1 on host threads in
the GPU's contiguous layout.  Host code: these run without a GPU, against the
oracle's prefix sums and partition (pinned by tests/test_oracle.py)."""
import numpy as np
import pytest

import fij
import oracle_lib as orc

CHUNKED, CONTIG = 0, 1


def keys_of(n, dtype, seed=1, hi=1 << 25):
    return orc.gen_uniform(n, seed, 3, 0, hi, dtype)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits,ignore", [(1, 0), (6, 0), (9, 0), (9, 9), (12, 3)])
@pytest.mark.parametrize("chunks", [1, 7, 64])
def test_prefix_sums_equal_the_oracle(dtype, bits, ignore, chunks):
    keys = keys_of(200_003, dtype)
    got = fij.cpu_prefix_sum(CHUNKED, keys, bits, chunks, ignore_bits=ignore, threads=5)
    np.testing.assert_array_equal(got, orc.chunked_prefix_sum(keys, chunks, bits, ignore))
    off, local = fij.cpu_prefix_sum(CONTIG, keys, bits, chunks, ignore_bits=ignore, threads=3)
    ref_off, ref_local = orc.contiguous_prefix_sum(keys, chunks, bits, ignore)
    np.testing.assert_array_equal(off, ref_off)
    np.testing.assert_array_equal(local, ref_local)


def test_prefix_sum_negative_and_empty():
    keys = np.array([-1, -2, 5, -(2 ** 31), 2 ** 31 - 1] * 1000, np.int32)
    np.testing.assert_array_equal(fij.cpu_prefix_sum(CHUNKED, keys, 4, 3), orc.chunked_prefix_sum(keys, 3, 4, 0))
    empty = np.zeros(0, np.int64)
    np.testing.assert_array_equal(fij.cpu_prefix_sum(CHUNKED, empty, 3, 2), orc.chunked_prefix_sum(empty, 2, 3, 0))


def test_prefix_sum_rejects_empty_last_chunk():
    # chunk length 32 (128 bytes of i32 keys) x 62 full chunks > 1954 keys
    with pytest.raises(fij.FijError):
        fij.cpu_prefix_sum(CHUNKED, keys_of(1954, np.int32), 4, 63)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits,ignore", [(2, 0), (8, 0), (9, 4)])
def test_partition_single_thread_equals_oracle(dtype, bits, ignore):
    n = 123_457
    keys = keys_of(n, dtype, seed=4)
    pays = orc.gen_iota(n, 0, dtype)
    k, v, off = fij.cpu_radix_partition(keys, pays, bits, ignore_bits=ignore, threads=1)
    ref_off, ref_local = orc.contiguous_prefix_sum(keys, 1, bits, ignore)
    np.testing.assert_array_equal(off, ref_off)
    ref = orc.partition(keys, pays, 1, bits, ignore, ref_local, len(k))
    # one thread scatters in input order, as the oracle's stable partition
    F = 1 << bits
    pad = orc.pad_len(dtype)
    for p in range(F):
        b = int(off[p])
        e = int(off[p + 1]) - pad if p + 1 < F else len(k)
        np.testing.assert_array_equal(k[b:e], ref["key"][b:e])
        np.testing.assert_array_equal(v[b:e], ref["value"][b:e])


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_partition_threads_keep_each_partitions_multiset(threads):
    dtype = np.int64
    n = 300_001
    keys = keys_of(n, dtype, seed=9, hi=1 << 40)
    keys[::97] = -1  # null keys are tuples like any other here
    pays = orc.gen_iota(n, 0, dtype)
    bits = 7
    k, v, off = fij.cpu_radix_partition(keys, pays, bits, threads=threads)
    k1, v1, off1 = fij.cpu_radix_partition(keys, pays, bits, threads=1)
    np.testing.assert_array_equal(off, off1)
    F = 1 << bits
    pad = orc.pad_len(dtype)
    for p in range(F):
        b = int(off[p])
        e = int(off[p + 1]) - pad if p + 1 < F else len(k)
        assert np.all((k[b:e].astype(np.uint64) & np.uint64(F - 1)) == p)
        a = np.sort(k[b:e] * 1_000_003 + v[b:e])
        c = np.sort(k1[b:e] * 1_000_003 + v1[b:e])
        np.testing.assert_array_equal(a, c)
    # every tuple exactly once
    ranges = [(int(off[p]), int(off[p + 1]) - pad if p + 1 < F else len(k)) for p in range(F)]
    got = np.sort(np.concatenate([v[b:e] for b, e in ranges]))
    np.testing.assert_array_equal(got, np.arange(n, dtype=dtype))


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("threads", [1, 5])
def test_partition_line_aligned_and_misaligned_output(dtype, threads):
    """Every whole 64-byte line of a thread's range is streamed
    (non-temporal stores) and only the lines it shares with another thread or
    with padding are copied, whether the output starts on a line or one tuple
    into it; an output not aligned to its tuple size (half a tuple off) copies
    every line.  All give the same relation."""
    n = 200_003
    keys = keys_of(n, dtype, seed=12, hi=1 << 30)
    pays = orc.gen_iota(n, 0, dtype)
    bits = 6
    rows = n + (1 << bits) * orc.pad_len(dtype)
    tb = 2 * np.dtype(dtype).itemsize
    got = []
    for shift in (0, tb, tb // 2):
        raw = np.zeros(rows * tb + 128, np.uint8)
        base = (-raw.ctypes.data) % 64 + shift
        out = raw[base: base + rows * tb].view(dtype).reshape(rows, 2)
        assert (out.ctypes.data % 64 == 0) == (shift == 0)
        k, v, off = fij.cpu_radix_partition(keys, pays, bits, threads=threads, out=out)
        got.append((k.copy(), v.copy(), off))
    for other in got[1:]:
        for a, b in zip(got[0], other):
            np.testing.assert_array_equal(a, b)
    ref_k, ref_v, ref_off = fij.cpu_radix_partition(keys, pays, bits, threads=threads)
    np.testing.assert_array_equal(got[0][2], ref_off)
    np.testing.assert_array_equal(got[0][0], ref_k)
    np.testing.assert_array_equal(got[0][1], ref_v)
    if threads == 1:  # one thread scatters in input order, as the oracle (its padding holds null keys)
        ref = orc.partition(keys, pays, 1, bits, 0, orc.contiguous_prefix_sum(keys, 1, bits, 0)[1], rows)
        off, pad, F = got[0][2], orc.pad_len(dtype), 1 << bits
        for p in range(F):
            b, e = int(off[p]), (int(off[p + 1]) - pad if p + 1 < F else rows)
            np.testing.assert_array_equal(got[0][0][b:e], ref["key"][b:e])
            np.testing.assert_array_equal(got[0][1][b:e], ref["value"][b:e])
