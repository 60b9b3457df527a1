"""Helpers shared by the GPU parity tests: inputs, uploads and the
partition validators of the reference's test suite
(sql-ops/tests/radix_partition/mod.rs:33-218) restated over numpy."""
import numpy as np

import fij
import oracle_lib as orc


def to_dev(a):
    return fij.DeviceArray.from_host(np.ascontiguousarray(a))


def sync():
    fij.synchronize()


def relation_ranges(host_offsets, chunks, radix_bits, padded_len, pad):
    """[(begin, end)] for every (chunk, partition), in offsets order."""
    F = 1 << radix_bits
    n = chunks * F
    out = []
    for ofi in range(n):
        b = int(host_offsets[ofi])
        e = int(host_offsets[ofi + 1]) - pad if ofi + 1 < n else padded_len
        out.append((b, e))
    return out


def range_ids(ranges, total_len):
    """Per-position index of the range that holds it, -1 for padding."""
    ids = np.full(total_len, -1, np.int64)
    for i, (b, e) in enumerate(ranges):
        if e > b:
            ids[b:e] = i
    return ids


def assert_same_partitions(gpu_keys, gpu_vals, ref_keys, ref_vals, ranges, null_key=-1):
    """Every (chunk, partition) range holds the same multiset of tuples in
    both relations, and padding outside the ranges is untouched (null key,
    zero payload)."""
    total = len(gpu_keys)
    ids = range_ids(ranges, total)
    inside = ids >= 0
    pad_k = gpu_keys[~inside]
    pad_v = gpu_vals[~inside]
    assert np.all(pad_k == null_key), "padding was written (keys)"
    assert np.all(pad_v == 0), "padding was written (payloads)"
    gi = ids[inside]
    o1 = np.lexsort((gpu_vals[inside], gpu_keys[inside], gi))
    o2 = np.lexsort((ref_vals[inside], ref_keys[inside], gi))
    np.testing.assert_array_equal(gpu_keys[inside][o1], ref_keys[inside][o2])
    np.testing.assert_array_equal(gpu_vals[inside][o1], ref_vals[inside][o2])


def assert_membership(keys, ranges, radix_bits, ignore_bits, F):
    """verify_partitions (mod.rs:90-114): each tuple sits in its partition."""
    for i, (b, e) in enumerate(ranges):
        if e > b:
            p = i % F
            d = orc.key_to_partition(keys[b:e], radix_bits, ignore_bits)
            assert np.all(d == p), f"tuple in range {i} has a foreign digit"
