"""C ABI host-side behaviour, no GPU needed.

* libfij.so loads and exports every function include/fij.h declares;
* layout helpers follow partitioned_relation.rs / partition_input_chunk.rs;
* argument validation returns the reference's error kinds
  (sql-ops/src/error.rs:17-29) before anything touches a device, as the
  reference's Rust checks do (gpu_radix_partition.rs:571-611, :1032-1080,
  cuda_radix_join.rs:183-199).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fij
from fij import FijError, Status

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAKE = 1 << 40  # never dereferenced: every call below fails validation first


def declared_functions():
    text = open(os.path.join(REPO, "include", "fij.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fij_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(fij.lib, n), f"libfij.so does not export {n}"


def test_version():
    assert "gfx950" in fij.version()


def test_padding_len():
    # partitioned_relation.rs:26-29: PADDING_BYTES / sizeof(Tuple)
    assert fij.padding_len(fij.KeyType.I32) == 16
    assert fij.padding_len(fij.KeyType.I64) == 8


@pytest.mark.parametrize("kt,n,chunks,expect", [
    (fij.KeyType.I32, 1000, 3, 320),          # ceil(1000/3)=334, & ~31
    (fij.KeyType.I64, 10 ** 6, 10, 100000),   # multiple of 16 already
    (fij.KeyType.I64, 100, 4, 16),            # 25 & ~15
    (fij.KeyType.I32, 100, 4, 0),             # 25 & ~31
    (fij.KeyType.I64, 0, 7, 0),
])
def test_input_chunk_size(kt, n, chunks, expect):
    assert fij.input_chunk_size(kt, n, chunks) == expect


def test_input_chunk_size_overflow():
    with pytest.raises(FijError) as e:
        fij.input_chunk_size(fij.KeyType.I64, 1 << 33, 1)
    assert e.value.status == Status.INTEGER_OVERFLOW


def test_layouts():
    chunks, olen, llen = C.c_uint32(), C.c_uint64(), C.c_uint64()
    fij.check(fij.lib.fij_partition_offsets_layout(0, 10, 9, C.byref(chunks), C.byref(olen), C.byref(llen)))
    assert (chunks.value, olen.value, llen.value) == (10, 5120, 0)
    fij.check(fij.lib.fij_partition_offsets_layout(1, 10, 9, C.byref(chunks), C.byref(olen), C.byref(llen)))
    assert (chunks.value, olen.value, llen.value) == (1, 512, 5120)
    rlen = C.c_uint64()
    # PartitionedRelation::new: len + fanout * chunks * pad
    fij.check(fij.lib.fij_partitioned_relation_layout(1, 1000, 0, 9, 10, C.byref(chunks), C.byref(rlen), C.byref(olen)))
    assert rlen.value == 1000 + 512 * 10 * 8 and olen.value == 5120
    fij.check(fij.lib.fij_partitioned_relation_layout(0, 1000, 1, 9, 10, C.byref(chunks), C.byref(rlen), C.byref(olen)))
    assert rlen.value == 1000 + 512 * 16 and olen.value == 512


def host_relation(n, bits, chunks, kt=fij.KeyType.I64):
    s = fij._PartitionedRelation()
    s.relation = FAKE
    s.offsets = FAKE
    s.relation_len = n + (1 << bits) * chunks * fij.padding_len(kt)
    s.len = n
    s.chunks = chunks
    s.radix_bits = bits
    s.type = int(kt)
    return s


def test_resize_and_capacity():
    s = host_relation(1000, 4, 2)
    assert fij.lib.fij_partitioned_relation_capacity(C.byref(s)) == 1000
    fij.check(fij.lib.fij_partitioned_relation_resize(C.byref(s), 600))
    assert s.len == 600
    assert fij.lib.fij_partitioned_relation_padded_len(C.byref(s)) == 600 + 16 * 2 * 8
    assert fij.lib.fij_partitioned_relation_resize(C.byref(s), 1001) == Status.INVALID_ARGUMENT
    assert s.len == 600  # an aborted resize leaves the relation intact


def test_partition_len_host():
    # two chunks, 4 partitions, i64 (pad 8): hand-built chunked offsets
    s = host_relation(20, 2, 2)
    offs = np.array([8, 19, 30, 41,  # chunk 0: sizes 3,3,3,? (end = next chunk start - pad)
                     60, 70, 80, 90], np.uint64)
    padded = fij.lib.fij_partitioned_relation_padded_len(C.byref(s))
    assert padded == 20 + 4 * 2 * 8
    out = C.c_uint64()
    lens = []
    for p in range(4):
        fij.check(fij.lib.fij_partitioned_relation_partition_len(C.byref(s), offs.ctypes.data_as(C.c_void_p), p,
                                                                 C.byref(out)))
        lens.append(out.value)
    # partition p: (offs[p+1]-8 - offs[p]) + (offs[4+p+1]-8 - offs[4+p]), last ends at padded_len
    assert lens == [(19 - 8 - 8) + (70 - 8 - 60), (30 - 8 - 19) + (80 - 8 - 70), (41 - 8 - 30) + (90 - 8 - 80),
                    (60 - 8 - 41) + (padded - 90)]
    assert fij.lib.fij_partitioned_relation_partition_len(C.byref(s), offs.ctypes.data_as(C.c_void_p), 4,
                                                          C.byref(out)) == Status.INVALID_ARGUMENT


def partitioner(hist=0, algo=0, bits=(9, None, None), grid=8, block=256):
    rb = fij.RadixBits(*bits)
    return fij.GpuRadixPartitioner(hist, algo, rb, grid, block)


def offsets_struct(chunks, max_chunks, bits, kt=fij.KeyType.I64, data_len=fij.LEN_UNSET, local=False):
    s = fij._PartitionOffsets()
    s.offsets = FAKE
    s.local_offsets = FAKE if local else None
    s.data_len = data_len
    s.chunks = chunks
    s.max_chunks = max_chunks
    s.radix_bits = bits
    s.type = int(kt)
    return s


def call_prefix_sum(p, pass_, offs, n=1000):
    return fij.lib.fij_prefix_sum(p.h, pass_, 1, C.c_void_p(FAKE), n, C.byref(offs), None)


def test_create_rejects_bad_block():
    with pytest.raises(FijError) as e:
        partitioner(block=100)
    assert e.value.status == Status.INVALID_ARGUMENT


def test_create_clamps_2048_block():
    # radix-join's default block is warp_size * 32 = 2048 on 64-wide waves
    p = partitioner(block=2048)
    assert p.h


def test_prefix_sum_pass_not_specified():
    p = partitioner(bits=(9, None, None))
    assert call_prefix_sum(p, 1, offsets_struct(8, 8, 9)) == Status.INVALID_ARGUMENT
    assert "not specified" in fij.lib.fij_last_error().decode()


def test_prefix_sum_mismatching_bits():
    p = partitioner()
    assert call_prefix_sum(p, 0, offsets_struct(8, 8, 8)) == Status.INVALID_ARGUMENT
    assert "radix bits" in fij.lib.fij_last_error().decode()


def test_prefix_sum_mismatching_chunks():
    assert call_prefix_sum(partitioner(), 0, offsets_struct(7, 7, 9)) == Status.INVALID_ARGUMENT
    contig = partitioner(hist=1)
    assert call_prefix_sum(contig, 0, offsets_struct(8, 8, 9, local=True)) == Status.INVALID_ARGUMENT
    assert call_prefix_sum(contig, 0, offsets_struct(1, 4, 9, local=True)) == Status.INVALID_ARGUMENT


def test_prefix_sum_integer_overflow():
    p = partitioner(grid=1)
    assert call_prefix_sum(p, 0, offsets_struct(1, 1, 9), n=1 << 33) == Status.INTEGER_OVERFLOW


def test_prefix_sum_grid_too_large_for_relation():
    # (grid - 1) * chunk_len > len would make the last chunk negative
    p = partitioner(grid=64)
    assert call_prefix_sum(p, 0, offsets_struct(64, 64, 9), n=961) == Status.INVALID_ARGUMENT


def test_too_many_radix_bits_unsupported():
    p = partitioner(bits=(16, None, None))
    assert call_prefix_sum(p, 0, offsets_struct(8, 8, 16)) == Status.UNSUPPORTED


def test_partition_without_prefix_sum():
    p = partitioner()
    offs = offsets_struct(8, 8, 9)
    rel = host_relation(1000, 9, 8)
    rc = fij.lib.fij_partition(p.h, 0, 1, C.c_void_p(FAKE), C.c_void_p(FAKE), 1000, C.byref(offs), C.byref(rel), None)
    assert rc == Status.INVALID_ARGUMENT
    assert "no length" in fij.lib.fij_last_error().decode()


def test_partition_mismatching_lengths_and_chunks():
    p = partitioner()
    offs = offsets_struct(8, 8, 9, data_len=1000)
    rel = host_relation(999, 9, 8)
    args = (p.h, 0, 1, C.c_void_p(FAKE), C.c_void_p(FAKE), 1000, C.byref(offs), C.byref(rel), None)
    assert fij.lib.fij_partition(*args) == Status.INVALID_ARGUMENT
    rel = host_relation(1000, 9, 4)
    args = (p.h, 0, 1, C.c_void_p(FAKE), C.c_void_p(FAKE), 1000, C.byref(offs), C.byref(rel), None)
    assert fij.lib.fij_partition(*args) == Status.INVALID_ARGUMENT
    assert "chunks" in fij.lib.fij_last_error().decode()
    rel = host_relation(1000, 8, 8)
    args = (p.h, 0, 1, C.c_void_p(FAKE), C.c_void_p(FAKE), 1000, C.byref(offs), C.byref(rel), None)
    assert fij.lib.fij_partition(*args) == Status.INVALID_ARGUMENT


def test_transform_requires_contiguous():
    p = partitioner(hist=0)
    offs = offsets_struct(1, 8, 9, local=True)
    src = host_relation(1000, 4, 8)
    rc = fij.lib.fij_prefix_sum_and_transform(p.h, 0, 1, 0, C.byref(src), C.c_void_p(FAKE), C.c_void_p(FAKE), 10,
                                              C.byref(offs), None)
    assert rc == Status.INVALID_ARGUMENT
    assert "Contiguous" in fij.lib.fij_last_error().decode()


def test_transform_invalid_partition_id():
    p = partitioner(hist=1, bits=(4, 9, None))
    offs = offsets_struct(1, 8, 9, local=True)
    src = host_relation(1000, 4, 8)
    rc = fij.lib.fij_prefix_sum_and_transform(p.h, 1, 1, 16, C.byref(src), C.c_void_p(FAKE), C.c_void_p(FAKE), 10,
                                              C.byref(offs), None)
    assert rc == Status.INVALID_ARGUMENT
    assert "partition ID" in fij.lib.fij_last_error().decode()


def test_copy_with_payload_requires_contiguous():
    p = partitioner(hist=0)
    offs = offsets_struct(8, 8, 9)
    rc = fij.lib.fij_prefix_sum_and_copy_with_payload(p.h, 0, 1, C.c_void_p(FAKE), C.c_void_p(FAKE), 100,
                                                      C.c_void_p(FAKE), C.c_void_p(FAKE), 100, C.byref(offs), None)
    assert rc == Status.INVALID_ARGUMENT


def join(scheme=0, grid=4, block=128, bits=(6, None, None), pass_=0):
    return fij.HipRadixJoin(pass_, fij.RadixBits(*bits), scheme, grid, block)


def call_join(j, build, probe, ta_len, result_len=4 * 128):
    return fij.lib.fij_radix_join_join(j.h, 1, C.byref(build), C.byref(probe), C.c_void_p(FAKE), result_len,
                                       C.c_void_p(FAKE), ta_len, None, None)


def test_join_rejects_chunked_relations():
    j = join()
    assert call_join(j, host_relation(100, 6, 2), host_relation(100, 6, 1), 5) == Status.INVALID_ARGUMENT
    assert "Chunked build" in fij.lib.fij_last_error().decode()
    assert call_join(j, host_relation(100, 6, 1), host_relation(100, 6, 2), 5) == Status.INVALID_ARGUMENT
    assert "Chunked probe" in fij.lib.fij_last_error().decode()


def test_join_task_assignment_length():
    j = join()
    assert call_join(j, host_relation(100, 6, 1), host_relation(100, 6, 1), 4) == Status.INVALID_ARGUMENT
    assert "grid size + 1" in fij.lib.fij_last_error().decode()


def test_join_linear_probing_unsupported():
    j = join(scheme=1)
    assert call_join(j, host_relation(100, 6, 1), host_relation(100, 6, 1), 5) == Status.UNSUPPORTED


def test_join_result_set_too_small():
    j = join()
    assert call_join(j, host_relation(100, 6, 1), host_relation(100, 6, 1), 5, result_len=10) == \
        Status.INVALID_ARGUMENT


def test_join_mismatching_bits():
    j = join()
    assert call_join(j, host_relation(100, 5, 1), host_relation(100, 5, 1), 5) == Status.INVALID_ARGUMENT


def test_np_build_table_too_small():
    rc = fij.lib.fij_np_build(1, 1, 0, C.c_void_p(FAKE), 16, C.c_void_p(FAKE), C.c_void_p(FAKE), 17, 1, 256, None)
    assert rc == Status.INVALID_ARGUMENT
    assert "too small" in fij.lib.fij_last_error().decode()


# ---- host-side rules and argument checks of the batch, server and
# out-of-core entry points (no device work)

@pytest.fixture
def rule_only():
    """The closed-form size-class rule, without the built-in tuning table."""
    fij.tuning_set(np.int32, [])
    fij.tuning_set(np.int64, [])
    yield
    fij.tuning_set(np.int32, builtin=True)
    fij.tuning_set(np.int64, builtin=True)


def test_autotune_rule(rule_only):
    assert fij.autotune(np.int32, 1000, 16000)[0]          # one block's LDS table
    assert not fij.autotune(np.int32, 12289, 16000)[0]      # build side beyond one table fill
    assert not fij.autotune(np.int32, 1000, (1 << 22) + 1)[0]
    assert fij.autotune(np.int64, 6144, 1 << 20)[0] and not fij.autotune(np.int64, 6145, 1 << 20)[0]
    small, b1, b2 = fij.autotune(np.int64, 1 << 27, 1 << 27)
    assert not small and (b1, b2) == (9, 9)
    # build sides from 2^18 tuples up to the single-pass limit: one pass of 9 bits
    for dtype, top in ((np.int32, 1 << 21), (np.int64, 1 << 20)):
        for n_r in (1 << 18, 300000, top):
            assert fij.autotune(dtype, n_r, 8 * n_r) == (False, 9, 0)
        assert fij.autotune(dtype, top + 1, 8 * top)[2] > 0
    for n_r in (1 << 14, (1 << 18) - 1, 5_000_000, 1 << 23, 1 << 26):
        for f in (1, 4, 16):
            small, b1, b2 = fij.autotune(np.int32, n_r, n_r * f)
            assert 1 <= b1 <= 9 and 1 <= b2 <= 12
            # at most 512 build and 512 probe tuples per final partition on average
            assert n_r >> (b1 + b2) <= 512 or (b1, b2) == (9, 12)
            assert (n_r * f) >> (b1 + b2) <= 512 or (b1, b2) == (9, 12)
            # a pass-1 partition's probe side spans at most 64 pass-2 tiles of 8192
            assert (n_r * f) >> b1 <= 64 * 8192 or b1 == 9


def test_builtin_tuning_table():
    """Every built-in entry is a valid configuration and is what batches of
    its class run with; classes outside the table get the rule."""
    import re

    text = open(os.path.join(REPO, "fast-interconnects_amd", "csrc", "tuning_table.inc")).read()
    for name, dtype in (("kTunedI32", np.int32), ("kTunedI64", np.int64)):
        body = text[text.index(name):]
        body = body[:body.index("};")]
        rows = [tuple(int(x) for x in m.groups()) for m in
                re.finditer(r"\{(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), 0\}", body)]
        assert rows[-1] == (0,) * 7  # terminator
        for bl, pl, small, b1, b2, ring, block in rows[:-1]:
            assert 1 <= b1 <= 9 and 1 <= b2 <= 12 and block in (512, 1024) and ring in (0, 1, 2, 4, 8, 16, 32, 64, 128, 256)
            n_r, n_s = (3 << bl) // 4, (3 << pl) // 4  # lengths inside the class
            t = fij.tuning_lookup(dtype, n_r, n_s)
            if not small and 1 << 18 <= n_r <= (1 << 21 if dtype == np.int32 else 1 << 20):
                # single-pass classes take precedence over the measured two-pass entries
                assert (t["radix_bits_1"], t["radix_bits_2"], t["from_table"]) == (9, 0, 0)
                continue
            assert t["from_table"] == 1 and (t["build_log2"], t["probe_log2"]) == (bl, pl)
            assert (t["radix_bits_1"], t["radix_bits_2"], t["ring_lines"], t["block"]) == (b1, b2, ring, block)
            assert t["small"] == small
    t = fij.tuning_lookup(np.int64, 1 << 35, 1 << 36)
    assert t["from_table"] == 0 and (t["radix_bits_1"], t["radix_bits_2"]) == (9, 12)


def test_tuning_set_validates_entries():
    bad = [dict(build_log2=20, probe_log2=23, radix_bits_1=10, radix_bits_2=5, block=1024),
           dict(build_log2=20, probe_log2=23, radix_bits_1=5, radix_bits_2=13, block=1024),
           dict(build_log2=20, probe_log2=23, radix_bits_1=5, radix_bits_2=5, block=256),
           dict(build_log2=20, probe_log2=23, radix_bits_1=5, radix_bits_2=5, block=512, ring_lines=3)]
    try:
        for e in bad:
            with pytest.raises(fij.FijError) as ex:
                fij.tuning_set(np.int32, [e])
            assert ex.value.status == fij.Status.INVALID_ARGUMENT
        fij.tuning_set(np.int32, [dict(build_log2=20, probe_log2=23, small=1, radix_bits_1=6, radix_bits_2=5,
                                       ring_lines=8, block=512)])
        t = fij.tuning_lookup(np.int32, 1 << 20, 1 << 23)
        assert (t["radix_bits_1"], t["radix_bits_2"], t["ring_lines"], t["block"], t["from_table"]) == (6, 5, 8, 512, 1)
        assert t["small"] == 0  # a 2^20 build side does not qualify for the small path
        assert fij.autotune(np.int32, 1 << 20, 1 << 23) == (False, 6, 5)
    finally:
        fij.tuning_set(np.int32, builtin=True)


def test_server_config_rejected_on_host():
    with pytest.raises(fij.FijError) as e:
        fij.JoinServer(np.int64, [(0, 1 << 20)], max_probe_len=1 << 20)  # no slot at all
    assert e.value.status == fij.Status.INVALID_ARGUMENT
    with pytest.raises(ValueError):
        fij.JoinServer(np.int64, [(1, 1 << 10)] * 5, max_probe_len=1 << 20)  # more classes than the ABI holds


def test_ooc_arguments_rejected_on_host():
    k = np.arange(1, 1001, dtype=np.int64)
    for bits in ((0, 4), (4, 0), (13, 4), (4, 13)):
        with pytest.raises(fij.FijError) as e:
            fij.ooc_join(np.int64, k, k, k, k, *bits, pass1_cpu=True)
        assert e.value.status == fij.Status.INVALID_ARGUMENT
    raw = np.zeros(8 * 1000 + 8, np.uint8)
    misaligned = np.frombuffer(raw.data, np.int64, 999, offset=4)
    with pytest.raises(fij.FijError):
        fij.ooc_join(np.int64, misaligned, k[:999], k, k, 4, 4, pass1_cpu=True)


def test_ooc_columns_checked_before_the_call():
    """Payload columns shorter than their keys, or columns of another type,
    would make the C call read past their ends: rejected in Python."""
    k = np.arange(1, 1001, dtype=np.int64)
    with pytest.raises(ValueError):
        fij.ooc_join(np.int64, k, k[:999], k, k, 4, 4, pass1_cpu=True)
    with pytest.raises(ValueError):
        fij.ooc_join(np.int64, k, k, k, k[:10], 4, 4, pass1_cpu=True)
    with pytest.raises(ValueError):
        fij.ooc_join(np.int64, k.astype(np.int32), k, k, k, 4, 4, pass1_cpu=True)
    with pytest.raises(ValueError):
        fij.ooc_join(np.int32, k, k, k, k, 4, 4, pass1_cpu=True)


@pytest.mark.parametrize("bytes_, parts, chunks", [
    (16 * (1 << 20) + 8, 512, 16),   # not whole words
    (16 * (1 << 20), 512, 0),        # no chunks
    (16 * (1 << 20), 512, 3),        # chunks do not divide the words
    (16 * 4096 * 2, 512, 2),         # chunks of 4096 words: not whole groups of 8192
])
def test_calib_line_scatter_rejects_shapes(bytes_, parts, chunks):
    st = fij.lib.fij_calib_line_scatter(C.c_void_p(FAKE), C.c_void_p(FAKE), bytes_, parts, chunks, None)
    assert st == Status.INVALID_ARGUMENT


def test_calib_line_scatter_rejects_uneven_pieces():
    # 8192-word chunks hold 1024 lines: 3000 partitions do not divide them
    st = fij.lib.fij_calib_line_scatter(C.c_void_p(FAKE), C.c_void_p(FAKE), 16 * 8192 * 4, 3000, 4, None)
    assert st == Status.INVALID_ARGUMENT


@pytest.mark.parametrize("part, f2, run, shift", [
    (262144, 512, 16, 17),   # shift past the run: the last words would leave the tile
    (262144, 512, 24, 1),    # runs do not divide an item
    (262144, 128, 16, 1),    # items of 2048 words
    (262143, 512, 16, 1),    # partitions do not divide the words
    (262144, 0, 16, 1),
])
def test_calib_run_gather_rejects_shapes(part, f2, run, shift):
    st = fij.lib.fij_calib_run_gather(C.c_void_p(FAKE), 16 * (1 << 20), part, f2, run, shift, C.c_void_p(FAKE), 512,
                                      None)
    assert st == Status.INVALID_ARGUMENT


@pytest.mark.parametrize("dtype, one_pass_max", [(np.int64, 1 << 21), (np.int32, 1 << 22)])
def test_plan_bits_one_pass_up_to_block_tables(dtype, one_pass_max):
    """One pass of 9 bits while |R| / 512 stays within 2/3 of a 128 KiB block
    table (6,144 int64 or 12,288 int32 build tuples); beyond it two passes:
    9 + b2 bits with items of about a third of a 64 KiB table's build tuples
    for probe sides of 16 |R| or more, else the size-class rule."""
    assert fij.plan_bits(dtype, one_pass_max, 1 << 28) == (9, 0)
    assert fij.plan_bits(dtype, 1000, 1 << 14) == (9, 0)
    per_item = 1024 if dtype == np.int64 else 2048
    for n_r in (one_pass_max + 1, 4 * one_pass_max, 1 << 26):
        b1, b2 = fij.plan_bits(dtype, n_r, 16 * n_r)
        assert b1 == 9 and (n_r >> 9) / (1 << b2) <= per_item < (n_r >> 9) / (1 << (b2 - 1))
        _, r1, r2 = fij.autotune(dtype, n_r, 8 * n_r)
        assert fij.plan_bits(dtype, n_r, 8 * n_r) == (r1, r2) and r2 > 0
    with pytest.raises(FijError):
        fij.check(fij.lib.fij_join_plan_bits(7, 10, 10, C.byref(C.c_uint32()), C.byref(C.c_uint32())))
