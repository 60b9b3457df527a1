"""Joins of relations loaded from files (fij_relation_file_read), the way the
reference feeds its joins from the generator's files
(data-store/src/join_data.rs:141-280, datagen/src/bin/generate.rs:343-365).

The files are written by Python's own text and gzip writers in the
generator's layout (header `key value`, one record per line, payload = row +
1 on the build side) and by the library's writer (comma-delimited), read back
through the C ABI, and joined by the execution method (JoinPlan), by the
operator API (prefix sum + partition + bucket-chaining join) and by the
out-of-core join reading the loaded columns from pinned host memory.  Every
result equals the oracle's radix join and the sort-merge join of the
in-memory columns the files were written from.
"""
import gzip

import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import sync, to_dev

pytestmark = pytest.mark.gpu

# the radix join takes single-range (contiguous) relations, as the
# reference's (cuda_radix_join.rs:183-193)
CONTIG = fij.HistogramAlgorithm.Contiguous


def write_generator_file(path, keys, pays, gz=False):
    text = "key value\n" + "".join(f"{int(k)} {int(v)}\n" for k, v in zip(keys, pays))
    if gz:
        with gzip.open(path, "wt", newline="") as f:
            f.write(text)
    else:
        with open(path, "w", newline="") as f:
            f.write(text)


def relations(dtype, n_r, n_s, seed):
    rk = orc.gen_perm(n_r, seed, dtype, 1)
    rp = np.arange(1, n_r + 1, dtype=dtype)  # the generator writes value = row + 1
    sk = orc.gen_uniform(n_s, seed, 1, 1, n_r + n_r // 4, dtype)  # a fifth of the probes miss
    sp = orc.gen_uniform(n_s, seed, 2, -(1 << 20), 1 << 21, dtype)
    return rk, rp, sk, sp


def operator_join(rk, rp, sk, sp, bits):
    kt = fij.KeyType.I32 if rk.dtype == np.int32 else fij.KeyType.I64
    rels = []
    for k, p in ((rk, rp), (sk, sp)):
        part = fij.GpuRadixPartitioner(CONTIG, fij.RadixPartitionAlgorithm.SSWWCv2, fij.RadixBits(bits), 8, 256)
        offs = fij.PartitionOffsets(CONTIG, 8, bits, kt)
        rel = fij.PartitionedRelation(len(k), CONTIG, bits, 8, kt)
        dk, dp = to_dev(k), to_dev(p)
        part.prefix_sum(fij.RadixPass.First, dk, offs)
        part.partition(fij.RadixPass.First, dk, dp, offs, rel)
        rels.append(rel)
    grid, block = 64, 256
    agg = fij.aggregates_array(grid)
    tasks = fij.DeviceArray.zeros(grid + 1, np.uint32)
    j = fij.HipRadixJoin(fij.RadixPass.First, fij.RadixBits(bits), fij.HashingScheme.BucketChaining, grid, block)
    j.join(rels[0], rels[1], None, tasks, agg)
    sync()
    assert fij.device_status() == 0
    return fij.sum_aggregates(agg)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("fmt", ["text", "gzip", "library_csv"])
def test_join_loaded_relations(tmp_path, dtype, fmt):
    rk, rp, sk, sp = relations(dtype, 20_000, 150_000, 9)
    paths = [tmp_path / f"r.{fmt}", tmp_path / f"s.{fmt}"]
    delim = "," if fmt == "library_csv" else None
    for path, (k, p) in zip(paths, ((rk, rp), (sk, sp))):
        if fmt == "library_csv":
            fij.write_relation_file(path, k, p, delimiter=",")
        else:
            write_generator_file(path, k, p, gz=(fmt == "gzip"))
    lk, lp = fij.read_relation_file(paths[0], dtype, delimiter=delim)
    qk, qp = fij.read_relation_file(paths[1], dtype, delimiter=delim)
    want = orc.radix_join(rk, rp, sk, sp, 6, 4, threads=4)
    assert want == orc.sort_merge_join(rk, rp, sk, sp)
    assert want[0] > 0

    plan = fij.JoinPlan(dtype, len(lk), len(qk), 6, 4)
    got, _ = plan.execute(to_dev(lk), to_dev(lp), to_dev(qk), to_dev(qp))
    assert fij.device_status() == 0
    assert got == want

    assert operator_join(lk, lp, qk, qp, 7) == want

    host = [fij.HostArray.from_numpy(x) for x in (lk, lp, qk, qp)]
    got, _ = fij.ooc_join(dtype, *host, 6, 4, cache_bytes=len(lk) * 2 * np.dtype(dtype).itemsize,
                          staging_tuples=1 << 16)
    assert got == want

