"""Skewed joins on the GPU (the Zipf-probe configuration at reduced size).

A Zipf-distributed probe side puts a large share of S on one key and so on
one (pass-1, pass-2) partition; the plan splits such partitions into work
units joined by many waves.  Results must equal the oracle's and an
independent sort-merge join on the same inputs.  The seeded Zipf generator
(datagen/src/relation.rs:218-290 uses the unseeded `zipf` crate) has its own
definition, restated identically in the oracle.
"""
import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("s", [0.5, 1.0, 1.25])
def test_gen_zipf_matches_oracle(dtype, s):
    n, N = 1 << 20, 1 << 16
    want = orc.gen_zipf_keys(n, N, s, seed=11, stream=3, perm_seed=42, base=1, dtype=dtype)
    out = fij.DeviceArray(n, dtype)
    fij.gen_zipf(out, N, s, 11, 3, 42, 1)
    got = out.to_host()
    # same formulas in double precision; a transcendental's last bit may
    # differ between the host and device math libraries at a rejection edge
    assert (got == want).mean() >= 0.9999
    keys = orc.gen_perm(N, 42, dtype, 1)
    assert np.isin(got, keys).all()


def zipf_case(n_r, n_s, s, dtype, seed=5):
    rk = orc.gen_perm(n_r, 42, dtype, 1)
    rp = orc.gen_iota(n_r, 1, dtype)
    sk = orc.gen_zipf_keys(n_s, n_r, s, seed=seed, stream=1, perm_seed=42, base=1, dtype=dtype)
    sp = orc.gen_iota(n_s, 0, dtype)
    return rk, rp, sk, sp


def run_plan(rk, rp, sk, sp, bits1, bits2):
    plan = fij.JoinPlan(rk.dtype, len(rk), len(sk), bits1, bits2)
    got, _ = plan.execute(to_dev(rk), to_dev(rp), to_dev(sk), to_dev(sp))
    assert fij.device_status() == 0
    return got


@pytest.mark.parametrize("s", [0.0, 0.5, 1.0, 1.25])
@pytest.mark.parametrize("bits", [(8, 8), (9, 9), (6, 0)])
def test_zipf_probe_i32(s, bits):
    rk, rp, sk, sp = zipf_case(1 << 20, 1 << 22, s, np.int32)
    got = run_plan(rk, rp, sk, sp, *bits)
    want = orc.sort_merge_join(rk, rp, sk, sp)
    assert got == want
    assert got[0] == len(sk)  # every probe key is one of R's keys


@pytest.mark.parametrize("bits", [(9, 9), (8, 0)])
def test_zipf_probe_i64(bits):
    rk, rp, sk, sp = zipf_case(1 << 19, 1 << 22, 1.25, np.int64, seed=9)
    got = run_plan(rk, rp, sk, sp, *bits)
    assert got == orc.radix_join(rk, rp, sk, sp, bits[0], bits[1], threads=8)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_heavy_build_key(dtype):
    """2 % of R carries one key: its partition's build side exceeds the LDS
    table, is built in several fills, and every probe match counts every
    duplicate (about 5 M matches on that key)."""
    n_r, n_s = 1 << 18, 1 << 20
    rk = orc.gen_uniform(n_r, 3, 1, 1, 1 << 17, dtype)
    rk[: n_r // 50] = 4242
    rp = orc.gen_uniform(n_r, 3, 2, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(n_s, 3, 3, 1, 1 << 17, dtype)
    sk[::997] = 4242
    sp = orc.gen_iota(n_s, 0, dtype)
    want = orc.sort_merge_join(rk, rp, sk, sp)
    for bits in ((9, 9), (7, 0)):
        assert run_plan(rk, rp, sk, sp, *bits) == want


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits", [(9, 9), (6, 0)])
def test_large_build_partitions_in_hbm_tables(dtype, bits):
    """Build keys repeated far beyond a block's LDS table (40,000 and 15,000
    copies of two keys, and a third key whose copies miss every probe): their
    items get hash tables in the plan's HBM arena, built once and probed by
    many units.  The same plan joins three different inputs in a row, so the
    arena must be back to all ones after every join."""
    n_r, n_s = 1 << 18, 1 << 21
    plan = fij.JoinPlan(dtype, n_r, n_s, *bits)
    for rep in range(3):
        rk = orc.gen_uniform(n_r, 40 + rep, 1, 1, 1 << 17, dtype)
        rk[:40_000] = 777 + rep
        rk[40_000:55_000] = 90_001
        rk[55_000:70_000] = 99_999_999  # never probed
        rp = orc.gen_uniform(n_r, 40 + rep, 2, -(1 << 20), 1 << 21, dtype)
        sk = orc.gen_uniform(n_s, 40 + rep, 3, 1, 1 << 17, dtype)
        sk[::1009] = 777 + rep
        sk[5::2003] = 90_001
        sp = orc.gen_iota(n_s, rep, dtype)
        got, t = plan.execute(to_dev(rk), to_dev(rp), to_dev(sk), to_dev(sp))
        assert fij.device_status() == 0
        assert got == orc.sort_merge_join(rk, rp, sk, sp), rep
        assert got[0] > 40_000 * (n_s // 1009)
