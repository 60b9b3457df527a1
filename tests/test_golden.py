"""The oracle against golden vectors written by the reference's own CPU hash
join (tests/golden/gen_golden.py; oracle/_ref built from the reference
sources).  Runs on CPU."""
import json
import os

import numpy as np
import pytest

import oracle_lib as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_fixtures():
    with open(os.path.join(GOLDEN, "join_fixtures.json")) as f:
        return json.load(f)["fixtures"]


def generate(case):
    # the generator specification is data in the fixture; rebuild the inputs
    dtype = np.int32 if case["dtype"] == "i32" else np.int64

    def gen(spec, n, keys=None):
        kind = spec[0]
        if kind == "perm":
            return orc.gen_perm(n, spec[1], dtype, spec[2])
        if kind == "uniform":
            return orc.gen_uniform(n, spec[1], spec[2], spec[3], spec[4], dtype)
        if kind == "iota":
            return orc.gen_iota(n, spec[1], dtype)
        if kind == "copy_keys":
            return keys.copy()
        raise ValueError(kind)

    rk = gen(case["r_keys"], case["r_n"])
    sk = gen(case["s_keys"], case["s_n"])
    rp = gen(case["r_pay"], case["r_n"], rk)
    sp = gen(case["s_pay"], case["s_n"], sk)
    return rk, rp, sk, sp


def digest(*arrays):
    import hashlib

    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


FIXTURES = load_fixtures()


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_generators_unchanged(case):
    assert digest(*generate(case)) == case["sha256"]


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
@pytest.mark.parametrize("bits", [(6, 0), (4, 3), (1, 1)])
def test_oracle_radix_join_matches_reference(case, bits):
    rk, rp, sk, sp = generate(case)
    want = (case["count"], case["sum_build"], case["sum_probe"])
    assert orc.radix_join(rk, rp, sk, sp, bits[0], bits[1], threads=3) == want


@pytest.mark.parametrize("case", FIXTURES, ids=[c["name"] for c in FIXTURES])
def test_oracle_np_join_matches_reference(case):
    rk, rp, sk, sp = generate(case)
    assert orc.np_join_lp(rk, rp, sk, sp) == (case["count"], case["sum_build"], case["sum_probe"])


def test_tiny_explicit_case():
    d = np.load(os.path.join(GOLDEN, "tiny_case.npz"))
    want = tuple(int(x) for x in d["expected"])
    assert orc.radix_join(d["rk"], d["rp"], d["sk"], d["sp"], 2, 1, threads=2) == want
    assert orc.sort_merge_join(d["rk"], d["rp"], d["sk"], d["sp"]) == want


@pytest.mark.skipif(not os.path.exists(orc.REF_SO), reason="reference CPU join not built (needs the reference sources)")
def test_live_reference_agrees_on_new_seed():
    """Where oracle/_ref exists, re-derive one fresh case live."""
    import ctypes as C

    ref = C.CDLL(orc.REF_SO)
    rk = orc.gen_uniform(3000, 99, 1, 1, 1000, np.int64)
    rp = orc.gen_uniform(3000, 99, 2, -100, 200, np.int64)
    sk = orc.gen_uniform(9000, 99, 3, 1, 1200, np.int64)
    sp = orc.gen_uniform(9000, 99, 4, -100, 200, np.int64)
    entries = 8192
    table = np.zeros(entries, orc.tuple_dtype(np.int64))
    table["key"] = -1
    ref.cpu_ht_build_linearprobing_int64(orc.P(table), C.c_uint64(entries), orc.P(rk), orc.P(rp), C.c_uint64(3000))
    out = np.zeros(1, np.uint64)
    ref.cpu_ht_probe_aggregate_linearprobing_int64(orc.P(table), C.c_uint64(entries), orc.P(sk), orc.P(sp),
                                                   C.c_uint64(9000), orc.P(out))
    assert int(out[0]) == orc.radix_join(rk, rp, sk, sp, 5, 2)[2]
