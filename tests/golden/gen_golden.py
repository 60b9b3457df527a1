"""Writes tests/golden/join_fixtures.json and tiny_case.npz from the
reference's own CPU hash join, compiled unmodified from its sources into
oracle/_ref/libref_np.so (make -C oracle ref).

The reference CPU join (sql-ops/cpputils/no_partitioning_join.cpp:27-163)
reports SUM(probe.value) only.  Payload substitution recovers all three
aggregates of an equi-join R |x| S from it:
  * SUM(S.value): build R, probe S with its payloads;
  * match count:  build R, probe S with every payload set to 1;
  * SUM(R.value): build S, probe R with R's payloads (each R tuple adds its
    payload once per matching S tuple).
Inputs come from the repository's seeded generators; each fixture records
the generator parameters and a SHA-256 of the generated columns, so a change
of the generators is detected instead of silently changing the inputs.

    python tests/golden/gen_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as orc  # noqa: E402

REF = C.CDLL(orc.REF_SO)


def ref_probe_sum(build_k, build_p, probe_k, probe_p):
    dtype = build_k.dtype
    suf = "int32" if dtype == np.int32 else "int64"
    entries = 2 * (1 << max(0, (len(build_k) - 1).bit_length()))
    table = np.zeros(entries, orc.tuple_dtype(dtype))
    table["key"] = -1  # HashTable::new_on_cpu sets the keys to the null key
    getattr(REF, f"cpu_ht_build_linearprobing_{suf}")(orc.P(table), C.c_uint64(entries), orc.P(build_k),
                                                      orc.P(build_p), C.c_uint64(len(build_k)))
    out = np.zeros(1, np.uint64)
    getattr(REF, f"cpu_ht_probe_aggregate_linearprobing_{suf}")(orc.P(table), C.c_uint64(entries), orc.P(probe_k),
                                                                orc.P(probe_p), C.c_uint64(len(probe_k)), orc.P(out))
    return int(out[0])


def ref_join_aggregates(rk, rp, sk, sp):
    sum_probe = ref_probe_sum(rk, rp, sk, sp)
    count = ref_probe_sum(rk, rp, sk, np.ones_like(sp))
    sum_build = ref_probe_sum(sk, sp, rk, rp)
    return count, sum_build, sum_probe


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


CASES = [
    # name, dtype, |R|, |S|, R key gen, S key gen, payload gens
    dict(name="pk_fk_i32", dtype="i32", r_n=4096, s_n=65536, r_keys=("perm", 11, 1), s_keys=("uniform", 11, 1, 1, 4096),
         r_pay=("copy_keys",), s_pay=("iota", 0)),
    dict(name="pk_fk_i64", dtype="i64", r_n=8192, s_n=32768, r_keys=("perm", 12, 1), s_keys=("uniform", 12, 1, 1, 8192),
         r_pay=("copy_keys",), s_pay=("iota", 0)),
    dict(name="dups_misses_i32", dtype="i32", r_n=6000, s_n=50000, r_keys=("uniform", 13, 1, 1, 2000),
         s_keys=("uniform", 13, 2, 1, 2600), r_pay=("uniform", 13, 3, -500, 1000), s_pay=("uniform", 13, 4, -500, 1000)),
    dict(name="dups_misses_i64", dtype="i64", r_n=6000, s_n=50000, r_keys=("uniform", 14, 1, 1, 2000),
         s_keys=("uniform", 14, 2, 1, 2600), r_pay=("uniform", 14, 3, -(1 << 40), 1 << 41),
         s_pay=("uniform", 14, 4, -(1 << 40), 1 << 41)),
    dict(name="c1_scaled_i32", dtype="i32", r_n=1 << 14, s_n=1 << 18, r_keys=("perm", 42, 1),
         s_keys=("uniform", 42, 1, 1, 1 << 14), r_pay=("copy_keys",), s_pay=("iota", 0)),
    # negative keys; -1 is the null key, which the reference reserves (its CPU
    # probe never terminates on it), so the ranges stop at -2
    dict(name="negative_keys_i64", dtype="i64", r_n=3000, s_n=20000, r_keys=("uniform", 15, 1, -5000, 4000),
         s_keys=("uniform", 15, 2, -5000, 4999), r_pay=("iota", 7), s_pay=("iota", -3)),
]


def generate(case):
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


def main():
    fixtures = []
    for case in CASES:
        rk, rp, sk, sp = generate(case)
        count, sb, sp_sum = ref_join_aggregates(rk, rp, sk, sp)
        fixtures.append(dict(case, sha256=digest(rk, rp, sk, sp), count=count, sum_build=sb, sum_probe=sp_sum))
    with open(os.path.join(HERE, "join_fixtures.json"), "w") as f:
        json.dump({"source": "reference sql-ops/cpputils/no_partitioning_join.cpp (linear probing), "
                             "compiled from the reference sources; aggregates by payload substitution",
                   "fixtures": fixtures}, f, indent=1)
    # a tiny explicit case that does not depend on the generators
    rk = np.array([5, 3, 9, 3, 12, -7, 0, 21], np.int64)
    rp = np.array([50, 30, 90, 31, 120, -70, 1, 210], np.int64)
    sk = np.array([3, 3, 9, 4, -7, 21, 21, 21, 100, 0, 5], np.int64)
    sp = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11], np.int64)
    count, sb, sp_sum = ref_join_aggregates(rk, rp, sk, sp)
    np.savez(os.path.join(HERE, "tiny_case.npz"), rk=rk, rp=rp, sk=sk, sp=sp,
             expected=np.array([count, sb, sp_sum], np.uint64))
    print(f"wrote {len(fixtures)} fixtures")


if __name__ == "__main__":
    main()
