"""GPU parity of the multiply-shift hash (sql-ops/include/gpu_common.h:143-168,
:214-217 `hash` = mult_shift_hash) as the no-partitioning join's table layout
shows it (sql-ops/cudautils/no_partitioning_join.cu:39-132: slot =
hash(key, log2 entries), then linear probing).

The table buffer is read back and compared slot for slot:
* a single key lands on the slot computed here from the golden-ratio
  constant (derived by integer square root, not taken from any source);
* keys whose hashes are distinct give one possible layout, equal to the
  oracle's table;
* with collisions, linear probing's occupied slots and each cluster's tuple
  multiset do not depend on insertion order, so they equal the oracle's too.
"""
import math

import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import sync, to_dev

pytestmark = pytest.mark.gpu


def golden(w):
    # floor(2^w * (sqrt(5) - 1) / 2)
    return (math.isqrt(5 << (2 * w)) - (1 << w)) // 2


def slot_closed_form(key, w, lg):
    return ((key % (1 << w)) * golden(w) % (1 << w)) >> (w - lg)


def kt_of(dtype):
    return fij.KeyType.I32 if dtype == np.int32 else fij.KeyType.I64


def gpu_table(keys, pays, entries, dtype, dims=(64, 256)):
    ht = fij.HashTable(entries, kt_of(dtype))
    hj = fij.HipHashJoin(fij.HashingScheme.LinearProbing, ht, dims, dims)
    hj.build(to_dev(keys), to_dev(pays))
    sync()
    assert fij.device_status() == 0
    raw = ht.mem.to_host().reshape(entries, 2)
    return raw[:, 0].copy(), raw[:, 1].copy()


def oracle_table(keys, pays, entries, dtype):
    # every 32-bit word all ones, as HashTable::new_on_gpu initialises device
    # tables (no_partitioning_join.rs:611-620, cuMemsetD32 of the null key)
    table = np.zeros(entries, orc.tuple_dtype(dtype))
    table["key"] = -1
    table["value"] = -1
    getattr(orc.o, f"orc_np_build_lp_{orc.suffix(dtype)}")(orc.P(table), entries, orc.P(keys), orc.P(pays),
                                                          len(keys))
    return table["key"].copy(), table["value"].copy()


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("lg", [1, 11, 20])
def test_single_key_slot_known_answer(dtype, lg):
    w = 32 if dtype == np.int32 else 64
    entries = 1 << lg
    for key in (1, 2, 3, 12345, 2**31 - 1, -5):
        keys = np.array([key], dtype)
        pays = np.array([7], dtype)
        k, v = gpu_table(keys, pays, entries, dtype, dims=(1, 64))
        occupied = np.nonzero(k != -1)[0]
        assert list(occupied) == [slot_closed_form(key, w, lg)], (key, lg)
        assert v[occupied[0]] == 7


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_collision_free_layout_equals_oracle(dtype):
    entries = 1 << 16
    lg = 16
    cand = orc.gen_uniform(40_000, 11, 1, 1, 1 << 30, dtype)
    cand = np.unique(cand)
    hsh = np.array([getattr(orc.o, f"orc_mult_shift_hash_{orc.suffix(dtype)}")(int(c), lg) for c in cand])
    _, first = np.unique(hsh, return_index=True)
    keys = cand[np.sort(first)]
    assert len(keys) > 10_000
    pays = orc.gen_uniform(len(keys), 11, 2, -1000, 1 << 20, dtype)
    gk, gv = gpu_table(keys, pays, entries, dtype)
    ok, ov = oracle_table(keys, pays, entries, dtype)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gv, ov)
    # and every key sits on its closed-form slot
    w = 32 if dtype == np.int32 else 64
    idx = np.nonzero(gk != -1)[0]
    for s in idx[:: max(1, len(idx) // 500)]:
        assert slot_closed_form(int(gk[s]), w, lg) == s


def clusters(keys):
    """Maximal runs of occupied slots of a circular table, as (start, end)."""
    occ = keys != -1
    n = len(keys)
    if occ.all():
        return [(0, n)]
    start = int(np.nonzero(~occ)[0][0]) + 1  # begin right after an empty slot
    out, i = [], 0
    while i < n:
        s = (start + i) % n
        if occ[s]:
            j = i
            while j < n and occ[(start + j) % n]:
                j += 1
            out.append(((start + i) % n, j - i))
            i = j
        else:
            i += 1
    return out


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_colliding_layout_equals_oracle(dtype):
    entries = 1 << 15
    keys = orc.gen_perm(entries // 2, 12, dtype, 1) * 3  # unique, load factor 1/2
    pays = orc.gen_uniform(len(keys), 12, 3, 0, 1 << 20, dtype)
    gk, gv = gpu_table(keys, pays, entries, dtype)
    ok, ov = oracle_table(keys, pays, entries, dtype)
    np.testing.assert_array_equal(gk != -1, ok != -1)
    cl = clusters(ok)
    assert any(length > 1 for _, length in cl)  # the case has collisions
    for s, length in cl:
        sl = [(s + i) % entries for i in range(length)]
        assert sorted(zip(gk[sl], gv[sl])) == sorted(zip(ok[sl], ov[sl]))
