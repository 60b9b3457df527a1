"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it.  It is
never used by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref_np.so")


def _load():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True, capture_output=True)
    return C.CDLL(ORACLE_SO)


o = _load()


class Res(C.Structure):
    _fields_ = [("count", C.c_uint64), ("sum_build", C.c_uint64), ("sum_probe", C.c_uint64)]

    def tuple(self):
        return (self.count, self.sum_build, self.sum_probe)


_u32, _u64, _i64, _vp, _i32 = C.c_uint32, C.c_uint64, C.c_int64, C.c_void_p, C.c_int32


def _proto(name, restype, *args):
    f = getattr(o, name)
    f.restype = restype
    f.argtypes = list(args)


_proto("orc_mix64", _u64, _u64)
_proto("orc_rand_u64", _u64, _u64, _u64, _u64)
_proto("orc_feistel_perm", _u64, _u64, _u64, _u64)
for s in ("i32", "i64"):
    _proto(f"orc_gen_perm_keys_{s}", None, _vp, _u64, _u64, _i64)
    _proto(f"orc_gen_uniform_{s}", None, _vp, _u64, _u64, _u64, _i64, _u64)
    _proto(f"orc_gen_iota_{s}", None, _vp, _u64, _i64)
    _proto(f"orc_key_to_partition_{s}", _u32, _i64 if s == "i64" else _i32, _u32, _u32)
    _proto(f"orc_chunked_prefix_sum_{s}", None, _vp, _u64, _u32, _u32, _u32, _vp)
    _proto(f"orc_contiguous_prefix_sum_{s}", None, _vp, _u64, _u32, _u32, _u32, _vp, _vp)
    _proto(f"orc_partition_{s}", None, _vp, _vp, _u64, _u32, _u32, _u32, _vp, _vp)
    _proto(f"orc_transform_{s}", _u64, _vp, _vp, _u32, _u32, _u64, _u32, _vp, _vp)
    _proto(f"orc_partitioned_join_{s}", C.c_int, _vp, _vp, _u64, _vp, _vp, _u64, _u32, _u32, _u64, C.POINTER(Res))
    _proto(f"orc_partitioned_join_perfect_{s}", C.c_int, _vp, _vp, _u64, _vp, _vp, _u64, _u32, _u32, _u64, C.POINTER(Res))
    _proto(f"orc_np_build_lp_{s}", None, _vp, _u64, _vp, _vp, _u64)
    _proto(f"orc_np_probe_lp_{s}", None, _vp, _u64, _vp, _vp, _u64, C.POINTER(Res))
    _proto(f"orc_radix_join_{s}", C.c_int, _vp, _vp, _u64, _vp, _vp, _u64, _u32, _u32, _u32, C.POINTER(Res))
_proto("orc_gen_zipf_ranks", C.c_int, _vp, _u64, _u64, C.c_double, _u64, _u64)
for s_ in ("i32", "i64"):
    _proto(f"orc_gen_zipf_keys_{s_}", C.c_int, _vp, _u64, _u64, C.c_double, _u64, _u64, _u64, _i64)
_proto("orc_input_chunk_size", C.c_int, _u64, _u32, _u32, C.POINTER(_u64))
_proto("orc_partition_len", _u64, _vp, _u32, _u32, _u64, _u32, _u32)
_proto("orc_assign_tasks", None, _vp, _u32, _u32, _u32, _u32, _vp)
_proto("orc_mult_shift_hash_i32", _i32, _i32, _u32)
_proto("orc_mult_shift_hash_i64", _i64, _i64, _u32)
_proto("orc_fmix32", _u32, _u32)
_proto("orc_fmix64", _u64, _u64)
_proto("orc_murmur3_hash_i32", _i32, _i32, _u32)
_proto("orc_murmur3_hash_i64", _i64, _i64, _u32)


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def suffix(dtype):
    return "i32" if np.dtype(dtype) == np.int32 else "i64"


def tuple_dtype(dtype):
    return np.dtype([("key", dtype), ("value", dtype)])


def pad_len(dtype):
    return 128 // (2 * np.dtype(dtype).itemsize)


# ---- generators -----------------------------------------------------------

def gen_perm(n, seed, dtype=np.int64, base=1):
    out = np.empty(n, dtype)
    getattr(o, f"orc_gen_perm_keys_{suffix(dtype)}")(P(out), n, seed, base)
    return out


def gen_uniform(n, seed, stream, lo, count, dtype=np.int64):
    out = np.empty(n, dtype)
    getattr(o, f"orc_gen_uniform_{suffix(dtype)}")(P(out), n, seed, stream, lo, count)
    return out


def gen_iota(n, base=0, dtype=np.int64):
    out = np.empty(n, dtype)
    getattr(o, f"orc_gen_iota_{suffix(dtype)}")(P(out), n, base)
    return out


def gen_zipf_keys(n, num_ranks, exponent, seed, stream, perm_seed, base=1, dtype=np.int64):
    """Zipf(exponent) keys over the keys gen_perm(num_ranks, perm_seed, dtype, base) makes."""
    out = np.empty(n, dtype)
    rc = getattr(o, f"orc_gen_zipf_keys_{suffix(dtype)}")(P(out), n, num_ranks, exponent, seed, stream, perm_seed,
                                                          base)
    assert rc == 0
    return out


def gen_zipf_ranks(n, num_ranks, exponent, seed, stream=7):
    out = np.empty(n, np.uint64)
    rc = o.orc_gen_zipf_ranks(P(out), n, num_ranks, exponent, seed, stream)
    assert rc == 0
    return out


# ---- partitioning ----------------------------------------------------------

def key_to_partition(keys, radix_bits, ignore_bits):
    """Vectorised restatement used for checking (gpu_radix_partition.h:121-127)."""
    k = np.asarray(keys).astype(np.int64).view(np.uint64)
    mask = np.uint64(((1 << radix_bits) - 1) << ignore_bits)
    return ((k & mask) >> np.uint64(ignore_bits)).astype(np.uint32)


def chunk_size(n, chunks, dtype):
    out = _u64()
    rc = o.orc_input_chunk_size(n, chunks, np.dtype(dtype).itemsize, C.byref(out))
    return out.value, rc


def chunked_prefix_sum(keys, chunks, radix_bits, ignore_bits=0):
    F = 1 << radix_bits
    out = np.empty(chunks * F, np.uint64)
    getattr(o, f"orc_chunked_prefix_sum_{suffix(keys.dtype)}")(P(keys), len(keys), chunks, radix_bits, ignore_bits, P(out))
    return out


def contiguous_prefix_sum(keys, chunks, radix_bits, ignore_bits=0):
    F = 1 << radix_bits
    off = np.empty(F, np.uint64)
    loc = np.empty(chunks * F, np.uint64)
    getattr(o, f"orc_contiguous_prefix_sum_{suffix(keys.dtype)}")(P(keys), len(keys), chunks, radix_bits, ignore_bits,
                                                                  P(off), P(loc))
    return off, loc


def partition(keys, pays, chunks, radix_bits, ignore_bits, chunk_offsets, relation_len):
    rel = np.zeros(relation_len, tuple_dtype(keys.dtype))
    rel["key"] = -1
    getattr(o, f"orc_partition_{suffix(keys.dtype)}")(P(keys), P(pays), len(keys), chunks, radix_bits, ignore_bits,
                                                      P(chunk_offsets), P(rel))
    return rel


def partition_len(offsets, chunks, radix_bits, padded_len, pad, pid):
    return o.orc_partition_len(P(offsets), chunks, radix_bits, padded_len, pad, pid)


def transform(src_rel, src_offsets, src_chunks, src_radix_bits, src_padded_len, pid, dtype):
    n = sum(partition_len(src_offsets, src_chunks, src_radix_bits, src_padded_len, pad_len(dtype), pid) for _ in [0])
    dk = np.empty(n, dtype)
    dp = np.empty(n, dtype)
    m = getattr(o, f"orc_transform_{suffix(dtype)}")(P(src_rel), P(src_offsets), src_chunks, src_radix_bits,
                                                     src_padded_len, pid, P(dk), P(dp))
    assert m == n
    return dk, dp


def assign_tasks(probe_offsets, probe_len, pad, radix_bits, grid):
    out = np.empty(grid + 1, np.uint32)
    o.orc_assign_tasks(P(probe_offsets), probe_len, pad, radix_bits, grid, P(out))
    return out


# ---- joins -----------------------------------------------------------------

def partitioned_join(build, boff, bplen, probe, poff, pplen, radix_bits, ignore_bits, dtype, max_build=0):
    r = Res()
    rc = getattr(o, f"orc_partitioned_join_{suffix(dtype)}")(P(build), P(boff), bplen, P(probe), P(poff), pplen,
                                                            radix_bits, ignore_bits, max_build, C.byref(r))
    return r.tuple(), rc


def partitioned_join_perfect(build, boff, bplen, probe, poff, pplen, radix_bits, ignore_bits, ht_entries, dtype):
    r = Res()
    rc = getattr(o, f"orc_partitioned_join_perfect_{suffix(dtype)}")(P(build), P(boff), bplen, P(probe), P(poff),
                                                                    pplen, radix_bits, ignore_bits, ht_entries,
                                                                    C.byref(r))
    return r.tuple(), rc


def np_join_lp(rk, rp, sk, sp, entries=None):
    dtype = rk.dtype
    if entries is None:
        entries = 2 * (1 << max(0, (len(rk) - 1).bit_length()))
    table = np.zeros(entries, tuple_dtype(dtype))
    table["key"] = -1
    getattr(o, f"orc_np_build_lp_{suffix(dtype)}")(P(table), entries, P(rk), P(rp), len(rk))
    r = Res()
    getattr(o, f"orc_np_probe_lp_{suffix(dtype)}")(P(table), entries, P(sk), P(sp), len(sk), C.byref(r))
    return r.tuple()


def radix_join(rk, rp, sk, sp, bits1, bits2=0, threads=4):
    r = Res()
    rc = getattr(o, f"orc_radix_join_{suffix(rk.dtype)}")(P(rk), P(rp), len(rk), P(sk), P(sp), len(sk), bits1, bits2,
                                                         threads, C.byref(r))
    assert rc == 0
    return r.tuple()


def sort_merge_join(rk, rp, sk, sp):
    """Independent textbook check: count and wrapping sums of an equi-join
    computed by sorting and grouping (numpy), not by hashing.  The null key
    (-1) never joins."""
    rk = np.asarray(rk, np.int64)
    sk = np.asarray(sk, np.int64)
    rp = np.asarray(rp, np.int64).view(np.uint64)
    sp = np.asarray(sp, np.int64).view(np.uint64)
    rl, sl = rk != -1, sk != -1
    rk, rp, sk, sp = rk[rl], rp[rl], sk[sl], sp[sl]
    ur, r_inv = np.unique(rk, return_inverse=True)
    r_cnt = np.bincount(r_inv).astype(np.uint64)
    with np.errstate(over="ignore"):
        r_sum = np.zeros(len(ur), np.uint64)
        np.add.at(r_sum, r_inv, rp)
        pos = np.searchsorted(ur, sk)
        pos_c = np.minimum(pos, len(ur) - 1) if len(ur) else pos
        hit = (pos < len(ur)) & (ur[pos_c] == sk) if len(ur) else np.zeros(len(sk), bool)
        cnt = np.uint64(r_cnt[pos_c[hit]].sum(dtype=np.uint64)) if hit.any() else np.uint64(0)
        sum_build = np.uint64(r_sum[pos_c[hit]].sum(dtype=np.uint64)) if hit.any() else np.uint64(0)
        sum_probe = np.uint64((sp[hit] * r_cnt[pos_c[hit]]).sum(dtype=np.uint64)) if hit.any() else np.uint64(0)
    return int(cnt), int(sum_build), int(sum_probe)
