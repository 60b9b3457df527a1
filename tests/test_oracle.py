"""Pins of the CPU oracle (oracle/oracle.c), CPU only.

Each oracle function is checked against something other than its own
formula: the reference's own test invariants (sql-ops/tests/radix_partition/
mod.rs:33-218), closed forms (test_cuda_radix_join.rs:179-182), a different
algorithm (sort-merge join in numpy), hand-computed known answers, brute
force on tiny inputs, and the compiled reference CPU hash join
(oracle/_ref, see test_golden.py).
"""
import math

import numpy as np
import pytest

import oracle_lib as orc


# ---- a1: partition function (gpu_radix_partition.h:121-127) ----------------

@pytest.mark.parametrize("key,bits,ignore,expect", [
    (182, 3, 2, 5),                  # 0b10110110 >> 2 = 0b101101, & 0b111 = 5
    (182, 0, 0, 0),                  # fanout 1
    (0x12345678, 9, 0, 0x078),       # low 9 bits
    (0x12345678, 9, 9, 0x12345678 >> 9 & 0x1FF),
    (-1, 4, 28, 15),                 # i32 -1: bits 28..31 are ones
    (-1, 4, 30, 15),                 # sign extension to 64 bits (static_cast<u64>(int))
    (-(2 ** 31), 4, 28, 8),
])
def test_key_to_partition_i32(key, bits, ignore, expect):
    assert orc.o.orc_key_to_partition_i32(key, bits, ignore) == expect
    assert orc.key_to_partition(np.array([key], np.int32), bits, ignore)[0] == expect


def test_key_to_partition_i64_high_bits():
    k = (0x5A << 40) | 0x3
    assert orc.o.orc_key_to_partition_i64(k, 8, 40) == 0x5A
    assert orc.o.orc_key_to_partition_i64(-1, 15, 49) == (1 << 15) - 1


# ---- chunking (partition_input_chunk.rs:24-35) ------------------------------

@pytest.mark.parametrize("n,chunks,dtype,expect", [
    (1000, 3, np.int32, 320), (1000, 3, np.int64, 320), (10 ** 6, 10, np.int64, 100000),
    (100, 4, np.int32, 0), (100, 4, np.int64, 16), (0, 5, np.int32, 0),
])
def test_chunk_size(n, chunks, dtype, expect):
    assert orc.chunk_size(n, chunks, dtype) == (expect, 0)


def test_chunk_size_overflow():
    assert orc.chunk_size(1 << 33, 1, np.int64)[1] == -1


# ---- a3/a4 prefix sums ---------------------------------------------------

def chunk_bounds(n, chunks, dtype):
    clen = orc.chunk_size(n, chunks, dtype)[0]
    return [(clen * c, n if c + 1 == chunks else clen * (c + 1)) for c in range(chunks)], clen


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("n,chunks,bits,ignore", [(100_003, 7, 5, 0), (1 << 16, 1, 9, 0), (50_000, 13, 4, 3),
                                                   (0, 3, 2, 0), (40, 4, 1, 0)])
def test_chunked_prefix_sum_invariants(dtype, n, chunks, bits, ignore):
    keys = orc.gen_uniform(n, 3, 0, -1000, 1 << 20, dtype)
    F = 1 << bits
    pad = orc.pad_len(dtype)
    off = orc.chunked_prefix_sum(keys, chunks, bits, ignore).astype(np.int64).reshape(chunks, F)
    bounds, clen = chunk_bounds(n, chunks, dtype)
    for c, (b, e) in enumerate(bounds):
        hist = np.bincount(orc.key_to_partition(keys[b:e], bits, ignore), minlength=F)
        # chunk c starts at its canonical slot, one padding in
        assert off[c, 0] == (clen + pad * F) * c + pad
        # consecutive partitions are separated by their size plus one padding
        assert np.all(np.diff(off[c]) == hist[:-1] + pad)
        # the chunk's last partition ends where the next chunk's first padding starts
        if c + 1 < chunks:
            assert off[c, -1] + hist[-1] == (clen + pad * F) * (c + 1)
    padded = n + F * chunks * pad
    total = sum(orc.partition_len(off.reshape(-1).astype(np.uint64), chunks, bits, padded, pad, p) for p in range(F))
    assert total == n


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("n,chunks,bits", [(100_003, 7, 5), (1 << 15, 64, 9), (33, 2, 3), (0, 4, 2)])
def test_contiguous_prefix_sum_invariants(dtype, n, chunks, bits):
    keys = orc.gen_uniform(n, 4, 0, 0, 1 << 30, dtype)
    F = 1 << bits
    pad = orc.pad_len(dtype)
    off, loc = orc.contiguous_prefix_sum(keys, chunks, bits)
    loc = loc.astype(np.int64).reshape(chunks, F)
    bounds, _ = chunk_bounds(n, chunks, dtype)
    hist = np.stack([np.bincount(orc.key_to_partition(keys[b:e], bits, 0), minlength=F) for b, e in bounds])
    # partitions are contiguous: global offsets advance by the total count plus one padding
    tot = hist.sum(axis=0)
    assert off[0] == pad
    assert np.all(np.diff(off.astype(np.int64)) == tot[:-1] + pad)
    # chunk c writes partition p right after chunk c-1's share of p (no padding between chunks)
    assert np.all(loc[0] == off.astype(np.int64))
    if chunks > 1:
        assert np.all(np.diff(loc, axis=0) == hist[:-1])
    # with a single chunk the contiguous and chunked layouts coincide
    off1, loc1 = orc.contiguous_prefix_sum(keys, 1, bits)
    np.testing.assert_array_equal(off1, orc.chunked_prefix_sum(keys, 1, bits))


# ---- a7 partition (validators of sql-ops/tests/radix_partition/mod.rs) --------

@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("contiguous", [False, True])
def test_partition_validators(dtype, contiguous):
    n, chunks, bits = 200_001, 5, 6
    F = 1 << bits
    pad = orc.pad_len(dtype)
    keys = orc.gen_perm(n, 8, dtype, base=0)
    pays = orc.gen_uniform(n, 8, 1, 0, 10000, dtype)
    if contiguous:
        off, loc = orc.contiguous_prefix_sum(keys, chunks, bits)
        rel_chunks, chunk_off = 1, loc
    else:
        off = orc.chunked_prefix_sum(keys, chunks, bits)
        rel_chunks, chunk_off = chunks, off
    rlen = n + F * rel_chunks * pad
    rel = orc.partition(keys, pays, chunks, bits, 0, chunk_off, rlen)
    # tuple_loss_or_duplicates: every input tuple exactly once, payload kept, padding null
    live = rel["key"] != -1
    assert live.sum() == n
    o1 = np.argsort(rel["key"][live], kind="stable")
    np.testing.assert_array_equal(rel["key"][live][o1], np.sort(keys))
    order = np.argsort(keys)
    np.testing.assert_array_equal(rel["value"][live][o1], pays[order])
    # verify_partitions: each tuple lies in its partition's range
    for ofi in range(rel_chunks * F):
        b = int(off[ofi])
        e = int(off[ofi + 1]) - pad if ofi + 1 < len(off) else rlen
        assert np.all(orc.key_to_partition(rel["key"][b:e], bits, 0) == ofi % F)
        assert np.all(rel["key"][b:e] != -1)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_partition_is_stable_and_transform_concatenates(dtype):
    """The oracle's NC partition keeps input order inside a partition, so the
    transform (chunk-order concatenation, mod.rs:190-218) of partition p must
    equal the input tuples with digit p in input order."""
    n, chunks, bits = 77_777, 6, 4
    keys = orc.gen_uniform(n, 9, 0, 0, 1 << 24, dtype)
    pays = orc.gen_iota(n, 0, dtype)
    off = orc.chunked_prefix_sum(keys, chunks, bits)
    rlen = n + (1 << bits) * chunks * orc.pad_len(dtype)
    rel = orc.partition(keys, pays, chunks, bits, 0, off, rlen)
    digits = orc.key_to_partition(keys, bits, 0)
    for p in range(1 << bits):
        tk, tp = orc.transform(rel, off, chunks, bits, rlen, p, dtype)
        sel = digits == p
        np.testing.assert_array_equal(tk, keys[sel])
        np.testing.assert_array_equal(tp, pays[sel])


# ---- a9 task assignment (radix_join.cu:58-98) -------------------------------

def test_assign_tasks_hand_example():
    # 4 partitions (2 bits), i64 pad 8, probe sizes 5, 1, 7, 3 -> 16 tuples, grid 3: avg = 6
    sizes = [5, 1, 7, 3]
    pad = 8
    offs, o = [], 0
    for s in sizes:
        o += pad
        offs.append(o)
        o += s
    probe_len = o
    ta = orc.assign_tasks(np.array(offs, np.uint64), probe_len, pad, 2, 3)
    # cumulative 5 (<6), 6 (>=6 cut after p1), 7 (>=6 cut after p2), then sentinel
    assert list(ta) == [0, 2, 3, 4]


def test_assign_tasks_more_blocks_than_partitions():
    offs = np.array([16, 40], np.uint64)  # i32 pad 16, 1 bit: sizes 8, 20
    ta = orc.assign_tasks(offs, 16 + 8 + 16 + 20, 16, 1, 5)
    # avg = ceil(28/5) = 6: p0 (8) cuts, p1 (20) cuts, remaining tasks get the sentinel 2
    assert list(ta) == [0, 1, 2, 2, 2, 2]


# ---- a12 hashes (gpu_common.h:143-212) ------------------------------------

def test_golden_ratio_constants():
    # floor(2^w * (sqrt(5) - 1) / 2), computed with exact integer square roots
    for w, c in ((32, 2654435769), (64, 11400714819323198485)):
        s = math.isqrt(5 << (2 * w))  # floor(sqrt(5) * 2^w)
        assert (s - (1 << w)) // 2 == c


def test_mult_shift_hash_known_answers():
    assert orc.o.orc_mult_shift_hash_i32(1, 11) == 2654435769 >> 21
    assert orc.o.orc_mult_shift_hash_i32(2, 11) == ((2 * 2654435769) % (1 << 32)) >> 21
    assert orc.o.orc_mult_shift_hash_i64(1, 20) == 11400714819323198485 >> 44
    assert orc.o.orc_mult_shift_hash_i32(0, 11) == 0


# MurmurHash3 (Appleby, SMHasher) built on the oracle's finalizers.  The body
# below is the published hash; only fmix32/fmix64 come from the oracle, so the
# published outputs of the whole hash pin both finalizers: a wrong constant or
# shift in either changes every vector.
_M32, _M64 = 0xFFFFFFFF, (1 << 64) - 1


def _rotl(x, r, bits):
    m = (1 << bits) - 1
    return ((x << r) | (x >> (bits - r))) & m


def _murmur3_x86_32(data, seed):
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h = seed & _M32
    nb = len(data) // 4
    for i in range(nb):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (_rotl((k * c1) & _M32, 15, 32) * c2) & _M32
        h = (_rotl(h ^ k, 13, 32) * 5 + 0xE6546B64) & _M32
    tail = data[4 * nb:]
    if tail:
        k = int.from_bytes(tail, "little")
        h ^= (_rotl((k * c1) & _M32, 15, 32) * c2) & _M32
    return orc.o.orc_fmix32(h ^ len(data))


def _murmur3_x64_128(data, seed):
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = seed & _M32
    nb = len(data) // 16
    for i in range(nb):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        h1 ^= (_rotl((k1 * c1) & _M64, 31, 64) * c2) & _M64
        h1 = ((_rotl(h1, 27, 64) + h2) * 5 + 0x52DCE729) & _M64
        h2 ^= (_rotl((k2 * c2) & _M64, 33, 64) * c1) & _M64
        h2 = ((_rotl(h2, 31, 64) + h1) * 5 + 0x38495AB5) & _M64
    tail = data[16 * nb:]
    if len(tail) > 8:
        k2 = int.from_bytes(tail[8:], "little")
        h2 ^= (_rotl((k2 * c2) & _M64, 33, 64) * c1) & _M64
    if tail:
        k1 = int.from_bytes(tail[:8], "little")
        h1 ^= (_rotl((k1 * c1) & _M64, 31, 64) * c2) & _M64
    h1 ^= len(data)
    h2 ^= len(data)
    h1 = (h1 + h2) & _M64
    h2 = (h2 + h1) & _M64
    h1 = orc.o.orc_fmix64(h1)
    h2 = orc.o.orc_fmix64(h2)
    h1 = (h1 + h2) & _M64
    h2 = (h2 + h1) & _M64
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def _smhasher_verification(hash_bytes):
    """SMHasher's VerificationTest: hash {}, {0}, {0,1}, ... {0..254} with seed
    256 - len, hash the concatenated results with seed 0, and read the first
    four bytes little-endian."""
    key = bytes(range(256))
    hashes = b"".join(hash_bytes(key[:i], 256 - i) for i in range(256))
    return int.from_bytes(hash_bytes(hashes, 0)[:4], "little")


def test_murmur3_x86_32_published_vectors():
    # SMHasher's verification value for MurmurHash3_x86_32
    assert _smhasher_verification(lambda d, s: _murmur3_x86_32(d, s).to_bytes(4, "little")) == 0xB0F57EE3
    # widely published single-input outputs of MurmurHash3_x86_32
    for data, seed, want in [
        (b"", 0, 0),
        (b"", 1, 0x514E28B7),
        (b"", 0xFFFFFFFF, 0x81F16F39),
        (b"\0\0\0\0", 0, 0x2362F9DE),
        (b"aaaa", 0x9747B28C, 0x5A97808A),
        (b"Hello, world!", 0x9747B28C, 0x24884CBA),
        (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
    ]:
        assert _murmur3_x86_32(data, seed) == want, (data, seed)


def test_murmur3_x64_128_published_vectors():
    # SMHasher's verification value for MurmurHash3_x64_128
    assert _smhasher_verification(_murmur3_x64_128) == 0x6384BA69
    got = _murmur3_x64_128(b"The quick brown fox jumps over the lazy dog", 0)
    assert got.hex() == "6c1b07bc7bbc4be347939ac4a93c437a"


def test_murmur3_hash_is_masked_finalizer():
    # gpu_common.h:181-212: fmix of the key, masked to log2_buckets bits; the
    # finalizer itself is pinned by the published vectors above
    for v in (0, 1, 7, -1, 123456789, -2**31):
        for lg in (1, 11, 16, 31):
            want = orc.o.orc_fmix32(v & _M32) & ((1 << lg) - 1)
            assert orc.o.orc_murmur3_hash_i32(v, lg) == want
    for v in (0, 1, -1, 2**40 + 5, -2**63):
        for lg in (1, 11, 33, 63):
            want = orc.o.orc_fmix64(v & _M64) & ((1 << lg) - 1)
            assert orc.o.orc_murmur3_hash_i64(v, lg) == want
    assert orc.o.orc_fmix32(1) == 0x514E28B7  # MurmurHash3_x86_32("", seed 1)


# ---- a10/a11/a13/a14 joins ---------------------------------------------------

@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits1,bits2", [(8, 0), (4, 5), (0, 0)])
def test_radix_join_closed_form(dtype, bits1, bits2):
    """test_cuda_radix_join.rs:179-182: PK/FK permutations, payload i+1 ->
    SUM(s.value) = N(N+1)/2, count N, SUM(r.value) = N(N+1)/2."""
    n = 6100 * 16
    rk = orc.gen_perm(n, 1, dtype, 0)
    sk = orc.gen_perm(n, 2, dtype, 0)
    rp = orc.gen_iota(n, 1, dtype)
    sp = orc.gen_iota(n, 1, dtype)
    res = orc.radix_join(rk, rp, sk, sp, bits1, bits2, threads=3)
    assert res == (n, n * (n + 1) // 2, n * (n + 1) // 2)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("bits1,bits2,threads", [(8, 0, 4), (5, 4, 2), (3, 3, 1)])
def test_radix_join_vs_sort_merge(dtype, bits1, bits2, threads):
    """Duplicates on both sides, misses, negative payloads: same count and
    wrapping sums as an independent sort-merge computation."""
    rk = orc.gen_uniform(30_000, 5, 1, 1, 10_000, dtype)
    rp = orc.gen_uniform(30_000, 5, 2, -(1 << 20), 1 << 21, dtype)
    sk = orc.gen_uniform(120_000, 5, 3, 1, 12_000, dtype)
    sp = orc.gen_uniform(120_000, 5, 4, -(1 << 20), 1 << 21, dtype)
    assert orc.radix_join(rk, rp, sk, sp, bits1, bits2, threads) == orc.sort_merge_join(rk, rp, sk, sp)


def test_radix_join_empty_sides():
    e = np.empty(0, np.int64)
    k = orc.gen_perm(100, 1, np.int64, 1)
    assert orc.radix_join(e, e, k, k, 4, 2) == (0, 0, 0)
    assert orc.radix_join(k, k, e, e, 4, 2) == (0, 0, 0)


def test_radix_join_brute_force_tiny():
    rk = np.array([1, 2, 2, 5, -7, 9], np.int64)
    rp = np.array([10, 20, 21, 50, 70, 90], np.int64)
    sk = np.array([2, 2, 5, 6, -7, 1, 1], np.int64)
    sp = np.array([1, 2, 3, 4, 5, 6, 7], np.int64)
    cnt = sb = ss = 0
    for a, x in zip(rk, rp):
        for b, y in zip(sk, sp):
            if a == b:
                cnt += 1
                sb += int(x)
                ss += int(y)
    assert orc.radix_join(rk, rp, sk, sp, 2, 1, threads=2) == (cnt, sb, ss)
    assert orc.sort_merge_join(rk, rp, sk, sp) == (cnt, sb, ss)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_null_keys_never_join(dtype):
    """-1 is the null key (gpu_common.h:109-123); inputs holding it join
    exactly like inputs without those tuples."""
    rk = orc.gen_uniform(20_000, 4, 1, -3, 500, dtype)   # includes -1
    sk = orc.gen_uniform(60_000, 4, 2, -3, 500, dtype)
    rp = orc.gen_iota(20_000, 1, dtype)
    sp = orc.gen_iota(60_000, 1, dtype)
    assert (rk == -1).any() and (sk == -1).any()
    rl, sl = rk != -1, sk != -1
    want = orc.sort_merge_join(rk[rl], rp[rl], sk[sl], sp[sl])
    assert orc.radix_join(rk, rp, sk, sp, 4, 2) == want
    assert orc.np_join_lp(rk, rp, sk, sp) == want


def test_wrapping_sums():
    """Sums wrap modulo 2^64 like the reference's release-mode i64 sums."""
    n = 1000
    rk = orc.gen_perm(n, 3, np.int64, 1)
    rp = np.full(n, np.iinfo(np.int64).max, np.int64)
    sk = rk.copy()
    sp = np.full(n, -1, np.int64)
    cnt, sb, ss = orc.radix_join(rk, rp, sk, sp, 3, 0)
    assert cnt == n
    assert sb == (n * (2 ** 63 - 1)) % (1 << 64)
    assert ss == (-n) % (1 << 64)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_perfect_join_matches_chaining_on_dense_keys(dtype):
    n, bits = 1 << 14, 4
    F = 1 << bits
    pad = orc.pad_len(dtype)
    rk = orc.gen_perm(n, 6, dtype, 0)
    sk = orc.gen_uniform(4 * n, 6, 1, 0, n + 100, dtype)  # some misses above n-1
    rp, sp = orc.gen_iota(n, 1, dtype), orc.gen_iota(4 * n, 1, dtype)
    rels = []
    for k, p in ((rk, rp), (sk, sp)):
        off = orc.chunked_prefix_sum(k, 1, bits)
        rlen = len(k) + F * pad
        rels.append((orc.partition(k, p, 1, bits, 0, off, rlen), off, rlen))
    (b, bo, bl), (q, qo, ql) = rels
    chain, rc1 = orc.partitioned_join(b, bo, bl, q, qo, ql, bits, bits, dtype)
    perfect, rc2 = orc.partitioned_join_perfect(b, bo, bl, q, qo, ql, bits, bits, (n >> bits) + 1, dtype)
    assert rc1 == 0 and rc2 == 0
    assert chain == perfect == orc.sort_merge_join(rk, rp, sk, sp)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_np_join_vs_sort_merge(dtype):
    rk = orc.gen_uniform(5000, 2, 1, 1, 3000, dtype)
    rp = orc.gen_uniform(5000, 2, 2, -9, 100, dtype)
    sk = orc.gen_uniform(20000, 2, 3, 1, 3500, dtype)
    sp = orc.gen_uniform(20000, 2, 4, -9, 100, dtype)
    assert orc.np_join_lp(rk, rp, sk, sp) == orc.sort_merge_join(rk, rp, sk, sp)


# ---- generators ---------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1 << 16, 100_003])
def test_permutation_is_bijection(n):
    k = orc.gen_perm(n, 12345, np.int64, base=1)
    np.testing.assert_array_equal(np.sort(k), np.arange(1, n + 1))


def test_permutation_depends_on_seed():
    a = orc.gen_perm(1000, 1, np.int64)
    b = orc.gen_perm(1000, 2, np.int64)
    assert not np.array_equal(a, b)


def test_uniform_chi_square():
    n, bins = 1 << 20, 64
    x = orc.gen_uniform(n, 77, 1, 1, bins, np.int64)
    assert x.min() == 1 and x.max() == bins
    h = np.bincount(x - 1, minlength=bins)
    chi2 = ((h - n / bins) ** 2 / (n / bins)).sum()
    assert chi2 < 120  # 63 degrees of freedom, p ~ 1e-5


@pytest.mark.parametrize("s", [0.5, 1.0, 1.25])
def test_zipf_rank_frequencies(s):
    N, n = 1000, 400_000
    r = orc.gen_zipf_ranks(n, N, s, seed=5)
    assert r.min() >= 1 and r.max() <= N
    H = sum(k ** -s for k in range(1, N + 1))
    h = np.bincount(r.astype(np.int64), minlength=N + 1)
    for k in (1, 2, 10):
        p = k ** -s / H
        sd = math.sqrt(n * p * (1 - p))
        assert abs(h[k] - n * p) < 6 * sd, (k, h[k], n * p)


def test_zipf_zero_is_uniform():
    r = orc.gen_zipf_ranks(200_000, 100, 0.0, seed=3)
    h = np.bincount(r.astype(np.int64), minlength=101)[1:]
    chi2 = ((h - 2000) ** 2 / 2000).sum()
    assert chi2 < 180


def test_zipf_keys_are_permutation_keys():
    """Zipf keys map ranks through the permutation relation's keys: the most
    frequent key is that relation's key at index 0 (rank 1)."""
    N = 4096
    keys = orc.gen_zipf_keys(100_000, N, 1.0, seed=2, stream=1, perm_seed=77, base=1, dtype=np.int64)
    rel = orc.gen_perm(N, 77, np.int64, 1)
    assert np.isin(keys, rel).all()
    vals, counts = np.unique(keys, return_counts=True)
    assert vals[np.argmax(counts)] == rel[0]
    ranks = orc.gen_zipf_ranks(100_000, N, 1.0, seed=2, stream=1)
    assert np.array_equal(np.sort(counts)[::-1][:5], np.sort(np.bincount(ranks.astype(np.int64)))[::-1][:5])
