"""Served joins (fij_join_server_*): build sides partitioned once and kept
resident, probe sides admitted to free slots as they arrive.  Every request's
result must be the same alone (submitted and waited for one at a time) and in
the mix (all submitted at once, overlapping on the device, waited for in
another order), and equal to the oracle's sort-merge join of the pair."""
import threading

import numpy as np
import pytest

import fij
import oracle_lib as orc
from gpu_helpers import to_dev

pytestmark = pytest.mark.gpu


def make_build(i, n, dtype, dup=False):
    if dup:  # duplicates and null keys
        rk = orc.gen_uniform(n, 300 + i, 1, -2, n // 3, dtype)
    else:
        rk = orc.gen_perm(n, 300 + i, dtype, 1)
    rp = orc.gen_uniform(n, 300 + i, 2, -(1 << 20), 1 << 21, dtype)
    return rk, rp


def make_probe(j, n_b, n, dtype):
    # keys 80% over the build side's keys 1..n_b, 20% above them
    sk = orc.gen_uniform(n, 700 + j, 1, 1, n_b + n_b // 4 + 1, dtype)
    sp = orc.gen_iota(n, j, dtype)
    return sk, sp


@pytest.mark.parametrize("table_build_max", [0, 1 << 17])
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_server_alone_equals_mix_equals_oracle(dtype, table_build_max):
    """table_build_max = 2^17: the two smaller builds (one with duplicates
    and null keys) are resident hash tables, the larger ones partitioned, in
    the same mix.  Builds of at most 2^21 (int64) or 2^22 (int32) tuples that
    are not tables are partitioned in one pass (fij_join_plan_bits; from 2^20
    or 2^21 on with 128 KiB block tables), the larger build in two."""
    big = (1 << (22 if dtype == np.int32 else 21)) + 4097
    builds = [make_build(0, 4096, dtype), make_build(1, 1 << 17, dtype, dup=True), make_build(2, 300000, dtype),
              make_build(3, 0, dtype), make_build(4, big, dtype), make_build(5, big - 8192, dtype)]
    assert fij.plan_bits(dtype, 300000, 1 << 20)[1] == 0 and fij.plan_bits(dtype, big - 8192, 1 << 20)[1] == 0
    assert fij.plan_bits(dtype, big, 1 << 20)[1] > 0
    srv = fij.JoinServer(dtype, [(3, 1 << 16), (2, 1 << 20)], max_probe_len=1 << 20, table_build_max=table_build_max)
    dev_builds = [tuple(to_dev(a) if len(a) else fij.DeviceArray(1, dtype).view(0, 0) for a in b) for b in builds]
    ids = [srv.add_build(*b) for b in dev_builds]
    del dev_builds  # resident: the input columns are no longer needed
    rng = np.random.default_rng(11)
    reqs = []
    for j in range(24):
        b = int(rng.integers(0, len(builds)))
        n = 0 if j == 5 else int(2 ** rng.uniform(10, 20))
        n_b = max(len(builds[b][0]), 1)
        host = make_probe(j, n_b, n, dtype)
        dev = tuple(to_dev(a) if len(a) else fij.DeviceArray(1, dtype).view(0, 0) for a in host)
        reqs.append((b, host, dev))
    want = [orc.sort_merge_join(*builds[b], *host) for b, host, _ in reqs]
    alone = []
    for b, _, dev in reqs:
        alone.append(srv.wait(srv.submit(ids[b], *dev))[0])
    rids = [srv.submit(ids[b], *dev) for b, _, dev in reqs]
    mix = {}
    for k in reversed(range(len(rids))):
        mix[k], t = srv.wait(rids[k])
        assert t["latency_ms"] >= t["queued_ms"] >= 0
    for k in range(len(reqs)):
        assert alone[k] == want[k], f"request {k} alone"
        assert mix[k] == want[k], f"request {k} in the mix"
    used, nb = srv.memory()
    assert nb == len(builds) and used > 0
    srv.close()
    assert fij.device_status() == 0


@pytest.mark.parametrize("table_build_max", [0, 1 << 16])
def test_server_concurrent_submitters(table_build_max):
    dtype = np.int64
    build = make_build(5, 50000, dtype)
    srv = fij.JoinServer(dtype, [(4, 1 << 18)], max_probe_len=1 << 18, table_build_max=table_build_max)
    bid = srv.add_build(*(to_dev(a) for a in build))
    probes = [make_probe(40 + j, 50000, int(2 ** (12 + j % 6)), dtype) for j in range(16)]
    devs = [tuple(to_dev(a) for a in p) for p in probes]
    got = [None] * len(probes)

    def client(ks):
        for k in ks:
            got[k] = srv.wait(srv.submit(bid, *devs[k]))[0]

    threads = [threading.Thread(target=client, args=(range(t, len(probes), 4),)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for k, p in enumerate(probes):
        assert got[k] == orc.sort_merge_join(*build, *p), f"request {k}"
    srv.close()


def test_server_matches_plain_plan():
    dtype = np.int64
    rk, rp = make_build(7, 1 << 20, dtype)
    sk, sp = make_probe(9, 1 << 20, 1 << 22, dtype)
    d = [to_dev(a) for a in (rk, rp, sk, sp)]
    srv = fij.JoinServer(dtype, [(1, 1 << 22)], max_probe_len=1 << 22)
    bid = srv.add_build(d[0], d[1])
    got, _ = srv.wait(srv.submit(bid, d[2], d[3]))
    _, b1, b2 = fij.autotune(dtype, 1 << 20, 1 << 22)
    want, _ = fij.JoinPlan(dtype, 1 << 20, 1 << 22, b1, b2).execute(*d)
    assert got == want
    srv.close()


def test_server_limits():
    dtype = np.int32
    with pytest.raises(fij.FijError):  # the slots alone exceed the budget
        fij.JoinServer(dtype, [(1, 1 << 20)], max_probe_len=1 << 20, hbm_budget=1 << 20)
    srv = fij.JoinServer(dtype, [(1, 1 << 12)], max_probe_len=1 << 12)
    bid = srv.add_build(*(to_dev(a) for a in make_build(8, 1000, dtype)))
    big = make_probe(1, 1000, 1 << 13, dtype)
    with pytest.raises(fij.FijError):  # larger than every slot
        srv.submit(bid, *(to_dev(a) for a in big))
    with pytest.raises(fij.FijError):
        srv.submit(bid + 1, *(to_dev(a) for a in big[:1] * 2))
    srv.close()
