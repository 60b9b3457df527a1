"""The calibration kernels that bench.py sets beside the scatter and the join
(fij_calib_line_scatter, fij_calib_run_gather): each is checked against a
numpy model of the access pattern it claims to have, so that the times
measured with them are times of that pattern."""
import numpy as np
import pytest

import fij

pytestmark = pytest.mark.gpu


def line_scatter_model(src, parts, chunks):
    """dst of the line-scatter copy: word r of chunk c goes to line r // 8
    of the chunk's line sequence, line l to partition l mod P, its (l / P)-th
    line in the chunk's piece of that partition."""
    w = len(src)
    chunk = w // chunks
    piece = chunk // parts
    region = piece * chunks
    i = np.arange(w, dtype=np.int64)
    c, r = i // chunk, i % chunk
    line = r // 8
    dst = np.empty_like(src)
    dst[(line % parts) * region + c * piece + (line // parts) * 8 + r % 8] = src
    return dst


@pytest.mark.parametrize("parts, chunks", [(512, 4), (64, 8), (1, 2)])
def test_line_scatter_is_the_scatter_layout(parts, chunks):
    rng = np.random.default_rng(3)
    words = chunks * 8192 * 2
    src = rng.integers(0, 2**63, size=(words, 2), dtype=np.int64)
    d_src = fij.DeviceArray.from_host(src.reshape(-1))
    d_dst = fij.DeviceArray.zeros(2 * words, np.int64)
    fij.calib_line_scatter(d_dst, d_src, parts, chunks)
    fij.synchronize()
    got = d_dst.to_host().reshape(words, 2)
    want = line_scatter_model(src, parts, chunks)
    assert np.array_equal(got, want)


def run_gather_model(words_u32, part, f2, run, shift):
    """xor of every 32-bit word the run gather loads (lanes past an item's
    end load its last word again)."""
    w = words_u32.reshape(-1, 4)
    fold = w[:, 0] ^ w[:, 1] ^ w[:, 2] ^ w[:, 3]
    item = part // f2
    loads = (item + 63) // 64
    e = np.minimum(np.arange(loads * 64), item - 1)
    acc = np.uint32(0)
    for p in range(len(w) // part):
        for q in range(f2):
            sh = shift if q + 1 < f2 else 0
            at = p * part + (e // run) * (f2 * run) + q * run + e % run + sh
            acc ^= np.bitwise_xor.reduce(fold[at])
    return int(acc)


@pytest.mark.parametrize("run, shift", [(16, 0), (16, 1), (32, 4), (8, 8)])
def test_run_gather_reads_the_runs(run, shift):
    rng = np.random.default_rng(5)
    part, f2 = 8192, 32  # items of 256 words, 4 loads per lane
    words = 4 * part
    src = rng.integers(0, 2**32, size=words * 4, dtype=np.uint32)
    d_src = fij.DeviceArray.from_host(src)
    ck = fij.DeviceArray.zeros(1, np.uint32)
    fij.calib_run_gather(d_src, part, f2, run, shift, ck, grid=8)
    fij.synchronize()
    assert int(ck.to_host()[0]) == run_gather_model(src, part, f2, run, shift)
