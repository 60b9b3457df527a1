"""Relation files (fij_relation_file_*): the text format of the reference's
generator and join data builder (datagen/src/bin/generate.rs:343-365,
data-store/src/join_data.rs:173-280).  Host code: these run without a GPU.

Files are written here independently of the library (Python's own text and
gzip writers, in the generator's layout: a `key value` header and value =
row + 1) and read back through the C ABI; the library's writer is checked by
reading its output with numpy."""
import gzip

import numpy as np
import pytest

import fij


def write_text(path, rows, header="key value", newline="\n", gz=False):
    text = header + newline + "".join(f"{k} {v}{newline}" for k, v in rows)
    if gz:
        with gzip.open(path, "wt", newline="") as f:
            f.write(text)
    else:
        with open(path, "w", newline="") as f:
            f.write(text)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("gz", [False, True])
def test_reads_generator_layout(tmp_path, dtype, gz):
    rng = np.random.default_rng(3)
    info = np.iinfo(dtype)
    keys = rng.integers(info.min, info.max, 5000, dtype=dtype, endpoint=True)
    keys[:4] = [info.min, info.max, 0, -1]
    rows = [(int(k), i + 1) for i, k in enumerate(keys)]  # value = row + 1, as the generator writes
    path = tmp_path / ("r.tsv.gz" if gz else "r.tsv")
    write_text(path, rows, gz=gz)
    assert fij.relation_file_len(path) == len(rows)
    k, v = fij.read_relation_file(path, dtype)
    assert k.dtype == dtype and v.dtype == dtype
    np.testing.assert_array_equal(k, keys)
    np.testing.assert_array_equal(v, np.arange(1, len(rows) + 1, dtype=dtype))


def test_crlf_blank_lines_and_no_final_newline(tmp_path):
    path = tmp_path / "r.tsv"
    with open(path, "w", newline="") as f:
        f.write("key value\r\n1 10\r\n\r\n\n-2 +20\r\n3 30")
    k, v = fij.read_relation_file(path, np.int64)
    assert k.tolist() == [1, -2, 3] and v.tolist() == [10, 20, 30]


def test_empty_and_header_only(tmp_path):
    for text in ["", "key value\n"]:
        path = tmp_path / "e.tsv"
        path.write_text(text)
        k, v = fij.read_relation_file(path, np.int32)
        assert len(k) == 0 and len(v) == 0


@pytest.mark.parametrize("line", ["1 x", "1  2", "1 2 3", "1", " 1 2", "1 2 ", "2147483648 1", "-2147483649 1",
                                  "1.0 2", "+ 1", "1 -"])
def test_malformed_records_are_parse_errors(tmp_path, line):
    path = tmp_path / "bad.tsv"
    path.write_text(f"key value\n1 1\n{line}\n")
    with pytest.raises(fij.FijError) as e:
        fij.read_relation_file(path, np.int32)
    assert e.value.status == fij.Status.PARSE_ERROR
    assert ":3:" in str(e.value)  # the line number


def test_i64_range_and_header_fields(tmp_path):
    path = tmp_path / "r.tsv"
    path.write_text("key value\n9223372036854775807 -9223372036854775808\n")
    k, v = fij.read_relation_file(path, np.int64)
    assert k[0] == np.iinfo(np.int64).max and v[0] == np.iinfo(np.int64).min
    path.write_text("key value\n9223372036854775808 1\n")
    with pytest.raises(fij.FijError):
        fij.read_relation_file(path, np.int64)
    path.write_text("keyvalue\n1 2\n")
    with pytest.raises(fij.FijError) as e:
        fij.read_relation_file(path, np.int64)
    assert e.value.status == fij.Status.PARSE_ERROR


def test_missing_file_is_io_error(tmp_path):
    with pytest.raises(fij.FijError) as e:
        fij.read_relation_file(tmp_path / "nope.tsv", np.int64)
    assert e.value.status == fij.Status.IO_ERROR


def test_capacity_is_checked(tmp_path):
    path = tmp_path / "r.tsv"
    write_text(path, [(i, i) for i in range(10)])
    keys = np.empty(4, np.int64)
    pays = np.empty(4, np.int64)
    n = fij._u64()
    st = fij.lib.fij_relation_file_read(str(path).encode(), b"\0", int(fij.KeyType.I64),
                                        keys.ctypes.data_as(fij._vp), pays.ctypes.data_as(fij._vp), 4,
                                        fij.C.byref(n))
    assert st == fij.Status.INVALID_ARGUMENT


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("name,delim", [("w.tsv", None), ("w.csv", ","), ("w.tsv.gz", None)])
def test_writer_round_trip(tmp_path, dtype, name, delim):
    rng = np.random.default_rng(5)
    info = np.iinfo(dtype)
    keys = rng.integers(info.min, info.max, 20000, dtype=dtype, endpoint=True)
    pays = np.arange(1, 20001, dtype=dtype)
    path = tmp_path / name
    fij.write_relation_file(path, keys, pays, delimiter=delim)
    opener = gzip.open if name.endswith("gz") else open
    with opener(path, "rt") as f:
        lines = f.read().splitlines()
    sep = delim or " "
    assert lines[0] == f"key{sep}value"
    parsed = np.array([[int(x) for x in l.split(sep)] for l in lines[1:]], dtype=dtype)
    np.testing.assert_array_equal(parsed[:, 0], keys)
    np.testing.assert_array_equal(parsed[:, 1], pays)
    k, v = fij.read_relation_file(path, dtype, delimiter=delim)
    np.testing.assert_array_equal(k, keys)
    np.testing.assert_array_equal(v, pays)
