"""Python mirror of the reference's sql-ops operator API over libfij.so.

The classes follow sql-ops' Rust types (RadixBits, PartitionOffsets,
PartitionedRelation, GpuRadixPartitioner, CudaRadixJoin, CudaHashJoin,
HashTable) and call the C ABI declared in include/fij.h through ctypes.
Device memory, copies and streams also go through the library (DeviceArray,
Stream, Event), so a process holds exactly one HIP runtime.  If libfij.so is
missing, importing this module raises, so no caller can silently fall back
to another implementation.
"""
import ctypes as C
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libfij.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python fast-interconnects_amd/build.py`"
    )
lib = C.CDLL(LIB_PATH)


# --------------------------------------------------------------------------
# enums (include/fij.h)


class Status(enum.IntEnum):
    OK = 0
    INVALID_ARGUMENT = 1
    INTEGER_OVERFLOW = 2
    DEVICE_ERROR = 3
    RUNTIME_ERROR = 4
    UNSUPPORTED = 5
    OUT_OF_MEMORY = 6
    IO_ERROR = 7
    PARSE_ERROR = 8


class KeyType(enum.IntEnum):
    I32 = 0
    I64 = 1


class HistogramAlgorithm(enum.IntEnum):
    Chunked = 0
    Contiguous = 1


class RadixPartitionAlgorithm(enum.IntEnum):
    NC = 0
    LASWWC = 1
    SSWWCv2 = 2
    SSWWCv2G = 3
    HSSWWCv4 = 4


class RadixPass(enum.IntEnum):
    First = 0
    Second = 1
    Third = 2


class HashingScheme(enum.IntEnum):
    Perfect = 0
    LinearProbing = 1
    BucketChaining = 2


class FijError(RuntimeError):
    def __init__(self, status, message):
        self.status = Status(status)
        super().__init__(f"{self.status.name}: {message}")


# --------------------------------------------------------------------------
# C structs


class _RadixBits(C.Structure):
    _fields_ = [("pass_bits", C.c_int32 * 3)]


class _PartitionOffsets(C.Structure):
    _fields_ = [
        ("offsets", C.c_void_p),
        ("local_offsets", C.c_void_p),
        ("data_len", C.c_uint64),
        ("chunks", C.c_uint32),
        ("max_chunks", C.c_uint32),
        ("radix_bits", C.c_uint32),
        ("type", C.c_uint32),
    ]


class _PartitionedRelation(C.Structure):
    _fields_ = [
        ("relation", C.c_void_p),
        ("offsets", C.c_void_p),
        ("relation_len", C.c_uint64),
        ("len", C.c_uint64),
        ("chunks", C.c_uint32),
        ("radix_bits", C.c_uint32),
        ("type", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class JoinAggregate(C.Structure):
    _fields_ = [("count", C.c_uint64), ("sum_build", C.c_uint64), ("sum_probe", C.c_uint64)]


class _DeviceLimits(C.Structure):
    _fields_ = [
        ("compute_units", C.c_uint32),
        ("wave_size", C.c_uint32),
        ("lds_bytes_per_block", C.c_uint32),
        ("max_threads_per_block", C.c_uint32),
        ("global_mem_bytes", C.c_uint64),
        ("arch_name", C.c_char * 64),
    ]


LEN_UNSET = 0xFFFFFFFFFFFFFFFF

_u32, _u64, _i32, _i64, _vp = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_void_p


def _proto(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_proto("fij_last_error", C.c_char_p)
_proto("fij_version", C.c_char_p)
_proto("fij_device_limits_get", C.c_int, C.c_int, C.POINTER(_DeviceLimits))
_proto("fij_device_status", C.c_int, _vp, C.c_int, C.POINTER(_u32))
_proto("fij_padding_len", _u32, C.c_int)
_proto("fij_input_chunk_size", C.c_int, C.c_int, _u64, _u32, C.POINTER(_u64))
_proto("fij_partition_offsets_layout", C.c_int, C.c_int, _u32, _u32, C.POINTER(_u32), C.POINTER(_u64), C.POINTER(_u64))
_proto("fij_partitioned_relation_layout", C.c_int, C.c_int, _u64, C.c_int, _u32, _u32, C.POINTER(_u32), C.POINTER(_u64), C.POINTER(_u64))
_proto("fij_partitioned_relation_padded_len", _u64, C.POINTER(_PartitionedRelation))
_proto("fij_partitioned_relation_capacity", _u64, C.POINTER(_PartitionedRelation))
_proto("fij_partitioned_relation_resize", C.c_int, C.POINTER(_PartitionedRelation), _u64)
_proto("fij_partitioned_relation_partition_len", C.c_int, C.POINTER(_PartitionedRelation), _vp, _u32, C.POINTER(_u64))
_proto("fij_partition_offsets_partition_len", C.c_int, C.POINTER(_PartitionOffsets), _vp, _u32, C.POINTER(_u64))
_proto("fij_partitioner_create", C.c_int, C.c_int, C.c_int, _RadixBits, _u32, _u32, _u64, C.POINTER(_vp))
_proto("fij_partitioner_destroy", C.c_int, _vp)
_proto("fij_prefix_sum", C.c_int, _vp, C.c_int, C.c_int, _vp, _u64, C.POINTER(_PartitionOffsets), _vp)
_proto("fij_prefix_sum_noalloc", C.c_int, _vp, C.c_int, C.c_int, _vp, _u64, C.POINTER(_PartitionOffsets), _vp)
_proto("fij_prefix_sum_and_copy_with_payload", C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _u64, _vp, _vp, _u64, C.POINTER(_PartitionOffsets), _vp)
_proto("fij_prefix_sum_and_transform", C.c_int, _vp, C.c_int, C.c_int, _u32, C.POINTER(_PartitionedRelation), _vp, _vp, _u64, C.POINTER(_PartitionOffsets), _vp)
_proto("fij_preallocate_partition_state", C.c_int, _vp, C.c_int, C.c_int)
_proto("fij_partition", C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _u64, C.POINTER(_PartitionOffsets), C.POINTER(_PartitionedRelation), _vp)
_proto("fij_radix_join_create", C.c_int, C.c_int, _RadixBits, C.c_int, _u32, _u32, C.POINTER(_vp))
_proto("fij_radix_join_destroy", C.c_int, _vp)
_proto("fij_radix_join_join", C.c_int, _vp, C.c_int, C.POINTER(_PartitionedRelation), C.POINTER(_PartitionedRelation), _vp, _u64, _vp, _u64, _vp, _vp)
_proto("fij_join_assign_tasks", C.c_int, C.POINTER(_PartitionedRelation), _u32, _u32, _vp, _vp)
_proto("fij_np_table_init", C.c_int, C.c_int, _vp, _u64, _vp)
_proto("fij_relation_fill_null", C.c_int, C.c_int, _vp, _u64, _vp)
_proto("fij_np_build", C.c_int, C.c_int, C.c_int, C.c_int, _vp, _u64, _vp, _vp, _u64, _u32, _u32, _vp)
_proto("fij_np_probe_sum", C.c_int, C.c_int, C.c_int, _vp, _u64, _vp, _vp, _u64, _vp, _u64, _vp, _u32, _u32, _vp)
_proto("fij_gen_permutation", C.c_int, C.c_int, _vp, _u64, _u64, _i64, _vp)
_proto("fij_gen_uniform", C.c_int, C.c_int, _vp, _u64, _u64, _u64, _i64, _u64, _vp)
_proto("fij_gen_iota", C.c_int, C.c_int, _vp, _u64, _i64, _vp)
_proto("fij_gen_zipf", C.c_int, C.c_int, _vp, _u64, _u64, C.c_double, _u64, _u64, _u64, _i64, _vp)


def check(status):
    if status != 0:
        raise FijError(status, lib.fij_last_error().decode())


def version():
    return lib.fij_version().decode()


# --------------------------------------------------------------------------
# runtime plumbing: device buffers and streams through the C ABI

_proto("fij_device_count", C.c_int, C.POINTER(C.c_int))
_proto("fij_mem_alloc", C.c_int, _u64, C.POINTER(_vp))
_proto("fij_mem_free", C.c_int, _vp)
_proto("fij_mem_alloc_host", C.c_int, _u64, C.POINTER(_vp))
_proto("fij_mem_free_host", C.c_int, _vp)
_proto("fij_memcpy", C.c_int, _vp, _vp, _u64, _vp)
_proto("fij_memcpy_async", C.c_int, _vp, _vp, _u64, _vp)
_proto("fij_mem_fill_u32", C.c_int, _vp, _u32, _u64, _vp)
_proto("fij_mem_fill_u64", C.c_int, _vp, _u64, _u64, _vp)
_proto("fij_copy_kernel", C.c_int, _vp, _vp, _u64, _vp)
_proto("fij_calib_line_scatter", C.c_int, _vp, _vp, _u64, _u32, _u32, _vp)
_proto("fij_calib_run_gather", C.c_int, _vp, _u64, _u32, _u32, _u32, _u32, _vp, _u32, _vp)
_proto("fij_stream_create", C.c_int, C.POINTER(_vp))
_proto("fij_stream_destroy", C.c_int, _vp)
_proto("fij_stream_synchronize", C.c_int, _vp)
_proto("fij_device_synchronize", C.c_int)
_proto("fij_event_create", C.c_int, C.POINTER(_vp))
_proto("fij_event_destroy", C.c_int, _vp)
_proto("fij_event_record", C.c_int, _vp, _vp)
_proto("fij_event_elapsed_ms", C.c_int, _vp, _vp, C.POINTER(C.c_float))


def device_count():
    c = C.c_int(0)
    check(lib.fij_device_count(C.byref(c)))
    return c.value


def synchronize():
    check(lib.fij_device_synchronize())


def copy_kernel(dst, src, nbytes=None, stream=None):
    """The library's plain device copy kernel (bandwidth reference)."""
    nbytes = src.nbytes() if nbytes is None else nbytes
    check(lib.fij_copy_kernel(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, _stream(stream)))


def calib_line_scatter(dst, src, partitions, chunks, nbytes=None, stream=None):
    """Calibration: the pass-1 scatter's line-write pattern as a plain copy."""
    nbytes = src.nbytes() if nbytes is None else nbytes
    check(lib.fij_calib_line_scatter(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, partitions, chunks,
                                     _stream(stream)))


def calib_run_gather(src, part_words, f2, run_words, shift, checksum, grid=512, nbytes=None, stream=None):
    """Calibration: the register join's run gathers without table work."""
    nbytes = src.nbytes() if nbytes is None else nbytes
    check(lib.fij_calib_run_gather(C.c_void_p(src.ptr), nbytes, part_words, f2, run_words, shift,
                                   C.c_void_p(checksum.ptr), grid, _stream(stream)))


class DeviceArray:
    """A typed device buffer (the role of rustacuda's DeviceBuffer)."""

    def __init__(self, n, dtype, _ptr=None, _owner=None):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self._owner = _owner
        if _ptr is not None:
            self.ptr = _ptr
        else:
            p = _vp()
            check(lib.fij_mem_alloc(max(self.n, 1) * self.dtype.itemsize, C.byref(p)))
            self.ptr = p.value
            self._owner = None

    def __del__(self):
        if self._owner is None and getattr(self, "ptr", None):
            lib.fij_mem_free(C.c_void_p(self.ptr))
            self.ptr = None

    @classmethod
    def empty(cls, n, dtype):
        return cls(n, dtype)

    @classmethod
    def zeros(cls, n, dtype):
        a = cls(n, dtype)
        a.fill(0)
        return a

    @classmethod
    def from_host(cls, host, stream=None):
        host = np.ascontiguousarray(host)
        a = cls(host.size, host.dtype)
        a.copy_from_host(host, stream)
        return a

    def numel(self):
        return self.n

    def __len__(self):
        return self.n

    def nbytes(self):
        return self.n * self.dtype.itemsize

    def data_ptr(self):
        return self.ptr

    def view(self, start, stop):
        """Contiguous sub-range [start, stop) sharing this buffer."""
        start, stop = int(start), int(stop)
        assert 0 <= start <= stop <= self.n
        return DeviceArray(stop - start, self.dtype, _ptr=self.ptr + start * self.dtype.itemsize,
                           _owner=self if self._owner is None else self._owner)

    def copy_from_host(self, host, stream=None):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size <= self.n
        check(lib.fij_memcpy(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), host.nbytes, _stream(stream)))

    def to_host(self, stream=None):
        out = np.empty(self.n, self.dtype)
        check(lib.fij_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, _stream(stream)))
        return out

    def fill(self, value, stream=None):
        if self.dtype.itemsize == 4:
            v = int(np.array(value, self.dtype).view(np.uint32))
            check(lib.fij_mem_fill_u32(C.c_void_p(self.ptr), v, self.n, _stream(stream)))
        elif self.dtype.itemsize == 8:
            v = int(np.array(value, self.dtype).view(np.uint64))
            check(lib.fij_mem_fill_u64(C.c_void_p(self.ptr), v, self.n, _stream(stream)))
        else:
            raise ValueError("fill supports 4- and 8-byte elements")


class Stream:
    def __init__(self):
        h = _vp()
        check(lib.fij_stream_create(C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_stream_destroy(C.c_void_p(self.h))
            self.h = None

    def synchronize(self):
        check(lib.fij_stream_synchronize(C.c_void_p(self.h)))


class Event:
    def __init__(self):
        h = _vp()
        check(lib.fij_event_create(C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_event_destroy(C.c_void_p(self.h))
            self.h = None

    def record(self, stream=None):
        check(lib.fij_event_record(C.c_void_p(self.h), _stream(stream)))

    def elapsed_ms(self, start):
        ms = C.c_float()
        check(lib.fij_event_elapsed_ms(C.c_void_p(start.h), C.c_void_p(self.h), C.byref(ms)))
        return ms.value


def _ktype(dtype):
    if isinstance(dtype, KeyType):
        return dtype
    if dtype in ("i32",):
        return KeyType.I32
    if dtype in ("i64",):
        return KeyType.I64
    d = np.dtype(dtype)
    if d == np.int32:
        return KeyType.I32
    if d == np.int64:
        return KeyType.I64
    raise ValueError(f"unsupported key type {dtype}")


def numpy_dtype(ktype):
    return np.int32 if ktype == KeyType.I32 else np.int64


def _ptr(x):
    if x is None:
        return None
    return C.c_void_p(x.data_ptr())


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.h)


def device_limits(device=0):
    out = _DeviceLimits()
    check(lib.fij_device_limits_get(device, C.byref(out)))
    return {
        "compute_units": out.compute_units,
        "wave_size": out.wave_size,
        "lds_bytes_per_block": out.lds_bytes_per_block,
        "max_threads_per_block": out.max_threads_per_block,
        "global_mem_bytes": out.global_mem_bytes,
        "arch_name": out.arch_name.decode(),
    }


def device_status(stream=None, clear=True):
    v = _u32(0)
    check(lib.fij_device_status(_stream(stream), int(clear), C.byref(v)))
    return v.value


def padding_len(ktype):
    return lib.fij_padding_len(int(ktype))


def input_chunk_size(ktype, data_len, num_chunks):
    out = _u64(0)
    check(lib.fij_input_chunk_size(int(ktype), data_len, num_chunks, C.byref(out)))
    return out.value


# --------------------------------------------------------------------------
# RadixBits (sql-ops/src/partition.rs:62-120)


class RadixBits:
    def __init__(self, first=None, second=None, third=None):
        self.pass_bits = [first, second, third]

    def c(self):
        s = _RadixBits()
        for i, b in enumerate(self.pass_bits):
            s.pass_bits[i] = -1 if b is None else int(b)
        return s

    def pass_radix_bits(self, p):
        return self.pass_bits[int(p)]

    def pass_fanout(self, p):
        b = self.pass_bits[int(p)]
        return None if b is None else 1 << b

    def pass_ignore_bits(self, p):
        return sum(b or 0 for b in self.pass_bits[: int(p)])

    def radix_bits(self):
        return sum(b or 0 for b in self.pass_bits)


# --------------------------------------------------------------------------
# PartitionOffsets / PartitionedRelation (partitioned_relation.rs)


class PartitionOffsets:
    """Device-resident partition offsets.  offsets: [chunks * F] u64;
    local_offsets: [max_chunks * F] for the Contiguous algorithm."""

    def __init__(self, histogram_algorithm, max_chunks, radix_bits, ktype):
        self.ktype = _ktype(ktype)
        chunks, olen, llen = _u32(), _u64(), _u64()
        check(lib.fij_partition_offsets_layout(int(histogram_algorithm), max_chunks, radix_bits,
                                               C.byref(chunks), C.byref(olen), C.byref(llen)))
        self.offsets = DeviceArray.zeros(olen.value, np.uint64)
        self.local_offsets = DeviceArray.zeros(llen.value, np.uint64) if llen.value else None
        self.s = _PartitionOffsets()
        self.s.offsets = self.offsets.data_ptr()
        self.s.local_offsets = self.local_offsets.data_ptr() if self.local_offsets is not None else None
        self.s.data_len = LEN_UNSET
        self.s.chunks = chunks.value
        self.s.max_chunks = max_chunks
        self.s.radix_bits = radix_bits
        self.s.type = int(self.ktype)

    def len(self):
        return None if self.s.data_len == LEN_UNSET else self.s.data_len

    def num_chunks(self):
        return self.s.chunks

    def radix_bits(self):
        return self.s.radix_bits

    def fanout(self):
        return 1 << self.s.radix_bits

    def host_offsets(self):
        return self.offsets.to_host()

    def host_local_offsets(self):
        return None if self.local_offsets is None else self.local_offsets.to_host()

    def partition_len(self, pid, host_offsets=None):
        h = self.host_offsets() if host_offsets is None else host_offsets
        out = _u64()
        check(lib.fij_partition_offsets_partition_len(C.byref(self.s), h.ctypes.data_as(C.c_void_p), pid, C.byref(out)))
        return out.value


class PartitionedRelation:
    """Row-layout partitioned relation {key, value}[relation_len] in device memory."""

    def __init__(self, length, histogram_algorithm, radix_bits, max_chunks, ktype, fill_null=False):
        self.ktype = _ktype(ktype)
        chunks, rlen, olen = _u32(), _u64(), _u64()
        check(lib.fij_partitioned_relation_layout(int(self.ktype), length, int(histogram_algorithm), radix_bits,
                                                  max_chunks, C.byref(chunks), C.byref(rlen), C.byref(olen)))
        self.relation = DeviceArray(rlen.value * 2, numpy_dtype(self.ktype))
        self.offsets = DeviceArray.zeros(olen.value, np.uint64)
        self.s = _PartitionedRelation()
        self.s.relation = self.relation.data_ptr()
        self.s.offsets = self.offsets.data_ptr()
        self.s.relation_len = rlen.value
        self.s.len = length
        self.s.chunks = chunks.value
        self.s.radix_bits = radix_bits
        self.s.type = int(self.ktype)
        if fill_null:
            self.fill_null()

    def len(self):
        return self.s.len

    def padded_len(self):
        return lib.fij_partitioned_relation_padded_len(C.byref(self.s))

    def capacity(self):
        return lib.fij_partitioned_relation_capacity(C.byref(self.s))

    def resize(self, new_len):
        check(lib.fij_partitioned_relation_resize(C.byref(self.s), new_len))

    def num_chunks(self):
        return self.s.chunks

    def radix_bits(self):
        return self.s.radix_bits

    def fanout(self):
        return 1 << self.s.radix_bits

    def padding_len(self):
        return padding_len(self.ktype)

    def host_offsets(self):
        return self.offsets.to_host()

    def host_relation(self):
        """(keys, values) numpy views of the whole relation, padding included."""
        a = self.relation.to_host()
        return a[0::2], a[1::2]

    def fill_null(self, stream=None):
        """Sets every tuple to {null key, 0} (as the reference's tests do)."""
        check(lib.fij_relation_fill_null(int(self.ktype), _ptr(self.relation), self.s.relation_len, _stream(stream)))

    def partition_len(self, pid, host_offsets=None):
        h = self.host_offsets() if host_offsets is None else host_offsets
        out = _u64()
        check(lib.fij_partitioned_relation_partition_len(C.byref(self.s), h.ctypes.data_as(C.c_void_p), pid, C.byref(out)))
        return out.value

    def partition_range(self, chunk, pid, host_offsets=None):
        """Index<(chunk, partition)> (partitioned_relation.rs:523-545) as a [begin, end) range."""
        h = self.host_offsets() if host_offsets is None else host_offsets
        F = self.fanout()
        ofi = chunk * F + pid
        begin = int(h[ofi])
        end = int(h[ofi + 1]) - self.padding_len() if ofi + 1 < len(h) else self.padded_len()
        return begin, end


# --------------------------------------------------------------------------
# GpuRadixPartitioner (gpu_radix_partition.rs:314-555)


class GpuRadixPartitioner:
    def __init__(self, histogram_algorithm, partition_algorithm, radix_bits, grid_size, block_size,
                 dmem_buffer_bytes=8 * 1024):
        if not isinstance(radix_bits, RadixBits):
            radix_bits = RadixBits(radix_bits)
        self.radix_bits = radix_bits
        self.histogram_algorithm = HistogramAlgorithm(histogram_algorithm)
        self.partition_algorithm = RadixPartitionAlgorithm(partition_algorithm)
        self.grid_size = grid_size
        self.block_size = block_size
        h = _vp()
        check(lib.fij_partitioner_create(int(histogram_algorithm), int(partition_algorithm), radix_bits.c(),
                                         grid_size, block_size, dmem_buffer_bytes, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_partitioner_destroy(self.h)
            self.h = None

    def prefix_sum(self, radix_pass, partition_attr, partition_offsets, stream=None):
        kt = _ktype(partition_attr.dtype)
        check(lib.fij_prefix_sum(self.h, int(radix_pass), int(kt), _ptr(partition_attr), partition_attr.numel(),
                                 C.byref(partition_offsets.s), _stream(stream)))

    def prefix_sum_noalloc(self, radix_pass, partition_attr, partition_offsets, stream=None):
        """prefix_sum without device allocation (fij_prefix_sum_noalloc): the
        contiguous histogram's scratch is the library's static per-GPU array,
        shared by every such call on that GPU."""
        kt = _ktype(partition_attr.dtype)
        check(lib.fij_prefix_sum_noalloc(self.h, int(radix_pass), int(kt), _ptr(partition_attr),
                                         partition_attr.numel(), C.byref(partition_offsets.s), _stream(stream)))

    def prefix_sum_and_copy_with_payload(self, radix_pass, src_key, src_pay, dst_key, dst_pay, partition_offsets,
                                         stream=None):
        kt = _ktype(src_key.dtype)
        check(lib.fij_prefix_sum_and_copy_with_payload(
            self.h, int(radix_pass), int(kt), _ptr(src_key), _ptr(src_pay), src_key.numel(), _ptr(dst_key),
            _ptr(dst_pay), dst_key.numel(), C.byref(partition_offsets.s), _stream(stream)))

    def prefix_sum_and_transform(self, radix_pass, partition_id, src_relation, dst_key, dst_pay, partition_offsets,
                                 stream=None):
        kt = _ktype(dst_key.dtype)
        check(lib.fij_prefix_sum_and_transform(
            self.h, int(radix_pass), int(kt), partition_id, C.byref(src_relation.s), _ptr(dst_key), _ptr(dst_pay),
            dst_key.numel(), C.byref(partition_offsets.s), _stream(stream)))

    def preallocate_partition_state(self, radix_pass, ktype):
        check(lib.fij_preallocate_partition_state(self.h, int(radix_pass), int(_ktype(ktype))))

    def partition(self, radix_pass, partition_attr, payload_attr, partition_offsets, partitioned_relation,
                  stream=None):
        if partition_attr.numel() != payload_attr.numel():
            raise FijError(Status.INVALID_ARGUMENT, "Partition and payload attributes have different sizes")
        kt = _ktype(partition_attr.dtype)
        check(lib.fij_partition(self.h, int(radix_pass), int(kt), _ptr(partition_attr), _ptr(payload_attr),
                                partition_attr.numel(), C.byref(partition_offsets.s),
                                C.byref(partitioned_relation.s), _stream(stream)))
        # The C call swapped the offset pointers; swap the owning buffers too.
        partition_offsets.offsets, partitioned_relation.offsets = (
            partitioned_relation.offsets, partition_offsets.offsets)
        assert partitioned_relation.s.offsets == partitioned_relation.offsets.data_ptr()


# --------------------------------------------------------------------------
# CudaRadixJoin (sql-ops/src/join/cuda_radix_join.rs:105-164)


class HipRadixJoin:
    def __init__(self, radix_pass, radix_bits, hashing_scheme, grid_size, block_size):
        if not isinstance(radix_bits, RadixBits):
            radix_bits = RadixBits(radix_bits)
        self.grid_size = grid_size
        self.block_size = block_size
        h = _vp()
        check(lib.fij_radix_join_create(int(radix_pass), radix_bits.c(), int(hashing_scheme), grid_size, block_size,
                                        C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_radix_join_destroy(self.h)
            self.h = None

    def join(self, build_rel, probe_rel, result_set, task_assignments, aggregates=None, stream=None):
        check(lib.fij_radix_join_join(
            self.h, int(build_rel.ktype), C.byref(build_rel.s), C.byref(probe_rel.s), _ptr(result_set),
            result_set.numel() if result_set is not None else 0, _ptr(task_assignments), task_assignments.numel(),
            _ptr(aggregates), _stream(stream)))


CudaRadixJoin = HipRadixJoin


def assign_tasks(probe_rel, radix_bits, grid_size, task_assignments, stream=None):
    check(lib.fij_join_assign_tasks(C.byref(probe_rel.s), radix_bits, grid_size, _ptr(task_assignments),
                                    _stream(stream)))


def aggregates_array(n):
    """Zeroed device array of n JoinAggregate records (3 x u64 each)."""
    return DeviceArray.zeros(3 * n, np.uint64)


def sum_aggregates(agg):
    a = agg.to_host().view(np.uint64).reshape(-1, 3)
    with np.errstate(over="ignore"):
        s = a.sum(axis=0, dtype=np.uint64)
    return int(s[0]), int(s[1]), int(s[2])


# --------------------------------------------------------------------------
# No-partitioning join (sql-ops/src/join/no_partitioning_join.rs)


class HashTable:
    """HashTable::new_on_gpu: `entries` {key, value} slots set to the null key."""

    def __init__(self, entries, ktype, stream=None):
        self.ktype = _ktype(ktype)
        self.entries = entries
        self.mem = DeviceArray(2 * entries, numpy_dtype(self.ktype))
        check(lib.fij_np_table_init(int(self.ktype), _ptr(self.mem), entries, _stream(stream)))


class HipHashJoin:
    def __init__(self, hashing_scheme, hash_table, build_dim=(1, 1024), probe_dim=(1, 1024), is_selective=False):
        self.scheme = HashingScheme(hashing_scheme)
        self.ht = hash_table
        self.build_dim = build_dim
        self.probe_dim = probe_dim
        self.is_selective = is_selective

    def build(self, keys, pays, stream=None):
        if keys.numel() != pays.numel():
            raise FijError(Status.INVALID_ARGUMENT, "Join and payload attributes have different sizes")
        check(lib.fij_np_build(int(self.ht.ktype), int(self.scheme), int(self.is_selective), _ptr(self.ht.mem),
                               self.ht.entries, _ptr(keys), _ptr(pays), keys.numel(), self.build_dim[0],
                               self.build_dim[1], _stream(stream)))

    def probe_sum(self, keys, pays, result_set=None, aggregates=None, stream=None):
        check(lib.fij_np_probe_sum(int(self.ht.ktype), int(self.scheme), _ptr(self.ht.mem), self.ht.entries,
                                   _ptr(keys), _ptr(pays), keys.numel(), _ptr(result_set),
                                   result_set.numel() if result_set is not None else 0, _ptr(aggregates),
                                   self.probe_dim[0], self.probe_dim[1], _stream(stream)))


CudaHashJoin = HipHashJoin


# --------------------------------------------------------------------------
# data generation


def gen_permutation(out, seed, base=1, stream=None):
    check(lib.fij_gen_permutation(int(_ktype(out.dtype)), _ptr(out), out.numel(), seed, base, _stream(stream)))


def gen_uniform(out, seed, stream_id, lo, count, stream=None):
    check(lib.fij_gen_uniform(int(_ktype(out.dtype)), _ptr(out), out.numel(), seed, stream_id, lo, count,
                              _stream(stream)))


def gen_iota(out, base=0, stream=None):
    check(lib.fij_gen_iota(int(_ktype(out.dtype)), _ptr(out), out.numel(), base, _stream(stream)))


def gen_zipf(out, num_ranks, exponent, seed, stream_id, perm_seed, base=1, stream=None):
    """Zipf(exponent) keys over the keys gen_permutation(num_ranks, perm_seed, base) makes."""
    check(lib.fij_gen_zipf(int(_ktype(out.dtype)), _ptr(out), out.numel(), num_ranks, float(exponent), seed,
                           stream_id, perm_seed, base, _stream(stream)))


# --------------------------------------------------------------------------
# radix-join execution method (include/fij.h: fij_join_plan_*)


class _PlanConfig(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("radix_bits_1", C.c_uint32),
        ("radix_bits_2", C.c_uint32),
        ("pass1_chunks", C.c_uint32),
        ("pass1_block", C.c_uint32),
        ("join_blocks", C.c_uint32),
        ("build_len", C.c_uint64),
        ("probe_len", C.c_uint64),
    ]


class _Timing(C.Structure):
    _fields_ = [("prefix_sum_ms", C.c_float), ("partition_ms", C.c_float), ("join_ms", C.c_float),
                ("total_ms", C.c_float), ("pass2_ms", C.c_float), ("build_probe_ms", C.c_float),
                ("deferred_ms", C.c_float)]


_proto("fij_join_plan_create", C.c_int, C.POINTER(_PlanConfig), C.POINTER(_vp))
_proto("fij_join_plan_destroy", C.c_int, _vp)
_proto("fij_join_plan_device_bytes", C.c_int, _vp, C.POINTER(_u64))
_proto("fij_join_plan_enqueue", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp)
_proto("fij_join_plan_result", C.c_int, _vp, _vp, C.POINTER(JoinAggregate), C.POINTER(_Timing))
_proto("fij_join_plan_execute", C.c_int, _vp, _vp, _vp, _vp, _vp, C.POINTER(JoinAggregate), C.POINTER(_Timing), _vp)
_proto("fij_join_plan_enqueue_sized", C.c_int, _vp, _vp, _vp, _u64, _vp, _vp, _u64, _vp, _vp)


class JoinPlan:
    """Radix join of R (build) and S (probe): pass 1 with bits1, pass 2 with
    bits2 (0 = single pass), LDS hash build/probe; returns
    (match count, SUM(R.value), SUM(S.value)) as wrapping u64."""

    def __init__(self, ktype, build_len, probe_len, bits1, bits2, pass1_chunks=0, pass1_block=0, join_blocks=0):
        cfg = _PlanConfig()
        cfg.type = int(_ktype(ktype))
        cfg.radix_bits_1 = bits1
        cfg.radix_bits_2 = bits2
        cfg.pass1_chunks = pass1_chunks
        cfg.pass1_block = pass1_block
        cfg.join_blocks = join_blocks
        cfg.build_len = build_len
        cfg.probe_len = probe_len
        h = _vp()
        check(lib.fij_join_plan_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.build_len, self.probe_len = build_len, probe_len

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_join_plan_destroy(self.h)
            self.h = None

    def device_bytes(self):
        out = _u64()
        check(lib.fij_join_plan_device_bytes(self.h, C.byref(out)))
        return out.value

    def enqueue(self, bk, bp, sk, sp, stream=None):
        """Joins columns of exactly the plan's lengths (the C call reads that
        many tuples from each pointer)."""
        for c, n in ((bk, self.build_len), (bp, self.build_len), (sk, self.probe_len), (sp, self.probe_len)):
            if c.numel() != n:
                raise ValueError(f"column of {c.numel()} tuples for a plan of {n}; use enqueue_sized")
        check(lib.fij_join_plan_enqueue(self.h, _ptr(bk), _ptr(bp), _ptr(sk), _ptr(sp), _stream(stream)))

    def result(self, stream=None):
        r = JoinAggregate()
        t = _Timing()
        check(lib.fij_join_plan_result(self.h, _stream(stream), C.byref(r), C.byref(t)))
        return (r.count, r.sum_build, r.sum_probe), {
            "prefix_sum_ms": t.prefix_sum_ms, "partition_ms": t.partition_ms, "join_ms": t.join_ms,
            "total_ms": t.total_ms, "pass2_ms": t.pass2_ms, "build_probe_ms": t.build_probe_ms,
            "deferred_ms": t.deferred_ms}

    def enqueue_sized(self, bk, bp, sk, sp, device_result, stream=None):
        """Enqueues a join of bk.numel() x sk.numel() tuples (at most the plan's
        lengths) whose {count, sums} go to device_result (a device buffer of
        at least 24 bytes; DeviceArray or raw address)."""
        addr = device_result if isinstance(device_result, int) else device_result.ptr
        if bp.numel() < bk.numel() or sp.numel() < sk.numel():
            raise ValueError("payload columns shorter than their key columns")
        check(lib.fij_join_plan_enqueue_sized(self.h, _ptr(bk), _ptr(bp), bk.numel(), _ptr(sk), _ptr(sp),
                                              sk.numel(), C.c_void_p(addr), _stream(stream)))

    def execute(self, bk, bp, sk, sp, stream=None):
        self.enqueue(bk, bp, sk, sp, stream)
        return self.result(stream)


# --------------------------------------------------------------------------
# batched joins (include/fij.h: fij_join_batch_*)

class JoinDesc(C.Structure):
    _fields_ = [("build_keys", _vp), ("build_payloads", _vp), ("build_len", _u64), ("probe_keys", _vp),
                ("probe_payloads", _vp), ("probe_len", _u64)]


class JoinBatchConfig(C.Structure):
    _fields_ = [("type", _u32), ("max_joins", _u32), ("use_graph", _u32), ("lanes", _u32)]


class JoinBatchStats(C.Structure):
    _fields_ = [("gpu_ms", C.c_float), ("prepare_ms", C.c_float), ("total_ms", C.c_float), ("small_joins", _u32),
                ("partitioned_joins", _u32), ("size_classes", _u32), ("graph", _u32), ("grow_ms", C.c_float),
                ("record_ms", C.c_float)]


_proto("fij_join_autotune", C.c_int, C.c_int, _u64, _u64, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32))
_proto("fij_join_batch_create", C.c_int, C.POINTER(JoinBatchConfig), C.POINTER(_vp))
_proto("fij_join_batch_destroy", C.c_int, _vp)
_proto("fij_join_batch_execute", C.c_int, _vp, _u32, C.POINTER(JoinDesc), _vp, C.POINTER(JoinBatchStats), _vp)
_proto("fij_join_batch_device_bytes", C.c_int, _vp, C.POINTER(_u64))


_proto("fij_join_plan_bits", C.c_int, C.c_int, _u64, _u64, C.POINTER(_u32), C.POINTER(_u32))


def plan_bits(dtype, build_len, probe_len):
    """(radix_bits_1, radix_bits_2) for one join run alone: a single pass of
    9 bits when the build side's partitions fit block tables; else two
    passes, with items of about a third of a block table for probe sides of
    16 |R| or more, else the size-class rule's."""
    b1, b2 = _u32(), _u32()
    check(lib.fij_join_plan_bits(int(_ktype(dtype)), build_len, probe_len, C.byref(b1), C.byref(b2)))
    return b1.value, b2.value


def autotune(dtype, build_len, probe_len):
    """(small, radix_bits_1, radix_bits_2) of the pair's size class (tuning
    table entry, or the closed-form rule)."""
    sm, b1, b2 = _u32(), _u32(), _u32()
    check(lib.fij_join_autotune(int(_ktype(dtype)), build_len, probe_len, C.byref(sm), C.byref(b1), C.byref(b2)))
    return bool(sm.value), b1.value, b2.value


class JoinTuning(C.Structure):
    _fields_ = [("build_log2", _u32), ("probe_log2", _u32), ("small", _u32), ("radix_bits_1", _u32),
                ("radix_bits_2", _u32), ("ring_lines", _u32), ("block", _u32), ("from_table", _u32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_proto("fij_join_tuning_lookup", C.c_int, C.c_int, _u64, _u64, C.POINTER(JoinTuning))
_proto("fij_join_tuning_set", C.c_int, C.c_int, C.POINTER(JoinTuning), _u32, C.c_int)


def tuning_lookup(dtype, build_len, probe_len):
    """The configuration batched joins of these lengths run with (a dict of
    fij_join_tuning's fields)."""
    t = JoinTuning()
    check(lib.fij_join_tuning_lookup(int(_ktype(dtype)), build_len, probe_len, C.byref(t)))
    return t.as_dict()


def tuning_set(dtype, entries=None, builtin=False):
    """Replaces the active tuning table of `dtype` with `entries` (dicts of
    fij_join_tuning fields; [] = the closed-form rule only), or restores the
    built-in table."""
    entries = entries or []
    arr = (JoinTuning * max(len(entries), 1))()
    for i, e in enumerate(entries):
        for k, _ in JoinTuning._fields_:
            setattr(arr[i], k, int(e.get(k, 0)))
    check(lib.fij_join_tuning_set(int(_ktype(dtype)), arr, len(entries), 1 if builtin else 0))


class JoinBatch:
    """Many independent joins with one submission.  execute() takes a list of
    (build_keys, build_payloads, probe_keys, probe_payloads) DeviceArrays and
    returns one (count, sum_build, sum_probe) per join and the stats."""

    def __init__(self, dtype, max_joins, use_graph=True, lanes=4):
        cfg = JoinBatchConfig(int(_ktype(dtype)), max_joins, 1 if use_graph else 0, lanes)
        h = _vp()
        check(lib.fij_join_batch_create(C.byref(cfg), C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.fij_join_batch_destroy(C.c_void_p(self.h))
            self.h = None

    def execute(self, joins, stream=None):
        n = len(joins)
        descs = (JoinDesc * max(n, 1))()
        for i, (rk, rp, sk, sp) in enumerate(joins):
            descs[i] = JoinDesc(rk.ptr, rp.ptr, rk.numel(), sk.ptr, sp.ptr, sk.numel())
        res = (JoinAggregate * max(n, 1))()
        st = JoinBatchStats()
        check(lib.fij_join_batch_execute(C.c_void_p(self.h), n, descs, C.cast(res, _vp), C.byref(st), _stream(stream)))
        out = [(res[i].count, res[i].sum_build, res[i].sum_probe) for i in range(n)]
        stats = {k: getattr(st, k) for k, _ in JoinBatchStats._fields_}
        return out, stats

    def device_bytes(self):
        b = _u64()
        check(lib.fij_join_batch_device_bytes(C.c_void_p(self.h), C.byref(b)))
        return b.value


# --------------------------------------------------------------------------
# served joins (include/fij.h: fij_join_server_*)

SERVER_MAX_SLOT_CLASSES = 4


class JoinServerConfig(C.Structure):
    _fields_ = [("type", _u32), ("slot_classes", _u32), ("slots", _u32 * SERVER_MAX_SLOT_CLASSES),
                ("slot_probe_len", _u64 * SERVER_MAX_SLOT_CLASSES), ("max_probe_len", _u64), ("hbm_budget", _u64),
                ("table_build_max", _u64)]


class JoinRequestTimes(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("latency_ms", C.c_double)]


_proto("fij_join_server_create", C.c_int, C.POINTER(JoinServerConfig), C.POINTER(_vp))
_proto("fij_join_server_destroy", C.c_int, _vp)
_proto("fij_join_server_add_build", C.c_int, _vp, _vp, _vp, _u64, C.POINTER(_u32))
_proto("fij_join_server_submit", C.c_int, _vp, _u32, _vp, _vp, _u64, C.POINTER(_u64))
_proto("fij_join_server_wait", C.c_int, _vp, _u64, C.POINTER(JoinAggregate), C.POINTER(JoinRequestTimes))
_proto("fij_join_server_memory", C.c_int, _vp, C.POINTER(_u64), C.POINTER(_u32))


class JoinServer:
    """Joins against resident build sides.  slot_classes is a list of
    (slots, largest probe length); add_build() partitions a build side and
    keeps it; submit() queues a probe side (DeviceArrays that stay valid until
    wait() returns) and returns a request id; wait() returns
    ((count, sum_build, sum_probe), {"queued_ms", "latency_ms"})."""

    def __init__(self, dtype, slot_classes, max_probe_len, hbm_budget=0, table_build_max=0):
        if not 0 < len(slot_classes) <= SERVER_MAX_SLOT_CLASSES:
            raise ValueError("between 1 and %d slot classes" % SERVER_MAX_SLOT_CLASSES)
        cfg = JoinServerConfig()
        cfg.type = int(_ktype(dtype))
        cfg.slot_classes = len(slot_classes)
        for i, (n, cap) in enumerate(slot_classes):
            cfg.slots[i] = n
            cfg.slot_probe_len[i] = cap
        cfg.max_probe_len = max_probe_len
        cfg.hbm_budget = hbm_budget
        cfg.table_build_max = table_build_max  # smaller build sides: one resident hash table
        h = _vp()
        check(lib.fij_join_server_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self._held = {}

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "h", None):
            lib.fij_join_server_destroy(C.c_void_p(self.h))
            self.h = None
            self._held = {}

    def add_build(self, keys, payloads):
        bid = _u32()
        check(lib.fij_join_server_add_build(C.c_void_p(self.h), keys.ptr, payloads.ptr, keys.numel(), C.byref(bid)))
        return bid.value

    def submit(self, build_id, keys, payloads):
        rid = _u64()
        check(lib.fij_join_server_submit(C.c_void_p(self.h), build_id, keys.ptr, payloads.ptr, keys.numel(),
                                         C.byref(rid)))
        self._held[rid.value] = (keys, payloads)
        return rid.value

    def wait(self, request_id):
        res = JoinAggregate()
        t = JoinRequestTimes()
        try:
            check(lib.fij_join_server_wait(C.c_void_p(self.h), request_id, C.byref(res), C.byref(t)))
        finally:
            self._held.pop(request_id, None)
        return (res.count, res.sum_build, res.sum_probe), {"queued_ms": t.queued_ms, "latency_ms": t.latency_ms}

    def memory(self):
        used, builds = _u64(), _u32()
        check(lib.fij_join_server_memory(C.c_void_p(self.h), C.byref(used), C.byref(builds)))
        return used.value, builds.value


# --------------------------------------------------------------------------
# relation files (include/fij.h: fij_relation_file_*)

_proto("fij_relation_file_len", C.c_int, C.c_char_p, C.c_char, C.POINTER(_u64))
_proto("fij_relation_file_read", C.c_int, C.c_char_p, C.c_char, C.c_int, _vp, _vp, _u64, C.POINTER(_u64))
_proto("fij_relation_file_write", C.c_int, C.c_char_p, C.c_char, C.c_int, _vp, _vp, _u64)


def _delim(d):
    return b"\0" if d is None else d.encode()


def relation_file_len(path, delimiter=None):
    n = _u64()
    check(lib.fij_relation_file_len(os.fsencode(path), _delim(delimiter), C.byref(n)))
    return n.value


def read_relation_file(path, dtype, delimiter=None):
    """(keys, payloads) host arrays of a relation file, as the reference's
    JoinDataBuilder::build_with_files reads it (two passes: count, read)."""
    dtype = np.dtype(dtype)
    n = relation_file_len(path, delimiter)
    keys = np.empty(n, dtype)
    pays = np.empty(n, dtype)
    got = _u64()
    check(lib.fij_relation_file_read(os.fsencode(path), _delim(delimiter), int(_ktype(dtype)),
                                     keys.ctypes.data_as(_vp), pays.ctypes.data_as(_vp), n, C.byref(got)))
    if got.value != n:
        raise FijError(Status.IO_ERROR, f"{path} changed while it was read")
    return keys, pays


def write_relation_file(path, keys, payloads, delimiter=None):
    keys = np.ascontiguousarray(keys)
    payloads = np.ascontiguousarray(payloads, dtype=keys.dtype)
    if len(keys) != len(payloads):
        raise ValueError("keys and payloads differ in length")
    check(lib.fij_relation_file_write(os.fsencode(path), _delim(delimiter), int(_ktype(keys.dtype)),
                                      keys.ctypes.data_as(_vp), payloads.ctypes.data_as(_vp), len(keys)))


# --------------------------------------------------------------------------
# host partitioner, pinned host arrays, out-of-core join

class HostArray:
    """Pinned host memory (fij_mem_alloc_host) with a numpy view `.a`; the
    GPU can read it in place."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = _vp()
        check(lib.fij_mem_alloc_host(max(self.n, 1) * self.dtype.itemsize, C.byref(p)))
        self.ptr = p.value
        buf = (C.c_char * (max(self.n, 1) * self.dtype.itemsize)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=self.n)

    @classmethod
    def from_numpy(cls, x):
        h = cls(len(x), np.asarray(x).dtype)
        h.a[:] = x
        return h

    def numel(self):
        return self.n

    def __del__(self):
        if getattr(self, "ptr", None):
            self.a = None
            lib.fij_mem_free_host(C.c_void_p(self.ptr))
            self.ptr = None


_proto("fij_cpu_prefix_sum", C.c_int, C.c_int, C.c_int, _vp, _u64, _u32, _u32, _u32, C.POINTER(_PartitionOffsets))
_proto("fij_cpu_radix_partition", C.c_int, C.c_int, _vp, _vp, _u64, _u32, _u32, _u32, _vp, _vp)


def cpu_prefix_sum(histogram_algorithm, keys, radix_bits, chunks, ignore_bits=0, threads=0):
    """Offsets of fij_prefix_sum computed on the host: Chunked -> [chunks * F];
    Contiguous -> ([F], [chunks * F] local offsets)."""
    keys = np.ascontiguousarray(keys)
    F = 1 << radix_bits
    contiguous = int(histogram_algorithm) == 1
    offs = np.zeros(F if contiguous else chunks * F, np.uint64)
    local = np.zeros(chunks * F, np.uint64) if contiguous else None
    o = _PartitionOffsets(offs.ctypes.data, local.ctypes.data if contiguous else None, 0, chunks, chunks,
                          radix_bits, int(_ktype(keys.dtype)))
    check(lib.fij_cpu_prefix_sum(int(histogram_algorithm), int(_ktype(keys.dtype)), keys.ctypes.data_as(_vp),
                                 len(keys), radix_bits, ignore_bits, threads, C.byref(o)))
    return (offs, local) if contiguous else offs


def cpu_radix_partition(keys, payloads, radix_bits, ignore_bits=0, threads=0, out=None):
    """Contiguous pass 1 on the host: (keys, values) of the padded relation
    (padding slots hold whatever was there: zeros here) and the partition
    starts.  `out`: a C-contiguous (len + F * pad, 2) array of the key type
    to write into instead of a new zeroed one."""
    keys = np.ascontiguousarray(keys)
    payloads = np.ascontiguousarray(payloads, dtype=keys.dtype)
    F = 1 << radix_bits
    pad = 128 // (2 * keys.dtype.itemsize)
    if out is None:
        rel = np.zeros((len(keys) + F * pad, 2), keys.dtype)
    else:
        rel = out
        if rel.shape != (len(keys) + F * pad, 2) or rel.dtype != keys.dtype or not rel.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous (len + F * pad, 2) array of the key type")
    offs = np.zeros(F, np.uint64)
    check(lib.fij_cpu_radix_partition(int(_ktype(keys.dtype)), keys.ctypes.data_as(_vp),
                                      payloads.ctypes.data_as(_vp), len(keys), radix_bits, ignore_bits, threads,
                                      rel.ctypes.data_as(_vp), offs.ctypes.data_as(_vp)))
    return rel[:, 0], rel[:, 1], offs


class OocConfig(C.Structure):
    _fields_ = [("type", _u32), ("radix_bits_1", _u32), ("radix_bits_2", _u32), ("pass1_cpu", _u32),
                ("cpu_threads", _u32), ("probe_slices", _u32), ("cache_bytes", _u64), ("staging_tuples", _u64)]


class OocTiming(C.Structure):
    _fields_ = [("prefix_sum_ms", C.c_float), ("partition_ms", C.c_float), ("join_ms", C.c_float),
                ("total_ms", C.c_float), ("cached_build_tuples", _u64), ("cached_probe_tuples", _u64),
                ("groups", _u32), ("reserved", _u32)]


_proto("fij_ooc_join", C.c_int, C.POINTER(OocConfig), _vp, _vp, _u64, _vp, _vp, _u64, C.POINTER(JoinAggregate),
       C.POINTER(OocTiming))


def _col_ptr(x, dtype):
    if x.dtype != dtype:
        raise ValueError(f"column of {x.dtype} in a join of {dtype}")
    if isinstance(x, np.ndarray):
        if not x.flags.c_contiguous:
            raise ValueError("columns must be contiguous")
        return x.ctypes.data, len(x)
    return x.ptr, x.numel()


def ooc_join(dtype, bk, bp, sk, sp, bits1, bits2, pass1_cpu=False, cache_bytes=0, staging_tuples=0,
             cpu_threads=0, probe_slices=0):
    """Out-of-core radix join; columns are numpy arrays (CPU pass 1 only),
    HostArray or DeviceArray.  Returns ((count, sum_build, sum_probe), timing)."""
    cfg = OocConfig(int(_ktype(dtype)), bits1, bits2, 1 if pass1_cpu else 0, cpu_threads, probe_slices,
                    cache_bytes, staging_tuples)
    want = numpy_dtype(_ktype(dtype))
    (bkp, nb), (bpp, nbp), (skp, ns), (spp, nsp) = (_col_ptr(x, want) for x in (bk, bp, sk, sp))
    if nbp != nb or nsp != ns:
        raise ValueError("Join and payload attributes have different sizes")
    res = JoinAggregate()
    t = OocTiming()
    check(lib.fij_ooc_join(C.byref(cfg), bkp, bpp, nb, skp, spp, ns, C.byref(res), C.byref(t)))
    return (res.count, res.sum_build, res.sum_probe), {k: getattr(t, k) for k, _ in OocTiming._fields_}
