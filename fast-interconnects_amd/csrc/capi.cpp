// C ABI of the library (declared in include/fij.h).
//
// Argument checks mirror the reference's operator methods and return the
// same error kinds as status codes:
//   prefix_sum            sql-ops/src/partition/gpu_radix_partition.rs:561-612
//   copy_with_payload     gpu_radix_partition.rs:698-760
//   prefix_sum_and_transform  gpu_radix_partition.rs:826-892
//   partition             gpu_radix_partition.rs:1022-1085
//   CudaRadixJoin::join   sql-ops/src/join/cuda_radix_join.rs:172-199
//   CudaHashJoin::build   sql-ops/src/join/no_partitioning_join.rs:313-375
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/fij.h"
#include "internal.hpp"

using namespace fij;

namespace {

thread_local std::string g_last_error;

fij_status fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

}  // namespace

namespace fij {
fij_status set_error(fij_status s, const char* msg) {
  g_last_error = msg;
  return s;
}
}  // namespace fij

namespace {

#define FIJ_HIP(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(FIJ_DEVICE_ERROR, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr uint64_t kU32Max = 0xFFFFFFFFull;
constexpr uint32_t kMaxRadixBits = 15;
constexpr int kMaxDevices = 64;

struct DeviceInfo {
  bool init = false;
  fij_device_limits limits{};
  uint32_t* status = nullptr;  // device status word
};
DeviceInfo g_devices[kMaxDevices];
std::mutex g_dev_mutex;

fij_status device_info(DeviceInfo** out) {
  int dev = 0;
  FIJ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FIJ_RUNTIME_ERROR, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  DeviceInfo& d = g_devices[dev];
  if (!d.init) {
    hipDeviceProp_t prop;
    FIJ_HIP(hipGetDeviceProperties(&prop, dev));
    d.limits.compute_units = static_cast<uint32_t>(prop.multiProcessorCount);
    d.limits.wave_size = static_cast<uint32_t>(prop.warpSize);
    uint64_t lds = prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
      lds = std::max<uint64_t>(lds, static_cast<uint64_t>(optin));
    d.limits.lds_bytes_per_block = static_cast<uint32_t>(lds);
    d.limits.max_threads_per_block = static_cast<uint32_t>(prop.maxThreadsPerBlock);
    d.limits.global_mem_bytes = prop.totalGlobalMem;
    std::snprintf(d.limits.arch_name, sizeof(d.limits.arch_name), "%s", prop.gcnArchName);
    FIJ_HIP(hipMalloc(&d.status, sizeof(uint32_t)));
    FIJ_HIP(hipMemset(d.status, 0, sizeof(uint32_t)));
    d.init = true;
  }
  *out = &d;
  return FIJ_OK;
}

bool valid_type(uint32_t t) { return t == FIJ_TYPE_I32 || t == FIJ_TYPE_I64; }
KeyType key_type(fij_type t) { return t == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64; }

fij_status pass_bits(const fij_radix_bits& b, fij_radix_pass pass, uint32_t* bits) {
  if (pass < FIJ_PASS_FIRST || pass > FIJ_PASS_THIRD || b.pass_bits[pass] < 0)
    return fail(FIJ_INVALID_ARGUMENT, "The requested partitioning pass is not specified");
  if (static_cast<uint32_t>(b.pass_bits[pass]) > kMaxRadixBits)
    return fail(FIJ_UNSUPPORTED, "at most %u radix bits per pass are supported", kMaxRadixBits);
  *bits = static_cast<uint32_t>(b.pass_bits[pass]);
  return FIJ_OK;
}

// RadixBits::pass_ignore_bits (sql-ops/src/partition.rs:104-110).
uint32_t pass_ignore_bits(const fij_radix_bits& b, fij_radix_pass pass) {
  uint32_t s = 0;
  for (int i = 0; i < static_cast<int>(pass); ++i)
    if (b.pass_bits[i] > 0) s += static_cast<uint32_t>(b.pass_bits[i]);
  return s;
}

uint64_t padded_len(uint64_t len, uint32_t radix_bits, uint32_t chunks, fij_type t) {
  return len + (static_cast<uint64_t>(1) << radix_bits) * chunks * pad_tuples(key_type(t));
}

bool aligned_to(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace

struct fij_partitioner {
  fij_histogram_algorithm hist;
  fij_partition_algorithm algo;
  fij_radix_bits bits;
  uint32_t grid;
  uint32_t block;
  uint64_t dmem_buffer_bytes;
  uint32_t* counts = nullptr;
  uint64_t counts_cap = 0;
  uint64_t* scan_state = nullptr;
  uint64_t scan_cap = 0;
};

struct fij_radix_join {
  fij_radix_pass pass;
  fij_radix_bits bits;
  fij_hashing_scheme scheme;
  uint32_t grid;
  uint32_t block;
};

namespace {

// Grows the partitioner's contiguous-scan state.  A grow waits for the
// stream first so that no in-flight kernel still uses the old buffers.
fij_status ensure_scan_state(fij_partitioner* p, uint32_t fanout, hipStream_t stream) {
  const uint64_t n = static_cast<uint64_t>(fanout) * p->grid;
  const uint64_t words = scan_state_words(n);
  if (n > p->counts_cap || words > p->scan_cap) FIJ_HIP(hipStreamSynchronize(stream));
  if (n > p->counts_cap) {
    if (p->counts) FIJ_HIP(hipFree(p->counts));
    p->counts = nullptr;
    if (hipMalloc(&p->counts, n * sizeof(uint32_t)) != hipSuccess)
      return fail(FIJ_OUT_OF_MEMORY, "cannot allocate %llu scan counters", static_cast<unsigned long long>(n));
    p->counts_cap = n;
  }
  if (words > p->scan_cap) {
    if (p->scan_state) FIJ_HIP(hipFree(p->scan_state));
    p->scan_state = nullptr;
    if (hipMalloc(&p->scan_state, words * sizeof(uint64_t)) != hipSuccess)
      return fail(FIJ_OUT_OF_MEMORY, "cannot allocate scan state");
    p->scan_cap = words;
  }
  return FIJ_OK;
}

fij_status check_block(uint32_t block) {
  if (block == 0 || block % 64 != 0 || block > 1024)
    return fail(FIJ_INVALID_ARGUMENT, "block size must be a multiple of 64 in [64, 1024], got %u", block);
  return FIJ_OK;
}

// The canonical chunking (partition_input_chunk.rs:24-35) gives every chunk
// but the last chunk_len tuples; when (grid - 1) * chunk_len exceeds the
// length, the last chunk's length would be negative and the padded layout
// inconsistent, so such grids are rejected.
fij_status check_chunking(fij_type type, uint64_t len, uint32_t grid) {
  const uint64_t align = 128 / key_bytes(key_type(type));
  const uint64_t c = ((len + grid - 1) / grid) & ~(align - 1);
  if (c * (grid - 1) > len)
    return fail(FIJ_INVALID_ARGUMENT, "Relation of %llu tuples is too short for %u chunks",
                static_cast<unsigned long long>(len), grid);
  return FIJ_OK;
}

}  // namespace

extern "C" {

const char* fij_last_error(void) { return g_last_error.c_str(); }

const char* fij_version(void) { return "fij 0.1.0 gfx950"; }

fij_status fij_device_limits_get(int device, fij_device_limits* out) {
  if (!out) return fail(FIJ_INVALID_ARGUMENT, "out is NULL");
  int cur = 0;
  FIJ_HIP(hipGetDevice(&cur));
  if (device != cur) FIJ_HIP(hipSetDevice(device));
  DeviceInfo* d = nullptr;
  fij_status s = device_info(&d);
  if (device != cur) (void)hipSetDevice(cur);
  if (s != FIJ_OK) return s;
  *out = d->limits;
  return FIJ_OK;
}

fij_status fij_device_status(void* stream, int clear, uint32_t* out) {
  DeviceInfo* d = nullptr;
  fij_status s = device_info(&d);
  if (s != FIJ_OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint32_t v = 0;
  FIJ_HIP(hipMemcpyAsync(&v, d->status, sizeof(v), hipMemcpyDeviceToHost, st));
  FIJ_HIP(hipStreamSynchronize(st));
  if (clear) {
    const uint32_t zero = 0;
    FIJ_HIP(hipMemcpyAsync(d->status, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    FIJ_HIP(hipStreamSynchronize(st));
  }
  if (out) *out = v;
  return FIJ_OK;
}

// ------------------------------------------------------------ layout ----

uint32_t fij_padding_len(fij_type type) { return pad_tuples(key_type(type)); }

fij_status fij_input_chunk_size(fij_type type, uint64_t data_len, uint32_t num_chunks, uint64_t* chunk_len) {
  if (!valid_type(type) || num_chunks == 0 || !chunk_len) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const uint64_t align = 128 / key_bytes(key_type(type));
  const uint64_t c = ((data_len + num_chunks - 1) / num_chunks) & ~(align - 1);
  *chunk_len = c;
  if (c >= kU32Max)
    return fail(FIJ_INTEGER_OVERFLOW,
                "Relation is too large and causes an integer overflow. Try using more chunks by setting a "
                "higher CUDA grid size");
  return FIJ_OK;
}

fij_status fij_partition_offsets_layout(fij_histogram_algorithm algo, uint32_t max_chunks, uint32_t radix_bits,
                                        uint32_t* chunks, uint64_t* offsets_len, uint64_t* local_offsets_len) {
  if (max_chunks == 0 || radix_bits > kMaxRadixBits || !chunks || !offsets_len || !local_offsets_len)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const uint64_t F = 1ull << radix_bits;
  *chunks = algo == FIJ_HISTOGRAM_CHUNKED ? max_chunks : 1;
  *offsets_len = F * *chunks;
  *local_offsets_len = algo == FIJ_HISTOGRAM_CONTIGUOUS ? F * max_chunks : 0;
  return FIJ_OK;
}

fij_status fij_partitioned_relation_layout(fij_type type, uint64_t len, fij_histogram_algorithm algo,
                                           uint32_t radix_bits, uint32_t max_chunks, uint32_t* chunks,
                                           uint64_t* relation_len, uint64_t* offsets_len) {
  if (!valid_type(type) || max_chunks == 0 || radix_bits > kMaxRadixBits || !chunks || !relation_len ||
      !offsets_len)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const uint64_t F = 1ull << radix_bits;
  *chunks = algo == FIJ_HISTOGRAM_CHUNKED ? max_chunks : 1;
  *relation_len = padded_len(len, radix_bits, *chunks, type);
  *offsets_len = F * *chunks;
  return FIJ_OK;
}

uint64_t fij_partitioned_relation_padded_len(const fij_partitioned_relation* rel) {
  if (!rel || !valid_type(rel->type)) return 0;
  return padded_len(rel->len, rel->radix_bits, rel->chunks, static_cast<fij_type>(rel->type));
}

uint64_t fij_partitioned_relation_capacity(const fij_partitioned_relation* rel) {
  if (!rel || !valid_type(rel->type)) return 0;
  const uint64_t pads = padded_len(0, rel->radix_bits, rel->chunks, static_cast<fij_type>(rel->type));
  return rel->relation_len >= pads ? rel->relation_len - pads : 0;
}

fij_status fij_partitioned_relation_resize(fij_partitioned_relation* rel, uint64_t new_len) {
  if (!rel) return fail(FIJ_INVALID_ARGUMENT, "relation is NULL");
  if (new_len <= fij_partitioned_relation_capacity(rel)) {
    rel->len = new_len;
    return FIJ_OK;
  }
  return fail(FIJ_INVALID_ARGUMENT, "Insufficient capacity to resize to new length");
}

static fij_status partition_len_impl(const uint64_t* off, uint32_t chunks, uint32_t radix_bits, uint64_t plen,
                                     uint32_t pad, uint32_t pid, uint64_t* len) {
  const uint32_t F = 1u << radix_bits;
  if (pid >= F) return fail(FIJ_INVALID_ARGUMENT, "Invalid partition ID");
  const uint64_t olen = static_cast<uint64_t>(chunks) * F;
  uint64_t total = 0;
  for (uint32_t c = 0; c < chunks; ++c) {
    const uint64_t ofi = static_cast<uint64_t>(c) * F + pid;
    const uint64_t begin = off[ofi];
    const uint64_t end = (ofi + 1 < olen) ? off[ofi + 1] - pad : plen;
    total += end - begin;
  }
  *len = total;
  return FIJ_OK;
}

fij_status fij_partitioned_relation_partition_len(const fij_partitioned_relation* rel, const uint64_t* host_offsets,
                                                  uint32_t partition_id, uint64_t* len) {
  if (!rel || !host_offsets || !len || !valid_type(rel->type)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  return partition_len_impl(host_offsets, rel->chunks, rel->radix_bits, fij_partitioned_relation_padded_len(rel),
                            pad_tuples(key_type(static_cast<fij_type>(rel->type))), partition_id, len);
}

fij_status fij_partition_offsets_partition_len(const fij_partition_offsets* offs, const uint64_t* host_offsets,
                                               uint32_t partition_id, uint64_t* len) {
  if (!offs || !host_offsets || !len || !valid_type(offs->type)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (offs->data_len == FIJ_LEN_UNSET) return fail(FIJ_RUNTIME_ERROR, "Data length not yet computed");
  const fij_type t = static_cast<fij_type>(offs->type);
  return partition_len_impl(host_offsets, offs->chunks, offs->radix_bits,
                            padded_len(offs->data_len, offs->radix_bits, offs->chunks, t),
                            pad_tuples(key_type(t)), partition_id, len);
}

// -------------------------------------------------------- partitioner ----

fij_status fij_partitioner_create(fij_histogram_algorithm histogram_algorithm,
                                  fij_partition_algorithm partition_algorithm, fij_radix_bits radix_bits,
                                  uint32_t grid_size, uint32_t block_size, uint64_t dmem_buffer_bytes,
                                  fij_partitioner** out) {
  if (!out) return fail(FIJ_INVALID_ARGUMENT, "out is NULL");
  if (histogram_algorithm != FIJ_HISTOGRAM_CHUNKED && histogram_algorithm != FIJ_HISTOGRAM_CONTIGUOUS)
    return fail(FIJ_INVALID_ARGUMENT, "unknown histogram algorithm");
  if (partition_algorithm < FIJ_PARTITION_NC || partition_algorithm > FIJ_PARTITION_HSSWWC_V4)
    return fail(FIJ_INVALID_ARGUMENT, "unknown partition algorithm");
  if (grid_size == 0) return fail(FIJ_INVALID_ARGUMENT, "grid size must be positive");
  const uint32_t block = std::min<uint32_t>(block_size, 1024);
  fij_status s = check_block(block);
  if (s != FIJ_OK) return s;
  fij_partitioner* p = new fij_partitioner();
  p->hist = histogram_algorithm;
  p->algo = partition_algorithm;
  p->bits = radix_bits;
  p->grid = grid_size;
  p->block = block;
  p->dmem_buffer_bytes = dmem_buffer_bytes;
  *out = p;
  return FIJ_OK;
}

fij_status fij_partitioner_destroy(fij_partitioner* p) {
  if (!p) return FIJ_OK;
  if (p->counts) (void)hipFree(p->counts);
  if (p->scan_state) (void)hipFree(p->scan_state);
  delete p;
  return FIJ_OK;
}

}  // extern "C"

namespace {

// Argument checks of prefix_sum (gpu_radix_partition.rs:561-696); sets *bits
// to the pass's radix bits.
fij_status prefix_sum_checks(const fij_partitioner* p, fij_radix_pass pass, fij_type type,
                             const void* partition_attr, uint64_t len, const fij_partition_offsets* offsets,
                             uint32_t* bits_out) {
  if (!p || !offsets || !valid_type(type) || (len && !partition_attr))
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  uint32_t bits = 0;
  fij_status s = pass_bits(p->bits, pass, &bits);
  if (s != FIJ_OK) return s;
  if (offsets->radix_bits != bits) return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching radix bits");
  if (offsets->type != static_cast<uint32_t>(type)) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has a different tuple type");
  if (p->hist == FIJ_HISTOGRAM_CHUNKED) {
    if (offsets->chunks != p->grid)
      return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching number of chunks");
  } else {
    if (offsets->chunks != 1) return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching number of chunks");
    if (!offsets->local_offsets || offsets->max_chunks != p->grid)
      return fail(FIJ_INVALID_ARGUMENT, "Contiguous offsets need local offsets for every chunk of the grid");
  }
  if (!offsets->offsets) return fail(FIJ_INVALID_ARGUMENT, "offsets buffer is NULL");
  if ((len + p->grid - 1) / p->grid >= kU32Max)
    return fail(FIJ_INTEGER_OVERFLOW,
                "Relation is too large and causes an integer overflow. Try using more chunks by setting a higher "
                "CUDA grid size");
  s = check_chunking(type, len, p->grid);
  if (s != FIJ_OK) return s;
  if (partition_attr && !aligned_to(partition_attr, key_bytes(key_type(type))))
    return fail(FIJ_INVALID_ARGUMENT, "key column is not aligned to its element size");
  *bits_out = bits;
  return FIJ_OK;
}

}  // namespace

extern "C" {

fij_status fij_prefix_sum(fij_partitioner* p, fij_radix_pass pass, fij_type type, const void* partition_attr,
                          uint64_t len, fij_partition_offsets* offsets, void* stream) {
  uint32_t bits = 0;
  fij_status s = prefix_sum_checks(p, pass, type, partition_attr, len, offsets, &bits);
  if (s != FIJ_OK) return s;
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  offsets->data_len = len;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t F = 1u << bits;
  const uint32_t ignore = pass_ignore_bits(p->bits, pass);
  if (p->hist == FIJ_HISTOGRAM_CHUNKED) {
    FIJ_HIP(launch_hist_chunked(key_type(type), partition_attr, len, p->grid, bits, ignore, offsets->offsets,
                                p->block, st));
  } else {
    s = ensure_scan_state(p, F, st);
    if (s != FIJ_OK) return s;
    FIJ_HIP(launch_hist_counts(key_type(type), partition_attr, nullptr, nullptr, nullptr, len, p->grid, bits, ignore,
                               p->counts, p->scan_state, p->scan_cap, p->block, st));
    FIJ_HIP(launch_offsets_scan(p->counts, F, p->grid, pad_tuples(key_type(type)), offsets->local_offsets,
                                offsets->offsets, p->scan_state, st));
  }
  return FIJ_OK;
}

// fij_prefix_sum without any device allocation: the contiguous histogram's
// counters and look-back state come from the library's static scratch array
// (noalloc_scratch), not from the partitioner's lazily grown buffers.  The
// device-info set-up (which allocates the status word) is not needed here.
fij_status fij_prefix_sum_noalloc(fij_partitioner* p, fij_radix_pass pass, fij_type type, const void* partition_attr,
                                  uint64_t len, fij_partition_offsets* offsets, void* stream) {
  uint32_t bits = 0;
  fij_status s = prefix_sum_checks(p, pass, type, partition_attr, len, offsets, &bits);
  if (s != FIJ_OK) return s;
  const uint32_t F = 1u << bits;
  const uint64_t n = static_cast<uint64_t>(F) * p->grid;
  if (p->hist == FIJ_HISTOGRAM_CONTIGUOUS && n > kNoAllocCounters)
    return fail(FIJ_INVALID_ARGUMENT,
                "fanout x grid = %llu exceeds the %llu counters of the allocation-free prefix sum's scratch",
                static_cast<unsigned long long>(n), static_cast<unsigned long long>(kNoAllocCounters));
  offsets->data_len = len;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t ignore = pass_ignore_bits(p->bits, pass);
  if (p->hist == FIJ_HISTOGRAM_CHUNKED) {
    // needs no scratch: the chunked histogram writes the offsets directly
    FIJ_HIP(launch_hist_chunked(key_type(type), partition_attr, len, p->grid, bits, ignore, offsets->offsets,
                                p->block, st));
    return FIJ_OK;
  }
  uint32_t* counts = nullptr;
  uint64_t* state = nullptr;
  FIJ_HIP(noalloc_scratch(&counts, &state));
  // the histogram kernel clears the scan state it needs before the scan
  FIJ_HIP(launch_hist_counts(key_type(type), partition_attr, nullptr, nullptr, nullptr, len, p->grid, bits, ignore,
                             counts, state, scan_state_words(n), p->block, st));
  FIJ_HIP(launch_offsets_scan(counts, F, p->grid, pad_tuples(key_type(type)), offsets->local_offsets,
                              offsets->offsets, state, st));
  return FIJ_OK;
}

fij_status fij_prefix_sum_and_copy_with_payload(fij_partitioner* p, fij_radix_pass pass, fij_type type,
                                                const void* src_partition_attr, const void* src_payload_attr,
                                                uint64_t len, void* dst_partition_attr, void* dst_payload_attr,
                                                uint64_t dst_len, fij_partition_offsets* offsets, void* stream) {
  if (!p || !offsets || !valid_type(type) ||
      (len && (!src_partition_attr || !src_payload_attr || !dst_partition_attr || !dst_payload_attr)))
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  uint32_t bits = 0;
  fij_status s = pass_bits(p->bits, pass, &bits);
  if (s != FIJ_OK) return s;
  if (offsets->radix_bits != bits) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has mismatching radix bits");
  if (offsets->type != static_cast<uint32_t>(type)) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has a different tuple type");
  if (p->hist == FIJ_HISTOGRAM_CHUNKED)
    return fail(FIJ_INVALID_ARGUMENT, "Unsupported histogram algorithm, try using Contiguous instead");
  if (offsets->chunks != 1) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has mismatching number of chunks");
  if (!offsets->local_offsets || offsets->max_chunks != p->grid || !offsets->offsets)
    return fail(FIJ_INVALID_ARGUMENT, "Contiguous offsets need local offsets for every chunk of the grid");
  if (len >= kU32Max) return fail(FIJ_INTEGER_OVERFLOW, "Relation is too large and causes an integer overflow.");
  if (dst_len < len) return fail(FIJ_INVALID_ARGUMENT, "destination is shorter than the source");
  s = check_chunking(type, len, p->grid);
  if (s != FIJ_OK) return s;
  const size_t kb = key_bytes(key_type(type));
  if (len && !(aligned_to(src_partition_attr, kb) && aligned_to(src_payload_attr, kb) &&
               aligned_to(dst_partition_attr, kb) && aligned_to(dst_payload_attr, kb)))
    return fail(FIJ_INVALID_ARGUMENT, "columns must be aligned to their element size");
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  offsets->data_len = len;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t F = 1u << bits;
  s = ensure_scan_state(p, F, st);
  if (s != FIJ_OK) return s;
  FIJ_HIP(launch_hist_counts(key_type(type), src_partition_attr, src_payload_attr, dst_partition_attr,
                             dst_payload_attr, len, p->grid, bits, pass_ignore_bits(p->bits, pass), p->counts,
                             p->scan_state, p->scan_cap, p->block, st));
  FIJ_HIP(launch_offsets_scan(p->counts, F, p->grid, pad_tuples(key_type(type)), offsets->local_offsets,
                              offsets->offsets, p->scan_state, st));
  return FIJ_OK;
}

fij_status fij_prefix_sum_and_transform(fij_partitioner* p, fij_radix_pass pass, fij_type type,
                                        uint32_t partition_id, const fij_partitioned_relation* src,
                                        void* dst_partition_attr, void* dst_payload_attr, uint64_t dst_len,
                                        fij_partition_offsets* offsets, void* stream) {
  if (!p || !offsets || !src || !valid_type(type) || (dst_len && (!dst_partition_attr || !dst_payload_attr)))
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  uint32_t bits = 0;
  fij_status s = pass_bits(p->bits, pass, &bits);
  if (s != FIJ_OK) return s;
  if (src->radix_bits > kMaxRadixBits || partition_id >= (1u << src->radix_bits))
    return fail(FIJ_INVALID_ARGUMENT, "Invalid partition ID");
  if (offsets->radix_bits != bits) return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching radix bits");
  if (offsets->type != static_cast<uint32_t>(type) || src->type != static_cast<uint32_t>(type))
    return fail(FIJ_INVALID_ARGUMENT, "tuple types differ");
  if (p->hist == FIJ_HISTOGRAM_CHUNKED)
    return fail(FIJ_INVALID_ARGUMENT, "Unsupported histogram algorithm, try using Contiguous instead");
  if (offsets->chunks != 1) return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching number of chunks");
  if (!offsets->local_offsets || offsets->max_chunks != p->grid || !offsets->offsets)
    return fail(FIJ_INVALID_ARGUMENT, "Contiguous offsets need local offsets for every chunk of the grid");
  const uint64_t src_plen = fij_partitioned_relation_padded_len(src);
  if (src_plen >= kU32Max) return fail(FIJ_INTEGER_OVERFLOW, "Relation is too large and causes an integer overflow.");
  if (!src->relation || !src->offsets || src->chunks == 0) return fail(FIJ_INVALID_ARGUMENT, "source relation is empty");
  s = check_chunking(type, dst_len, p->grid);
  if (s != FIJ_OK) return s;
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  const uint32_t F = 1u << bits;
  const uint64_t lds = 256 + (2ull * src->chunks + 1) * 8 + static_cast<uint64_t>(F) * 4;
  if (lds > d->limits.lds_bytes_per_block)
    return fail(FIJ_UNSUPPORTED, "too many source chunks for the transform's LDS piece table");
  offsets->data_len = dst_len;
  hipStream_t st = static_cast<hipStream_t>(stream);
  s = ensure_scan_state(p, F, st);
  if (s != FIJ_OK) return s;
  FIJ_HIP(launch_transform_counts(key_type(type), src->relation, src->offsets, src->chunks, src->radix_bits, src_plen,
                                  partition_id, dst_len, dst_partition_attr, dst_payload_attr, p->grid, bits,
                                  pass_ignore_bits(p->bits, pass), p->counts, p->scan_state, p->scan_cap, d->status,
                                  p->block, st));
  FIJ_HIP(launch_offsets_scan(p->counts, F, p->grid, pad_tuples(key_type(type)), offsets->local_offsets,
                              offsets->offsets, p->scan_state, st));
  return FIJ_OK;
}

fij_status fij_preallocate_partition_state(fij_partitioner* p, fij_radix_pass pass, fij_type type) {
  if (!p || !valid_type(type)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  uint32_t bits = 0;
  fij_status s = pass_bits(p->bits, pass, &bits);
  if (s != FIJ_OK) return s;
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  if (p->hist == FIJ_HISTOGRAM_CONTIGUOUS) return ensure_scan_state(p, 1u << bits, nullptr);
  return FIJ_OK;
}

fij_status fij_partition(fij_partitioner* p, fij_radix_pass pass, fij_type type, const void* partition_attr,
                         const void* payload_attr, uint64_t len, fij_partition_offsets* offsets,
                         fij_partitioned_relation* relation, void* stream) {
  if (!p || !offsets || !relation || !valid_type(type) || (len && (!partition_attr || !payload_attr)))
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  uint32_t bits = 0;
  fij_status s = pass_bits(p->bits, pass, &bits);
  if (s != FIJ_OK) return s;
  if (offsets->data_len == FIJ_LEN_UNSET) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has no length");
  if (relation->len != offsets->data_len)
    return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets and PartitionedRelation have mismatching lengths");
  if (len != offsets->data_len)
    return fail(FIJ_INVALID_ARGUMENT, "Partition attribute length differs from the prefix sum's length");
  if (relation->radix_bits != bits) return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching radix bits");
  if ((len + p->grid - 1) / p->grid >= kU32Max)
    return fail(FIJ_INTEGER_OVERFLOW,
                "Relation is too large and causes an integer overflow. Try using more chunks by setting a higher "
                "CUDA grid size");
  if (offsets->chunks != relation->chunks)
    return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets and PartitionedRelation have mismatching chunks");
  if (offsets->radix_bits != bits) return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has mismatching radix bits");
  if (offsets->type != static_cast<uint32_t>(type) || relation->type != static_cast<uint32_t>(type))
    return fail(FIJ_INVALID_ARGUMENT, "tuple types differ");
  const bool contiguous = offsets->local_offsets != nullptr;
  if (contiguous ? offsets->max_chunks != p->grid : offsets->chunks != p->grid)
    return fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets were computed for a different grid size");
  const uint64_t plen = fij_partitioned_relation_padded_len(relation);
  if (relation->relation_len < plen || !relation->relation)
    return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation is smaller than its padded length");
  const uint32_t F = 1u << bits;
  const uint32_t pad = pad_tuples(key_type(type));
  uint64_t chunk_len = 0;
  s = fij_input_chunk_size(type, len, p->grid, &chunk_len);
  if (s != FIJ_OK) return s;
  s = check_chunking(type, len, p->grid);
  if (s != FIJ_OK) return s;
  // Offsets are kept relative to a chunk's start in 32 bits inside the
  // kernels (as in radix_partition.cu:495-511).
  const uint64_t last_chunk = len - chunk_len * (p->grid - 1);
  const uint64_t span = contiguous ? plen + 2ull * pad
                                   : std::max(chunk_len, last_chunk) + static_cast<uint64_t>(pad) * F + 2ull * pad;
  if (span >= kU32Max) return fail(FIJ_INTEGER_OVERFLOW, "Partitioned chunk exceeds 32-bit offsets");
  if (!aligned_to(relation->relation, 2 * key_bytes(key_type(type))))
    return fail(FIJ_INVALID_ARGUMENT, "relation is not aligned to its tuple size");
  const size_t kb = key_bytes(key_type(type));
  if (len && !(aligned_to(partition_attr, kb) && aligned_to(payload_attr, kb)))
    return fail(FIJ_INVALID_ARGUMENT, "columns must be aligned to their element size");
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  const uint64_t* chunk_offsets = contiguous ? offsets->local_offsets : offsets->offsets;
  // gpu_radix_partition.rs:1117-1121: the relation takes over the offsets.
  std::swap(relation->offsets, offsets->offsets);
  PartitionLaunch a{};
  a.type = key_type(type);
  a.keys = partition_attr;
  a.pays = payload_attr;
  a.n = len;
  a.chunks = p->grid;
  a.radix_bits = bits;
  a.ignore_bits = pass_ignore_bits(p->bits, pass);
  a.chunk_offsets = chunk_offsets;
  a.relation = relation->relation;
  a.block = p->block;
  a.lds_bytes = d->limits.lds_bytes_per_block;
  a.buffered = p->algo != FIJ_PARTITION_NC;
  FIJ_HIP(launch_partition(a, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

// --------------------------------------------------------- radix join ----

fij_status fij_radix_join_create(fij_radix_pass pass, fij_radix_bits radix_bits, fij_hashing_scheme scheme,
                                 uint32_t grid_size, uint32_t block_size, fij_radix_join** out) {
  if (!out || grid_size == 0) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (scheme != FIJ_HASH_PERFECT && scheme != FIJ_HASH_LINEAR_PROBING && scheme != FIJ_HASH_BUCKET_CHAINING)
    return fail(FIJ_INVALID_ARGUMENT, "unknown hashing scheme");
  const uint32_t block = std::min<uint32_t>(block_size, 1024);
  fij_status s = check_block(block);
  if (s != FIJ_OK) return s;
  fij_radix_join* j = new fij_radix_join();
  j->pass = pass;
  j->bits = radix_bits;
  j->scheme = scheme;
  j->grid = grid_size;
  j->block = block;
  *out = j;
  return FIJ_OK;
}

fij_status fij_radix_join_destroy(fij_radix_join* j) {
  delete j;
  return FIJ_OK;
}

fij_status fij_join_assign_tasks(const fij_partitioned_relation* probe_rel, uint32_t radix_bits, uint32_t grid_size,
                                 uint32_t* task_assignments, void* stream) {
  if (!probe_rel || !task_assignments || grid_size == 0 || !valid_type(probe_rel->type) ||
      radix_bits != probe_rel->radix_bits || probe_rel->chunks != 1)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const uint64_t plen = fij_partitioned_relation_padded_len(probe_rel);
  if (plen >= kU32Max) return fail(FIJ_INTEGER_OVERFLOW, "probe relation exceeds 32-bit lengths");
  FIJ_HIP(launch_assign_tasks(probe_rel->offsets, plen, pad_tuples(key_type(static_cast<fij_type>(probe_rel->type))),
                              radix_bits, grid_size, task_assignments, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_radix_join_join(fij_radix_join* j, fij_type type, const fij_partitioned_relation* build_rel,
                               const fij_partitioned_relation* probe_rel, int64_t* result_set,
                               uint64_t result_set_len, uint32_t* task_assignments, uint64_t task_assignments_len,
                               fij_join_aggregate* aggregates, void* stream) {
  if (!j || !build_rel || !probe_rel || !valid_type(type) || !task_assignments)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (build_rel->chunks != 1)
    return fail(FIJ_INVALID_ARGUMENT,
                "Chunked build relations are not supported, use a contiguous relation instead");
  if (probe_rel->chunks != 1)
    return fail(FIJ_INVALID_ARGUMENT,
                "Chunked probe relations are not supported, use a contiguous relation instead");
  if (static_cast<uint64_t>(j->grid) + 1 != task_assignments_len)
    return fail(FIJ_INVALID_ARGUMENT, "Task assignement array must have length: grid size + 1");
  if (j->scheme == FIJ_HASH_LINEAR_PROBING)
    return fail(FIJ_UNSUPPORTED, "linear probing is not implemented for the radix join");
  if (result_set && result_set_len < static_cast<uint64_t>(j->grid) * j->block)
    return fail(FIJ_INVALID_ARGUMENT, "result set must hold grid * block partial sums");
  if (build_rel->type != static_cast<uint32_t>(type) || probe_rel->type != static_cast<uint32_t>(type))
    return fail(FIJ_INVALID_ARGUMENT, "tuple types differ");
  uint32_t bits = 0;
  fij_status s = pass_bits(j->bits, j->pass, &bits);
  if (s != FIJ_OK) return s;
  if (build_rel->radix_bits != bits || probe_rel->radix_bits != bits)
    return fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching radix bits");
  const uint64_t blen = fij_partitioned_relation_padded_len(build_rel);
  const uint64_t plen = fij_partitioned_relation_padded_len(probe_rel);
  if (blen >= kU32Max || plen >= kU32Max) return fail(FIJ_INTEGER_OVERFLOW, "relation exceeds 32-bit lengths");
  DeviceInfo* d = nullptr;
  s = device_info(&d);
  if (s != FIJ_OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t pad = pad_tuples(key_type(type));
  FIJ_HIP(launch_assign_tasks(probe_rel->offsets, plen, pad, bits, j->grid, task_assignments, st));
  JoinLaunch a{};
  a.type = key_type(type);
  a.scheme = j->scheme == FIJ_HASH_PERFECT ? JoinScheme::Perfect : JoinScheme::BucketChaining;
  a.build_rel = build_rel->relation;
  a.build_offsets = build_rel->offsets;
  a.build_len = blen;
  a.probe_rel = probe_rel->relation;
  a.probe_offsets = probe_rel->offsets;
  a.probe_len = plen;
  a.radix_bits = bits;
  // cuda_radix_join.rs:207-208: ignore every bit partitioned so far.
  a.ignore_bits = pass_ignore_bits(j->bits, j->pass) + bits;
  a.result_set = result_set;
  a.aggregates = reinterpret_cast<JoinAggregate*>(aggregates);
  a.task_assignments = task_assignments;
  a.grid = j->grid;
  a.block = j->block;
  a.lds_bytes = d->limits.lds_bytes_per_block;
  a.status = d->status;
  FIJ_HIP(launch_join(a, st));
  return FIJ_OK;
}

// ------------------------------------------------- no-partitioning join ----

fij_status fij_np_table_init(fij_type type, void* table, uint64_t entries, void* stream) {
  if (!valid_type(type) || (entries && !table)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (entries == 0) return FIJ_OK;
  FIJ_HIP(launch_np_table_init(key_type(type), table, entries, true, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_relation_fill_null(fij_type type, void* relation, uint64_t tuples, void* stream) {
  if (!valid_type(type) || (tuples && !relation)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (tuples == 0) return FIJ_OK;
  FIJ_HIP(launch_np_table_init(key_type(type), relation, tuples, false, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_np_build(fij_type type, fij_hashing_scheme scheme, int is_selective, void* table, uint64_t entries,
                        const void* keys, const void* payloads, uint64_t len, uint32_t grid_size,
                        uint32_t block_size, void* stream) {
  if (!valid_type(type) || !table || (len && (!keys || !payloads)) || grid_size == 0)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  fij_status s = check_block(block_size);
  if (s != FIJ_OK) return s;
  if (len > entries) return fail(FIJ_INVALID_ARGUMENT, "Hash table is too small for the build data");
  if (scheme == FIJ_HASH_BUCKET_CHAINING) return fail(FIJ_UNSUPPORTED, "bucket chaining is not implemented for the no-partitioning join");
  if (scheme == FIJ_HASH_LINEAR_PROBING && entries < 2)
    return fail(FIJ_INVALID_ARGUMENT, "linear probing needs at least two table entries");
  if (len == 0) return FIJ_OK;
  FIJ_HIP(launch_np_build(key_type(type), scheme == FIJ_HASH_PERFECT ? JoinScheme::Perfect : JoinScheme::LinearProbing,
                          is_selective != 0, table, entries, keys, payloads, len, grid_size, block_size,
                          static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_np_probe_sum(fij_type type, fij_hashing_scheme scheme, const void* table, uint64_t entries,
                            const void* keys, const void* payloads, uint64_t len, uint64_t* result_set,
                            uint64_t result_set_len, fij_join_aggregate* aggregates, uint32_t grid_size,
                            uint32_t block_size, void* stream) {
  if (!valid_type(type) || !table || (len && (!keys || !payloads)) || grid_size == 0)
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  fij_status s = check_block(block_size);
  if (s != FIJ_OK) return s;
  if (scheme == FIJ_HASH_BUCKET_CHAINING) return fail(FIJ_UNSUPPORTED, "bucket chaining is not implemented for the no-partitioning join");
  if (scheme == FIJ_HASH_LINEAR_PROBING && entries < 2)
    return fail(FIJ_INVALID_ARGUMENT, "linear probing needs at least two table entries");
  if (result_set && result_set_len < static_cast<uint64_t>(grid_size) * block_size)
    return fail(FIJ_INVALID_ARGUMENT, "result set must hold grid * block partial sums");
  FIJ_HIP(launch_np_probe(key_type(type), scheme == FIJ_HASH_PERFECT ? JoinScheme::Perfect : JoinScheme::LinearProbing,
                          table, entries, keys, payloads, len, result_set,
                          reinterpret_cast<JoinAggregate*>(aggregates), grid_size, block_size,
                          static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

// ---------------------------------------------------- data generation ----

fij_status fij_gen_permutation(fij_type type, void* out, uint64_t n, uint64_t seed, int64_t base, void* stream) {
  if (!valid_type(type) || (n && !out)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (n == 0) return FIJ_OK;
  FIJ_HIP(launch_gen_perm_keys(key_type(type), out, n, seed, base, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_gen_uniform(fij_type type, void* out, uint64_t n, uint64_t seed, uint64_t stream_id, int64_t lo,
                           uint64_t range, void* stream) {
  if (!valid_type(type) || (n && !out)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (n == 0) return FIJ_OK;
  FIJ_HIP(launch_gen_uniform(key_type(type), out, n, seed, stream_id, lo, range, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_gen_zipf(fij_type type, void* out, uint64_t n, uint64_t num_ranks, double exponent, uint64_t seed,
                        uint64_t stream_id, uint64_t perm_seed, int64_t base, void* stream) {
  if (!valid_type(type) || (n && !out) || num_ranks == 0 || !(exponent >= 0.0))
    return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (n == 0) return FIJ_OK;
  FIJ_HIP(launch_gen_zipf(key_type(type), out, n, num_ranks, exponent, seed, stream_id, perm_seed, base,
                          static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_gen_iota(fij_type type, void* out, uint64_t n, int64_t base, void* stream) {
  if (!valid_type(type) || (n && !out)) return fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (n == 0) return FIJ_OK;
  FIJ_HIP(launch_gen_iota(key_type(type), out, n, base, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

}  // extern "C"
