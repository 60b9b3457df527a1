// Runtime plumbing of the C ABI: device and pinned host buffers, copies,
// fills, streams and events.  These stand where the reference's host code
// uses rustacuda's DeviceBuffer / LockedBuffer / Stream / Event and
// numa-gpu's Mem types (numa-gpu/src/runtime/memory.rs:36-95), so that a
// caller needs nothing but this library and the HIP driver.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/fij.h"
#include "internal.hpp"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

namespace {

fij_status hip_fail(hipError_t e, const char* what) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  return fij::set_error(e == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, buf);
}

#define RT_HIP(expr)                                  \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

}  // namespace

extern "C" {

fij_status fij_mem_alloc(uint64_t bytes, void** out) {
  if (!out) return fij::set_error(FIJ_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return FIJ_OK;
  RT_HIP(hipMalloc(out, bytes));
  return FIJ_OK;
}

fij_status fij_mem_free(void* ptr) {
  if (ptr) RT_HIP(hipFree(ptr));
  return FIJ_OK;
}

fij_status fij_mem_alloc_host(uint64_t bytes, void** out) {
  if (!out) return fij::set_error(FIJ_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return FIJ_OK;
  RT_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return FIJ_OK;
}

fij_status fij_mem_free_host(void* ptr) {
  if (ptr) RT_HIP(hipHostFree(ptr));
  return FIJ_OK;
}

fij_status fij_memcpy(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return FIJ_OK;
  if (!dst || !src) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL copy operand");
  hipStream_t st = static_cast<hipStream_t>(stream);
  RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st));
  RT_HIP(hipStreamSynchronize(st));
  return FIJ_OK;
}

fij_status fij_memcpy_async(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return FIJ_OK;
  if (!dst || !src) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL copy operand");
  RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_mem_fill_u32(void* ptr, uint32_t value, uint64_t count, void* stream) {
  if (count == 0) return FIJ_OK;
  if (!ptr) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL buffer");
  RT_HIP(fij::launch_fill_u32(static_cast<uint32_t*>(ptr), count, value, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_mem_fill_u64(void* ptr, uint64_t value, uint64_t count, void* stream) {
  if (count == 0) return FIJ_OK;
  if (!ptr) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL buffer");
  RT_HIP(fij::launch_fill_u64(static_cast<uint64_t*>(ptr), count, value, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_copy_kernel(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes == 0) return FIJ_OK;
  if (!dst || !src) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL buffer");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "copy kernel needs 16-byte aligned buffers and sizes");
  RT_HIP(fij::launch_copy(dst, src, bytes / 16, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_calib_line_scatter(void* dst, const void* src, uint64_t bytes, uint32_t partitions, uint32_t chunks,
                                  void* stream) {
  if (!dst || !src) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL buffer");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "line scatter needs 16-byte aligned buffers and sizes");
  const uint64_t words = bytes / 16;
  if (chunks == 0 || partitions == 0 || words % chunks)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "line scatter: the chunks must divide the words");
  const uint64_t chunk = words / chunks;
  // whole groups of eight 1024-word steps per block, whole lines per piece
  if (chunk == 0 || chunk % (1024 * 8) || (chunk / 8) % partitions)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "line scatter: a chunk must be a multiple of 8192 words and of "
                                                "8 x partitions words");
  RT_HIP(fij::launch_line_scatter(dst, src, words, partitions, chunks, static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_calib_run_gather(const void* src, uint64_t bytes, uint32_t part_words, uint32_t f2, uint32_t run_words,
                                uint32_t shift, uint32_t* checksum, uint32_t grid, void* stream) {
  if (!src || !checksum) return fij::set_error(FIJ_INVALID_ARGUMENT, "NULL buffer");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "run gather needs a 16-byte aligned buffer and size");
  const uint64_t words = bytes / 16;
  if (part_words == 0 || words % part_words || f2 == 0 || part_words % f2 || grid == 0)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "run gather: partitions must divide the words, f2 a partition");
  const uint32_t item = part_words / f2;
  // every run stays inside its tile: shift <= run_words
  if (item > 1024 || run_words == 0 || item % run_words || shift > run_words)
    return fij::set_error(FIJ_INVALID_ARGUMENT, "run gather: items of at most 1024 words, whole runs, "
                                                "shift <= run length");
  RT_HIP(fij::launch_run_gather(src, words, part_words, f2, run_words, shift, checksum,
                                static_cast<hipStream_t>(stream), grid));
  return FIJ_OK;
}

fij_status fij_stream_create(void** out) {
  if (!out) return fij::set_error(FIJ_INVALID_ARGUMENT, "out is NULL");
  hipStream_t s = nullptr;
  RT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return FIJ_OK;
}

fij_status fij_stream_destroy(void* stream) {
  if (stream) RT_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_stream_synchronize(void* stream) {
  RT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_device_synchronize(void) {
  RT_HIP(hipDeviceSynchronize());
  return FIJ_OK;
}

fij_status fij_event_create(void** out) {
  if (!out) return fij::set_error(FIJ_INVALID_ARGUMENT, "out is NULL");
  hipEvent_t e = nullptr;
  RT_HIP(hipEventCreate(&e));
  *out = e;
  return FIJ_OK;
}

fij_status fij_event_destroy(void* event) {
  if (event) RT_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return FIJ_OK;
}

fij_status fij_event_record(void* event, void* stream) {
  RT_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return FIJ_OK;
}

fij_status fij_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!ms) return fij::set_error(FIJ_INVALID_ARGUMENT, "ms is NULL");
  RT_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  RT_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return FIJ_OK;
}

fij_status fij_device_count(int* count) {
  if (!count) return fij::set_error(FIJ_INVALID_ARGUMENT, "count is NULL");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e == hipErrorNoDevice) {
    *count = 0;
    return FIJ_OK;
  }
  RT_HIP(e);
  return FIJ_OK;
}

}  // extern "C"
