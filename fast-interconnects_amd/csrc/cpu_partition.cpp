// Host radix partitioner: pass 1 on the CPU, producing exactly the
// contiguous layout the GPU's pass 1 writes (one range per partition, `pad`
// tuples of padding in front of each, off[p] = sum of the counts of the
// partitions before p + (p + 1) * pad), so the GPU's pass 2 and join consume
// either.  The reference's CPU-partitioned radix join runs its pass 1 on the
// CPU the same way, one chunk per thread
// (radix-join/src/execution_methods/cpu_partitioned_radix_join.rs:238-290,
// 455-555; sql-ops/cpputils/radix_partition.cpp:228-267 histogram,
// :448-557 software write-combining scatter); its chunked layout has one
// range per (chunk, partition), this one per partition with per-thread
// write cursors.
//
// Each thread histograms its chunk; the counts are scanned into partition
// and per-thread offsets; each thread then scatters its chunk through
// per-partition 64-byte write-combining buffers, each standing for one
// 64-byte line of the output.  A full line is written with non-temporal
// stores (no read for ownership); the partial lines at either end of a
// thread's range in a partition, which other threads or padding share, are
// copied.  Order within a partition is that of the threads, each thread's
// tuples in input order.
#include <emmintrin.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace fij {

namespace {

template <typename K>
inline uint32_t host_digit(K key, uint64_t mask, uint32_t ignore) {
  return static_cast<uint32_t>((static_cast<uint64_t>(static_cast<int64_t>(key)) & mask) >> ignore);
}

template <typename K>
struct HostTuple {
  K key;
  K value;
};

template <typename F>
void parallel_for(uint32_t threads, F&& f) {
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(f, t);
  f(0u);
  for (auto& th : pool) th.join();
}

template <typename K>
void partition_t(const K* keys, const K* pays, uint64_t n, uint32_t bits, uint32_t ignore, uint32_t threads,
                 void* out_v, uint64_t* off) {
  using T = HostTuple<K>;
  constexpr uint32_t L = 64 / sizeof(T);  // tuples per write-combining buffer
  const uint32_t F = 1u << bits;
  const uint64_t pad = 128 / sizeof(T);  // 128-byte padding, as on the GPU
  const uint64_t mask = (static_cast<uint64_t>(F) - 1) << ignore;
  T* out = static_cast<T*>(out_v);
  threads = std::max<uint32_t>(1, std::min<uint64_t>(threads, std::max<uint64_t>(1, n / 4096)));
  const uint64_t per = (n + threads - 1) / threads;
  std::vector<uint64_t> hist(static_cast<size_t>(threads) * F, 0);
  parallel_for(threads, [&](uint32_t t) {
    const uint64_t b = std::min(n, per * t), e = std::min(n, b + per);
    uint64_t* h = &hist[static_cast<size_t>(t) * F];
    for (uint64_t i = b; i < e; ++i) ++h[host_digit(keys[i], mask, ignore)];
  });
  // partition offsets, then each thread's cursor inside each partition
  uint64_t run = 0;
  for (uint32_t p = 0; p < F; ++p) {
    off[p] = run + (p + 1) * pad;
    uint64_t c = off[p];
    for (uint32_t t = 0; t < threads; ++t) {
      uint64_t& h = hist[static_cast<size_t>(t) * F + p];
      const uint64_t cnt = h;
      h = c;  // becomes thread t's cursor in partition p
      c += cnt;
      run += cnt;
    }
  }
  // A buffer stands for one 64-byte line of memory: tuple i of the output
  // lies in line (i + sh) / L, sh = the output's offset into its first line
  // in tuples (any output aligned to its tuple size; otherwise every line is
  // copied).
  const uintptr_t addr = reinterpret_cast<uintptr_t>(out);
  const bool lines = addr % sizeof(T) == 0;
  const uint32_t sh = lines ? static_cast<uint32_t>(addr % 64 / sizeof(T)) : 0u;
  parallel_for(threads, [&](uint32_t t) {
    const uint64_t b = std::min(n, per * t), e = std::min(n, b + per);
    uint64_t* cur = &hist[static_cast<size_t>(t) * F];  // first tuple of the buffered line
    std::vector<uint8_t> storage(static_cast<size_t>(F) * 64 + 64);  // F lines, 64-byte aligned
    T* buf = reinterpret_cast<T*>(storage.data() + (-reinterpret_cast<uintptr_t>(storage.data()) & 63));
    std::vector<uint32_t> fill(F), lo(F);  // buffered slots [lo, fill) of the line
    for (uint32_t d = 0; d < F; ++d) {
      lo[d] = fill[d] = lines ? static_cast<uint32_t>((cur[d] + sh) % L) : 0u;
      cur[d] -= lo[d];  // out + cur[d] is 64-byte aligned when lines
    }
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t d = host_digit(keys[i], mask, ignore);
      T* slot = buf + static_cast<size_t>(d) * L;
      slot[fill[d]] = T{keys[i], pays[i]};
      if (++fill[d] == L) {
        if (lo[d] == 0 && lines) {
          const __m128i* src = reinterpret_cast<const __m128i*>(slot);
          __m128i* dst = reinterpret_cast<__m128i*>(out + cur[d]);
          for (uint32_t k = 0; k < 4; ++k) _mm_stream_si128(dst + k, _mm_load_si128(src + k));
        } else {
          memcpy(out + cur[d] + lo[d], slot + lo[d], sizeof(T) * (L - lo[d]));
        }
        cur[d] += L;
        fill[d] = lo[d] = 0;
      }
    }
    for (uint32_t d = 0; d < F; ++d)
      if (fill[d] > lo[d])
        memcpy(out + cur[d] + lo[d], buf + static_cast<size_t>(d) * L + lo[d], sizeof(T) * (fill[d] - lo[d]));
    _mm_sfence();  // the streamed lines are visible before the threads join
  });
}

// Per-chunk histograms of the canonical input chunks (chunk_len keys each,
// the last chunk the rest), computed by up to `threads` threads.
template <typename K>
std::vector<uint64_t> chunk_histograms(const K* keys, uint64_t n, uint32_t chunks, uint64_t chunk_len, uint32_t F,
                                       uint64_t mask, uint32_t ignore, uint32_t threads) {
  std::vector<uint64_t> hist(static_cast<size_t>(chunks) * F, 0);
  threads = std::max(1u, std::min(threads, chunks));
  parallel_for(threads, [&](uint32_t t) {
    for (uint32_t c = t; c < chunks; c += threads) {
      const uint64_t b = std::min(n, chunk_len * c);
      const uint64_t e = c + 1 == chunks ? n : std::min(n, b + chunk_len);
      uint64_t* h = &hist[static_cast<size_t>(c) * F];
      for (uint64_t i = b; i < e; ++i) ++h[host_digit(keys[i], mask, ignore)];
    }
  });
  return hist;
}

// The GPU prefix sums' offsets computed on the host (histogram mixing:
// sql-ops/src/partition/cpu_radix_partition.rs:46-62).  Chunked:
// offsets[c][p] = (chunk_len + F pad) c + sum_{q<p} (hist_c[q] + pad) + pad
// (radix_partition.cpp:228-267 = radix_partition.cu:40-96).  Contiguous:
// local[c][p] = sum over (p', c') before (p, c) in partition-major order of
// hist + (p + 1) pad, offsets[p] = local[0][p] (radix_partition.cu:104-199).
template <typename K>
void prefix_sum_t(const K* keys, uint64_t n, uint32_t chunks, uint64_t chunk_len, uint32_t bits, uint32_t ignore,
                  bool contiguous, uint32_t threads, uint64_t* offsets, uint64_t* local) {
  const uint32_t F = 1u << bits;
  const uint64_t pad = 128 / (2 * sizeof(K));
  const uint64_t mask = (static_cast<uint64_t>(F) - 1) << ignore;
  const std::vector<uint64_t> hist = chunk_histograms(keys, n, chunks, chunk_len, F, mask, ignore, threads);
  if (!contiguous) {
    for (uint32_t c = 0; c < chunks; ++c) {
      uint64_t o = (chunk_len + pad * F) * c + pad;
      for (uint32_t p = 0; p < F; ++p) {
        offsets[static_cast<size_t>(c) * F + p] = o;
        o += hist[static_cast<size_t>(c) * F + p] + pad;
      }
    }
    return;
  }
  uint64_t run = 0;
  for (uint32_t p = 0; p < F; ++p)
    for (uint32_t c = 0; c < chunks; ++c) {
      local[static_cast<size_t>(c) * F + p] = run + (p + 1) * pad;
      run += hist[static_cast<size_t>(c) * F + p];
    }
  for (uint32_t p = 0; p < F; ++p) offsets[p] = local[p];
}

}  // namespace

void cpu_prefix_sum(KeyType t, const void* keys, uint64_t n, uint32_t chunks, uint64_t chunk_len, uint32_t bits,
                    uint32_t ignore, bool contiguous, uint32_t threads, uint64_t* offsets, uint64_t* local) {
  if (t == KeyType::I32)
    prefix_sum_t<int32_t>(static_cast<const int32_t*>(keys), n, chunks, chunk_len, bits, ignore, contiguous, threads,
                          offsets, local);
  else
    prefix_sum_t<int64_t>(static_cast<const int64_t*>(keys), n, chunks, chunk_len, bits, ignore, contiguous, threads,
                          offsets, local);
}

void cpu_partition_contiguous(KeyType t, const void* keys, const void* pays, uint64_t n, uint32_t bits,
                              uint32_t ignore, uint32_t threads, void* out, uint64_t* offsets) {
  if (t == KeyType::I32)
    partition_t<int32_t>(static_cast<const int32_t*>(keys), static_cast<const int32_t*>(pays), n, bits, ignore,
                         threads, out, offsets);
  else
    partition_t<int64_t>(static_cast<const int64_t*>(keys), static_cast<const int64_t*>(pays), n, bits, ignore,
                         threads, out, offsets);
}

}  // namespace fij

// ---------------------------------------------------------------- C ABI ----

#include <cstdarg>
#include <cstdio>

#include "../../include/fij.h"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

namespace {

fij_status cpu_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return fij::set_error(s, buf);
}

uint32_t host_threads(uint32_t t) { return t ? t : std::max(1u, std::thread::hardware_concurrency()); }

}  // namespace

extern "C" {

fij_status fij_cpu_prefix_sum(fij_histogram_algorithm algorithm, fij_type type, const void* keys, uint64_t len,
                              uint32_t radix_bits, uint32_t ignore_bits, uint32_t threads,
                              fij_partition_offsets* offsets) {
  if (!offsets || (type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) || (len && !keys) || !offsets->offsets ||
      offsets->max_chunks == 0)
    return cpu_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (offsets->radix_bits != radix_bits || radix_bits > 16)
    return cpu_fail(FIJ_INVALID_ARGUMENT, "PartitionedRelation has mismatching radix bits");
  if (offsets->type != static_cast<uint32_t>(type))
    return cpu_fail(FIJ_INVALID_ARGUMENT, "PartitionOffsets has a different tuple type");
  const bool contiguous = algorithm == FIJ_HISTOGRAM_CONTIGUOUS;
  if (contiguous && !offsets->local_offsets)
    return cpu_fail(FIJ_INVALID_ARGUMENT, "Contiguous offsets need local offsets for every chunk");
  const uint32_t chunks = offsets->max_chunks;
  uint64_t chunk_len = 0;
  const fij_status s = fij_input_chunk_size(type, len, chunks, &chunk_len);
  if (s != FIJ_OK) return s;
  if (chunks > 1 && chunk_len * (chunks - 1) > len)
    return cpu_fail(FIJ_INVALID_ARGUMENT, "%u chunks of %llu keys leave the last chunk of %llu keys empty", chunks,
                    static_cast<unsigned long long>(chunk_len), static_cast<unsigned long long>(len));
  fij::cpu_prefix_sum(type == FIJ_TYPE_I32 ? fij::KeyType::I32 : fij::KeyType::I64, keys, len, chunks, chunk_len,
                      radix_bits, ignore_bits, contiguous, host_threads(threads), offsets->offsets,
                      offsets->local_offsets);
  offsets->chunks = contiguous ? 1 : chunks;
  offsets->data_len = len;
  return FIJ_OK;
}

fij_status fij_cpu_radix_partition(fij_type type, const void* keys, const void* payloads, uint64_t len,
                                   uint32_t radix_bits, uint32_t ignore_bits, uint32_t threads, void* relation,
                                   uint64_t* offsets) {
  if ((type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) || (len && (!keys || !payloads)) || !relation || !offsets ||
      radix_bits > 16)
    return cpu_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  fij::cpu_partition_contiguous(type == FIJ_TYPE_I32 ? fij::KeyType::I32 : fij::KeyType::I64, keys, payloads, len,
                                radix_bits, ignore_bits, host_threads(threads), relation, offsets);
  return FIJ_OK;
}

}  // extern "C"
