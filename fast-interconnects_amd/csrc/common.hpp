// Device-side building blocks shared by every kernel of the library:
// tuple layout, the radix partition function, and wave64 primitives
// (ballot/mbcnt ranking, DPP scans) for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fij {

// Layout constants kept identical to the reference
// (sql-ops/build.rs:44-72, sql-ops/src/partition/partitioned_relation.rs:26-29):
// partitions are padded by 128 bytes in front, and buffered writes are
// aligned to 128-byte lines.
constexpr uint32_t kAlignBytes = 128;
constexpr uint32_t kPaddingBytes = 128;
constexpr uint32_t kWave = 64;

// Row tuple {key, value}; 8 bytes for i32, 16 bytes for i64
// (sql-ops/include/gpu_common.h:50-103).
template <typename K>
struct alignas(2 * sizeof(K)) Tuple {
  K key;
  K value;
};

template <typename K>
__host__ __device__ constexpr K null_key() {
  return static_cast<K>(-1);
}

template <typename K>
__host__ __device__ constexpr uint32_t padding_len() {
  return kPaddingBytes / sizeof(Tuple<K>);
}

// Radix digit of a key: bits [ignore_bits, ignore_bits + radix_bits) of the
// key widened to 64 bits (sql-ops/include/gpu_radix_partition.h:121-127).
template <typename K>
__device__ __forceinline__ uint32_t key_to_partition(K key, uint64_t mask, uint32_t ignore_bits) {
  return static_cast<uint32_t>((static_cast<uint64_t>(static_cast<int64_t>(key)) & mask) >> ignore_bits);
}

__host__ __device__ inline uint64_t radix_mask(uint32_t radix_bits, uint32_t ignore_bits) {
  return static_cast<uint64_t>((1u << radix_bits) - 1u) << ignore_bits;
}

// Canonical input chunk length: ceil(n / chunks) rounded down to a multiple of
// 128 bytes of keys; the last chunk takes the remainder
// (sql-ops/src/partition/partition_input_chunk.rs:24-35).
template <typename K>
__host__ __device__ inline uint64_t canonical_chunk_len(uint64_t n, uint32_t chunks) {
  const uint64_t align_mask = ~(static_cast<uint64_t>(kAlignBytes / sizeof(K)) - 1ull);
  return ((n + chunks - 1) / chunks) & align_mask;
}

// Multiply-shift hash (sql-ops/include/gpu_common.h:143-168).
__device__ __forceinline__ uint32_t mult_shift_hash32(uint32_t v, uint32_t log2_buckets) {
  return (v * 2654435769u) >> (32u - log2_buckets);
}
__device__ __forceinline__ uint64_t mult_shift_hash64(uint64_t v, uint32_t log2_buckets) {
  return (v * 11400714819323198485ull) >> (64u - log2_buckets);
}
template <typename K>
__device__ __forceinline__ uint32_t mult_shift_hash(K v, uint32_t log2_buckets) {
  if constexpr (sizeof(K) == 4)
    return mult_shift_hash32(static_cast<uint32_t>(v), log2_buckets);
  else
    return static_cast<uint32_t>(mult_shift_hash64(static_cast<uint64_t>(v), log2_buckets));
}

// ---------------------------------------------------------------- wave64 ----

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` in lanes below the calling lane (mbcnt).
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t ballot(bool pred) { return __ballot(pred); }

// Adds one per active lane to counters[digit] and returns each lane's rank.
// Lanes that share the digit of the wave's first active lane are aggregated
// into a single LDS atomic (ballot + mbcnt), which removes the same-address
// serialisation of skewed inputs; the remaining lanes use one atomic each.
__device__ __forceinline__ uint32_t wave_rank_add(uint32_t* counters, uint32_t digit, bool active) {
  const uint64_t act = ballot(active);
  if (act == 0) return 0;
  const int leader = __ffsll(static_cast<unsigned long long>(act)) - 1;
  const uint32_t lead_digit = __builtin_amdgcn_readlane(digit, leader);
  const bool same = active && digit == lead_digit;
  const uint64_t peers = ballot(same);
  uint32_t base = 0;
  if (static_cast<int>(lane_id()) == leader)
    base = atomicAdd(&counters[lead_digit], static_cast<uint32_t>(__popcll(peers)));
  base = __builtin_amdgcn_readlane(base, leader);
  uint32_t rank = 0;
  if (same)
    rank = base + mbcnt(peers);
  else if (active)
    rank = atomicAdd(&counters[digit], 1u);
  return rank;
}

// Count-only variant of wave_rank_add (histograms).
__device__ __forceinline__ void wave_count_add(uint32_t* counters, uint32_t digit, bool active) {
  const uint64_t act = ballot(active);
  if (act == 0) return;
  const int leader = __ffsll(static_cast<unsigned long long>(act)) - 1;
  const uint32_t lead_digit = __builtin_amdgcn_readlane(digit, leader);
  const bool same = active && digit == lead_digit;
  const uint64_t peers = ballot(same);
  if (static_cast<int>(lane_id()) == leader)
    atomicAdd(&counters[lead_digit], static_cast<uint32_t>(__popcll(peers)));
  else if (active && !same)
    atomicAdd(&counters[digit], 1u);
}

// DPP inclusive scan over the 64 lanes of a wave: row_shr:1,2,4,8 builds
// 16-lane row scans, row_bcast:15 and row_bcast:31 carry row totals upward.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(x), CTRL, ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x) {
  const uint32_t lo = dpp_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(x));
  const uint32_t hi = dpp_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(x >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  x += dpp_u32<0x111, 0xf>(x);
  x += dpp_u32<0x112, 0xf>(x);
  x += dpp_u32<0x114, 0xf>(x);
  x += dpp_u32<0x118, 0xf>(x);
  x += dpp_u32<0x142, 0xa>(x);
  x += dpp_u32<0x143, 0xc>(x);
  return x;
}
__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t x) {
  x += dpp_u64<0x111, 0xf>(x);
  x += dpp_u64<0x112, 0xf>(x);
  x += dpp_u64<0x114, 0xf>(x);
  x += dpp_u64<0x118, 0xf>(x);
  x += dpp_u64<0x142, 0xa>(x);
  x += dpp_u64<0x143, 0xc>(x);
  return x;
}

// Inclusive max-scan over the 64 lanes (same DPP pattern as the sum scan;
// lanes a DPP step does not reach combine with 0).
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t x) {
  x = max(x, dpp_u32<0x111, 0xf>(x));
  x = max(x, dpp_u32<0x112, 0xf>(x));
  x = max(x, dpp_u32<0x114, 0xf>(x));
  x = max(x, dpp_u32<0x118, 0xf>(x));
  x = max(x, dpp_u32<0x142, 0xa>(x));
  x = max(x, dpp_u32<0x143, 0xc>(x));
  return x;
}

template <typename U>
__device__ __forceinline__ U wave_broadcast_last(U x) {
  if constexpr (sizeof(U) == 4) {
    return static_cast<U>(__builtin_amdgcn_readlane(static_cast<uint32_t>(x), 63));
  } else {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(x), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(x >> 32), 63);
    return (static_cast<U>(hi) << 32) | lo;
  }
}

// Exclusive scan of data[0..n) in LDS by the whole block, in place.  Each
// thread scans a contiguous run serially, runs are combined with wave DPP
// scans and a scan of per-wave totals.  Returns the total.  Must be called by
// all threads of the block; `wave_tmp` holds at least blockDim.x / 64 + 1
// entries.  Ends with a barrier.
template <typename U>
__device__ U block_exclusive_scan(U* data, uint32_t n, U* wave_tmp) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  const uint32_t nwaves = nthreads / kWave;
  const uint32_t per = (n + nthreads - 1) / nthreads;
  const uint32_t lo = min(tid * per, n);
  const uint32_t hi = min(lo + per, n);
  U local = 0;
  for (uint32_t i = lo; i < hi; ++i) local += data[i];
  const U incl = wave_inclusive_scan(local);
  const uint32_t w = tid / kWave;
  if (lane_id() == kWave - 1) wave_tmp[w] = incl;
  __syncthreads();
  if (tid == 0) {
    U run = 0;
    for (uint32_t i = 0; i < nwaves; ++i) {
      U t = wave_tmp[i];
      wave_tmp[i] = run;
      run += t;
    }
    wave_tmp[nwaves] = run;
  }
  __syncthreads();
  U run = wave_tmp[w] + incl - local;
  for (uint32_t i = lo; i < hi; ++i) {
    U t = data[i];
    data[i] = run;
    run += t;
  }
  const U total = wave_tmp[nwaves];
  __syncthreads();
  return total;
}

// Dynamic LDS a kernel may request so that dynamic + static LDS stays within
// `total` bytes per block (rounded down to 16 bytes); 0 if it cannot launch.
inline uint32_t dynamic_lds_budget(const void* kernel, uint32_t total) {
  hipFuncAttributes attr{};
  if (hipFuncGetAttributes(&attr, kernel) != hipSuccess) return 0;
  if (attr.sharedSizeBytes >= total) return 0;
  return static_cast<uint32_t>((total - attr.sharedSizeBytes) & ~static_cast<size_t>(15));
}

// ------------------------------------------------------------- memory ----

// Global-memory pointers.  A pointer read from memory (a descriptor of a
// batched launch) is generic, and accesses through it would be flat
// instructions, which count on both vmcnt and lgkmcnt; as_global makes
// them global loads and stores.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> as_global(T* p) {
  return (global_ptr<T>)(p);
}

// Wave-uniform load of a word or a struct (a multiple of 4 bytes) through
// the scalar data cache, for data no kernel in flight writes.
template <typename U>
__device__ __forceinline__ U scalar_load(const U* ptr) {
  static_assert(sizeof(U) % 4 == 0, "scalar loads move whole words");
  U out;
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
  const __attribute__((address_space(4))) uint32_t* in = (const __attribute__((address_space(4))) uint32_t*)(ptr);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(U) / 4; ++i) o[i] = in[i];
  return out;
}

template <typename K>
struct Vec2;
template <>
struct Vec2<int32_t> {
  typedef int32_t type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<int64_t> {
  typedef int64_t type __attribute__((ext_vector_type(2)));
};

// 8- or 16-byte tuple store as one vector store.
template <typename K>
__device__ __forceinline__ void store_tuple(Tuple<K>* dst, K key, K value) {
  typename Vec2<K>::type v;
  v.x = key;
  v.y = value;
  *reinterpret_cast<typename Vec2<K>::type*>(dst) = v;
}
template <typename K>
__device__ __forceinline__ void store_tuple_nt(Tuple<K>* dst, K key, K value) {
  typename Vec2<K>::type v;
  v.x = key;
  v.y = value;
  __builtin_nontemporal_store(v, reinterpret_cast<typename Vec2<K>::type*>(dst));
}
template <typename K>
__device__ __forceinline__ void store_tuple(global_ptr<Tuple<K>> dst, K key, K value) {
  typename Vec2<K>::type v;
  v.x = key;
  v.y = value;
  *(global_ptr<typename Vec2<K>::type>)(dst) = v;
}
template <typename K>
__device__ __forceinline__ void store_tuple_nt(global_ptr<Tuple<K>> dst, K key, K value) {
  typename Vec2<K>::type v;
  v.x = key;
  v.y = value;
  __builtin_nontemporal_store(v, (global_ptr<typename Vec2<K>::type>)(dst));
}
template <typename K>
__device__ __forceinline__ Tuple<K> load_tuple(const Tuple<K>* src) {
  typename Vec2<K>::type v = *reinterpret_cast<const typename Vec2<K>::type*>(src);
  Tuple<K> t;
  t.key = v.x;
  t.value = v.y;
  return t;
}
template <typename K>
__device__ __forceinline__ Tuple<K> load_tuple_nt(const Tuple<K>* src) {
  typename Vec2<K>::type v = __builtin_nontemporal_load(reinterpret_cast<const typename Vec2<K>::type*>(src));
  Tuple<K> t;
  t.key = v.x;
  t.value = v.y;
  return t;
}

}  // namespace fij
