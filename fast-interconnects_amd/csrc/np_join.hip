// No-partitioning hash join in HBM: the comparison point for the radix join.
//
// Semantics follow sql-ops/cudautils/no_partitioning_join.cu: a table of a
// power-of-two number of {key, value} entries initialised to the null key;
// linear probing from the multiply-shift hash (:39-132 build, :134-249
// probe, which keeps probing after a match so duplicate build keys all
// count); perfect hashing uses the key as the index (:251-379).  Partial
// SUM(probe.value) per thread as in the reference, plus per-block
// {count, SUM(build.value), SUM(probe.value)}.
#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {

template <typename K>
__global__ void np_init_kernel(Tuple<K>* table, uint64_t entries, K value) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < entries; i += stride)
    store_tuple<K>(table + i, null_key<K>(), value);
}

__device__ __forceinline__ uint32_t log2_floor(uint64_t x) { return 63u - __builtin_clzll(x); }

// i32: key and payload are installed together with one 64-bit CAS
// (no_partitioning_join.cu:49-59); i64: CAS on the key, then the payload.
template <typename K>
__device__ __forceinline__ void lp_insert(Tuple<K>* table, uint32_t lg, K key, K pay) {
  const uint64_t entries = 1ull << lg;
  const uint64_t mask = entries - 1;
  uint64_t idx = mult_shift_hash<K>(key, lg);
  for (uint64_t a = 0; a < entries; ++a, idx = (idx + 1) & mask) {
    if constexpr (sizeof(K) == 4) {
      const unsigned long long want =
          (static_cast<unsigned long long>(static_cast<uint32_t>(key))) |
          (static_cast<unsigned long long>(static_cast<uint32_t>(pay)) << 32);
      const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(table + idx), ~0ull, want);
      if (old == ~0ull) return;
    } else {
      const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(&table[idx].key), ~0ull,
                                               static_cast<unsigned long long>(key));
      if (old == ~0ull) {
        table[idx].value = pay;
        return;
      }
    }
  }
}

template <typename K>
__global__ __launch_bounds__(1024) void np_build_lp_kernel(Tuple<K>* table, uint64_t entries, const K* keys, const K* pays, uint64_t n) {
  const uint32_t lg = log2_floor(entries);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K key = keys[i];
    if (key != null_key<K>()) lp_insert<K>(table, lg, key, pays[i]);  // null keys never join
  }
}

template <typename K>
__global__ __launch_bounds__(1024) void np_build_perfect_kernel(Tuple<K>* table, uint64_t entries, const K* keys, const K* pays,
                                        uint64_t n, bool selective) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const K key = keys[i];
    if (selective && key == null_key<K>()) continue;
    const uint64_t idx = static_cast<uint64_t>(static_cast<int64_t>(key));
    if (idx < entries) store_tuple<K>(table + idx, key, pays[i]);
  }
}

__device__ void np_block_aggregate(JoinAggregate* out, uint64_t c, uint64_t sb, uint64_t sp) {
  __shared__ uint64_t red[3][16];
  const uint32_t w = threadIdx.x / kWave;
  const uint64_t wc = wave_broadcast_last(wave_inclusive_scan(c));
  const uint64_t wb = wave_broadcast_last(wave_inclusive_scan(sb));
  const uint64_t wp = wave_broadcast_last(wave_inclusive_scan(sp));
  if (lane_id() == 0) {
    red[0][w] = wc;
    red[1][w] = wb;
    red[2][w] = wp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0, d = 0;
    for (uint32_t i = 0; i < blockDim.x / kWave; ++i) {
      a += red[0][i];
      b += red[1][i];
      d += red[2][i];
    }
    out[blockIdx.x].count += a;
    out[blockIdx.x].sum_build += b;
    out[blockIdx.x].sum_probe += d;
  }
}

template <typename K, bool PERFECT>
__global__ __launch_bounds__(1024) void np_probe_kernel(const Tuple<K>* table, uint64_t entries, const K* keys,
                                                        const K* pays, uint64_t n, uint64_t* result_set,
                                                        JoinAggregate* aggregates) {
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t cnt = 0, sb = 0, sp = 0;
  if (PERFECT) {
    for (uint64_t i = gid; i < n; i += stride) {
      const K key = keys[i];
      const uint64_t idx = static_cast<uint64_t>(static_cast<int64_t>(key));
      if (idx < entries) {
        const Tuple<K> e = load_tuple<K>(table + idx);
        if (e.key == key) {
          cnt += 1;
          sb += static_cast<uint64_t>(static_cast<int64_t>(e.value));
          sp += static_cast<uint64_t>(static_cast<int64_t>(pays[i]));
        }
      }
    }
  } else {
    const uint32_t lg = log2_floor(entries);
    const uint64_t mask = (1ull << lg) - 1;
    for (uint64_t i = gid; i < n; i += stride) {
      const K key = keys[i];
      const K pay = pays[i];
      if (key == null_key<K>()) continue;  // would match every empty slot
      uint64_t idx = mult_shift_hash<K>(key, lg);
      for (uint64_t a = 0; a <= mask; ++a, idx = (idx + 1) & mask) {
        const Tuple<K> e = load_tuple<K>(table + idx);
        if (e.key == key) {
          cnt += 1;
          sb += static_cast<uint64_t>(static_cast<int64_t>(e.value));
          sp += static_cast<uint64_t>(static_cast<int64_t>(pay));
        } else if (e.key == null_key<K>()) {
          break;
        }
      }
    }
  }
  if (result_set) result_set[gid] += sp;
  if (aggregates) np_block_aggregate(aggregates, cnt, sb, sp);
}

// Probe of a resident linear-probing table of 2^lg entries (a served build
// side that fits the GPU's cache): U probes in flight per thread, their first
// slots loaded together; every match counts, probing ends at a free slot.
// Block b writes (not adds) its {count, SUM(build.value), SUM(probe.value)}
// to out[b].
template <typename K, uint32_t U>
__global__ __launch_bounds__(256) void np_probe_resident_kernel(const Tuple<K>* table, uint32_t lg, const K* keys,
                                                                const K* pays, uint64_t n, JoinAggregate* out) {
  const uint64_t mask = (1ull << lg) - 1;
  const uint64_t T = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t cnt = 0, sb = 0, sp = 0;
  for (uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i0 < n; i0 += U * T) {
    K key[U], pay[U];
    uint64_t h[U];
    Tuple<K> e[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint64_t i = i0 + u * T;
      key[u] = i < n ? keys[i] : null_key<K>();
      pay[u] = i < n ? pays[i] : K(0);
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      h[u] = mult_shift_hash<K>(key[u], lg);
      e[u] = load_tuple<K>(table + h[u]);
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (key[u] == null_key<K>()) continue;  // past the end, or a null key: never joins
      uint64_t hh = h[u];
      Tuple<K> x = e[u];
      for (uint64_t a = 0; a <= mask; ++a) {
        if (x.key == key[u]) {
          cnt += 1;
          sb += static_cast<uint64_t>(static_cast<int64_t>(x.value));
          sp += static_cast<uint64_t>(static_cast<int64_t>(pay[u]));
        } else if (x.key == null_key<K>()) {
          break;
        }
        hh = (hh + 1) & mask;
        x = load_tuple<K>(table + hh);
      }
    }
  }
  __shared__ uint64_t red[3][256 / kWave];
  const uint32_t w = threadIdx.x / kWave;
  const uint64_t wc = wave_broadcast_last(wave_inclusive_scan(cnt));
  const uint64_t wb = wave_broadcast_last(wave_inclusive_scan(sb));
  const uint64_t wp = wave_broadcast_last(wave_inclusive_scan(sp));
  if (lane_id() == 0) {
    red[0][w] = wc;
    red[1][w] = wb;
    red[2][w] = wp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0, d = 0;
    for (uint32_t i = 0; i < blockDim.x / kWave; ++i) {
      a += red[0][i];
      b += red[1][i];
      d += red[2][i];
    }
    out[blockIdx.x].count = a;
    out[blockIdx.x].sum_build = b;
    out[blockIdx.x].sum_probe = d;
  }
}

template <typename K>
hipError_t build_t(JoinScheme s, bool selective, void* table, uint64_t entries, const void* keys, const void* pays, uint64_t n,
                   uint32_t grid, uint32_t block, hipStream_t stream) {
  if (s == JoinScheme::LinearProbing)
    hipLaunchKernelGGL(np_build_lp_kernel<K>, dim3(grid), dim3(block), 0, stream, static_cast<Tuple<K>*>(table),
                       entries, static_cast<const K*>(keys), static_cast<const K*>(pays), n);
  else
    hipLaunchKernelGGL(np_build_perfect_kernel<K>, dim3(grid), dim3(block), 0, stream,
                       static_cast<Tuple<K>*>(table), entries, static_cast<const K*>(keys),
                       static_cast<const K*>(pays), n, selective);
  return hipGetLastError();
}

template <typename K>
hipError_t probe_t(JoinScheme s, const void* table, uint64_t entries, const void* keys, const void* pays, uint64_t n,
                   uint64_t* result_set, JoinAggregate* aggregates, uint32_t grid, uint32_t block,
                   hipStream_t stream) {
  if (s == JoinScheme::LinearProbing)
    hipLaunchKernelGGL((np_probe_kernel<K, false>), dim3(grid), dim3(block), 0, stream,
                       static_cast<const Tuple<K>*>(table), entries, static_cast<const K*>(keys),
                       static_cast<const K*>(pays), n, result_set, aggregates);
  else
    hipLaunchKernelGGL((np_probe_kernel<K, true>), dim3(grid), dim3(block), 0, stream,
                       static_cast<const Tuple<K>*>(table), entries, static_cast<const K*>(keys),
                       static_cast<const K*>(pays), n, result_set, aggregates);
  return hipGetLastError();
}

}  // namespace

// null_values: every word of the entry becomes the null key, as the
// reference's cuMemsetD32 initialisation does (no_partitioning_join.rs:
// 611-624) -- the i32 build's 64-bit CAS expects that; otherwise the payload
// is zeroed (the reference tests' Tuple { null_key, default } padding).
hipError_t launch_np_table_init(KeyType t, void* table, uint64_t entries, bool null_values, hipStream_t stream) {
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((entries + 255) / 256, 8192));
  if (t == KeyType::I32)
    hipLaunchKernelGGL(np_init_kernel<int32_t>, dim3(grid ? grid : 1), dim3(256), 0, stream,
                       static_cast<Tuple<int32_t>*>(table), entries, null_values ? null_key<int32_t>() : 0);
  else
    hipLaunchKernelGGL(np_init_kernel<int64_t>, dim3(grid ? grid : 1), dim3(256), 0, stream,
                       static_cast<Tuple<int64_t>*>(table), entries, null_values ? null_key<int64_t>() : 0);
  return hipGetLastError();
}

// `selective` skips null build keys in the perfect build, mirroring the
// reference's is_selective flag (no_partitioning_join.cu:266-281).
hipError_t launch_np_build(KeyType t, JoinScheme s, bool selective, void* table, uint64_t entries, const void* keys,
                           const void* pays, uint64_t n, uint32_t grid, uint32_t block, hipStream_t stream) {
  if (t == KeyType::I32) return build_t<int32_t>(s, selective, table, entries, keys, pays, n, grid, block, stream);
  return build_t<int64_t>(s, selective, table, entries, keys, pays, n, grid, block, stream);
}

hipError_t launch_np_probe_resident(KeyType t, const void* table, uint32_t lg, const void* keys, const void* pays,
                                    uint64_t n, JoinAggregate* out, uint32_t grid, hipStream_t stream) {
  constexpr uint32_t U = 4;
  if (t == KeyType::I32)
    hipLaunchKernelGGL((np_probe_resident_kernel<int32_t, U>), dim3(grid), dim3(256), 0, stream,
                       static_cast<const Tuple<int32_t>*>(table), lg, static_cast<const int32_t*>(keys),
                       static_cast<const int32_t*>(pays), n, out);
  else
    hipLaunchKernelGGL((np_probe_resident_kernel<int64_t, U>), dim3(grid), dim3(256), 0, stream,
                       static_cast<const Tuple<int64_t>*>(table), lg, static_cast<const int64_t*>(keys),
                       static_cast<const int64_t*>(pays), n, out);
  return hipGetLastError();
}

hipError_t launch_np_probe(KeyType t, JoinScheme s, const void* table, uint64_t entries, const void* keys,
                           const void* pays, uint64_t n, uint64_t* result_set, JoinAggregate* aggregates,
                           uint32_t grid, uint32_t block, hipStream_t stream) {
  if (t == KeyType::I32)
    return probe_t<int32_t>(s, table, entries, keys, pays, n, result_set, aggregates, grid, block, stream);
  return probe_t<int64_t>(s, table, entries, keys, pays, n, result_set, aggregates, grid, block, stream);
}

}  // namespace fij
