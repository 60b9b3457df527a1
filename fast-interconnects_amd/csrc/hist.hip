// Histogram and prefix-sum kernels of the radix partitioner.
//
// Semantics follow the reference's GPU prefix sums
// (sql-ops/cudautils/radix_partition.cu:40-96 chunked, :104-199 contiguous,
// :207-315 copy-with-payload, :323-467 transform); the implementation is
// gfx950-specific: wave64 ballot/mbcnt aggregation into per-wave LDS
// sub-histograms, DPP wave scans, and a decoupled look-back device scan
// with a dynamic tile ticket instead of cooperative grid barriers.
#include <mutex>

#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {

constexpr uint32_t kHistBlockMax = 1024;
constexpr uint32_t kHistLdsBudget = 48 * 1024;
constexpr uint32_t kWaveTmpWords = 32;  // u64 scratch for block scans


// 16 bytes of keys, only as aligned as one key: gfx950 loads 16 bytes from
// any 4-byte aligned global address, so a column that starts anywhere
// (batched joins' columns sit at arbitrary element offsets of one arena)
// still takes the vector path.
template <typename K>
struct VecN {
  static constexpr uint32_t N = 16 / sizeof(K);
  typedef K type __attribute__((ext_vector_type(16 / sizeof(K)), aligned(sizeof(K))));
};

// Shared tail of both histogram kernels: sub-histogram reduction and output.
template <bool CHUNKED>
__device__ void finish_histogram(uint32_t* hist, uint64_t* wave_tmp, uint32_t F, uint32_t S,
                                 const HistParams& p, uint32_t c, uint32_t G) {
  const uint32_t tid = threadIdx.x;
  __syncthreads();
  for (uint32_t i = tid; i < F; i += blockDim.x) {
    uint32_t s = 0;
    for (uint32_t j = 0; j < S; ++j) s += hist[j * F + i];
    hist[i] = s;
  }
  __syncthreads();
  if constexpr (CHUNKED) {
    block_exclusive_scan<uint32_t>(hist, F, reinterpret_cast<uint32_t*>(wave_tmp));
    // radix_partition.cu:85-95: chunk base + exclusive count + (p+1) pads.
    const uint64_t base = (p.chunk_len + static_cast<uint64_t>(p.pad) * F) * c;
    for (uint32_t i = tid; i < F; i += blockDim.x)
      p.offsets[static_cast<uint64_t>(c) * F + i] =
          base + hist[i] + static_cast<uint64_t>(i + 1) * p.pad;
  } else {
    for (uint32_t i = tid; i < F; i += blockDim.x) p.counts[static_cast<uint64_t>(i) * G + c] = hist[i];
  }
}

// Histogram of chunk c of G (one block).
template <typename K, bool CHUNKED, bool COPY>
__device__ __forceinline__ void hist_body(const HistParams& p, uint32_t c, uint32_t G) {
  extern __shared__ uint64_t lds_u64[];
  uint64_t* wave_tmp = lds_u64;
  uint32_t* hist = reinterpret_cast<uint32_t*>(lds_u64 + kWaveTmpWords);
  const uint32_t F = 1u << p.radix_bits;
  const uint32_t S = p.subhists;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x, B = blockDim.x;

  if (p.scan_state) {
    for (uint64_t i = static_cast<uint64_t>(c) * B + tid; i < p.scan_state_len;
         i += static_cast<uint64_t>(G) * B)
      p.scan_state[i] = 0;
  }
  for (uint32_t i = tid; i < S * F; i += B) hist[i] = 0;
  __syncthreads();

  const uint64_t begin = min(p.chunk_len * c, p.n);
  const uint64_t len = (c + 1 == G) ? p.n - begin : min(p.chunk_len, p.n - begin);
  const K* keys = static_cast<const K*>(p.keys) + begin;
  uint32_t* my = hist + ((tid / kWave) % S) * F;

  using V = VecN<K>;
  // the copying variant also stores 16 bytes at a time, and keeps its vector
  // path for columns that are all 16-byte aligned
  const bool aligned = !COPY || ((reinterpret_cast<uintptr_t>(keys) |
                                  reinterpret_cast<uintptr_t>(static_cast<const K*>(p.pays) + begin) |
                                  reinterpret_cast<uintptr_t>(static_cast<K*>(p.dst_keys) + begin) |
                                  reinterpret_cast<uintptr_t>(static_cast<K*>(p.dst_pays) + begin)) % 16) == 0;
  uint64_t scalar_from = 0;
  if (aligned) {
    const uint64_t nv = len / V::N;
    const typename V::type* kv = reinterpret_cast<const typename V::type*>(keys);
    // U vectors per thread and step; the next step's vectors are loaded while
    // this step's keys are counted.  Lanes past the end re-read the last
    // vector (unconditional loads) and count nothing.
    constexpr uint32_t U = 4;
    using VT = typename V::type;
    auto load = [&](VT (&x)[U], uint64_t v0) {
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) x[u] = __builtin_nontemporal_load(as_global(kv) + min(v0 + u * B + tid, nv - 1));
    };
    auto count = [&](const VT (&x)[U], uint64_t v0) {
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint64_t v = v0 + u * B + tid;
        const bool act = v < nv;
        if constexpr (COPY) {
          if (act) {
            const VT* pv = reinterpret_cast<const VT*>(static_cast<const K*>(p.pays) + begin);
            const VT y = __builtin_nontemporal_load(pv + v);
            reinterpret_cast<VT*>(static_cast<K*>(p.dst_keys) + begin)[v] = x[u];
            reinterpret_cast<VT*>(static_cast<K*>(p.dst_pays) + begin)[v] = y;
          }
        }
#pragma unroll
        for (uint32_t j = 0; j < V::N; ++j)
          wave_count_add(my, key_to_partition<K>(x[u][j], mask, p.ignore_bits), act);
      }
    };
    if (nv) {
      VT xa[U], xb[U];
      const uint64_t step = static_cast<uint64_t>(U) * B;
      load(xa, 0);
      for (uint64_t v0 = 0;; v0 += 2 * step) {
        load(xb, v0 + step);
        count(xa, v0);
        if (v0 + step >= nv) break;
        load(xa, v0 + 2 * step);
        count(xb, v0 + step);
        if (v0 + 2 * step >= nv) break;
      }
    }
    scalar_from = nv * V::N;
  }
  for (uint64_t i0 = scalar_from; i0 < len; i0 += B) {
    const uint64_t i = i0 + tid;
    const bool act = i < len;
    K key = 0;
    if (act) {
      key = keys[i];
      if constexpr (COPY) {
        static_cast<K*>(p.dst_keys)[begin + i] = key;
        static_cast<K*>(p.dst_pays)[begin + i] = static_cast<const K*>(p.pays)[begin + i];
      }
    }
    wave_count_add(my, key_to_partition<K>(key, mask, p.ignore_bits), act);
  }
  finish_histogram<CHUNKED>(hist, wave_tmp, F, S, p, c, G);
}

template <typename K, bool CHUNKED, bool COPY>
__global__ __launch_bounds__(kHistBlockMax) void hist_kernel(HistParams p) {
  hist_body<K, CHUNKED, COPY>(p, blockIdx.x, gridDim.x);
}

// Index of the job that owns `block`: the last j with prefix[j] <= block
// (prefix[0] = 0, prefix[jobs] = total blocks).
__device__ __forceinline__ uint32_t job_of_block(const uint32_t* prefix, uint32_t jobs, uint32_t block) {
  uint32_t lo = 0, hi = jobs;
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) / 2;
    if (scalar_load(prefix + m) <= block) lo = m; else hi = m;
  }
  return lo;
}

// Histograms of many relations in one launch: block b counts chunk
// b - prefix[j] of relation j (batched joins).
template <typename K>
__global__ __launch_bounds__(kHistBlockMax) void hist_batched_kernel(const HistParams* jobs, const uint32_t* prefix,
                                                                    uint32_t njobs) {
  const uint32_t j = job_of_block(prefix, njobs, blockIdx.x);
  const HistParams p = scalar_load(jobs + j);
  const uint32_t c0 = scalar_load(prefix + j);
  hist_body<K, false, false>(p, blockIdx.x - c0, scalar_load(prefix + j + 1) - c0);
}

// Offsets of many relations in one launch (one block per relation): the
// exclusive scan of counts[F][G] in partition-major order gives
// local[c][p] = scan + (p + 1) * pad (relative to the relation's region) and
// off1[p] = region + local[0][p].
__global__ __launch_bounds__(1024) void offsets_batched_kernel(const BatchScanJob* jobs) {
  __shared__ uint64_t wave_tmp[kWaveTmpWords];
  const BatchScanJob j = jobs[blockIdx.x];
  const uint32_t tid = threadIdx.x, B = blockDim.x;
  const uint64_t n = static_cast<uint64_t>(j.fanout) * j.chunks;
  const uint64_t per = (n + B - 1) / B;
  const uint64_t lo = min(tid * per, n), hi = min(lo + per, n);
  uint64_t local = 0;
  for (uint64_t e = lo; e < hi; ++e) local += j.counts[e];
  const uint64_t incl = wave_inclusive_scan(local);
  const uint32_t w = tid / kWave;
  if (lane_id() == kWave - 1) wave_tmp[w] = incl;
  __syncthreads();
  uint64_t before = 0;
  for (uint32_t i = 0; i < w; ++i) before += wave_tmp[i];
  uint64_t run = before + incl - local;
  for (uint64_t e = lo; e < hi; ++e) {
    const uint32_t part = static_cast<uint32_t>(e / j.chunks);
    const uint32_t chunk = static_cast<uint32_t>(e - static_cast<uint64_t>(part) * j.chunks);
    const uint64_t off = run + static_cast<uint64_t>(part + 1) * j.pad;
    j.local[static_cast<uint64_t>(chunk) * j.fanout + part] = off;
    if (chunk == 0) j.off1[part] = j.region + off;
    run += j.counts[e];
  }
}

struct TransformParams {
  const void* src_rel;
  const uint64_t* src_offsets;
  uint32_t src_chunks;
  uint32_t src_radix_bits;
  uint64_t src_padded_len;
  uint32_t pid;
  uint64_t n_expected;
  void* dst_keys;
  void* dst_pays;
  uint32_t radix_bits;
  uint32_t ignore_bits;
  uint32_t pad;
  uint32_t subhists;
  uint32_t* counts;
  uint64_t* scan_state;
  uint64_t scan_state_len;
  uint32_t* status;
};

// Gathers one partition of a (possibly chunked) row-layout relation into
// columnar arrays, in chunk order (radix_partition.cu:343-410), and counts
// the next pass's digits.
template <typename K>
__global__ __launch_bounds__(kHistBlockMax) void transform_counts_kernel(TransformParams p) {
  extern __shared__ uint64_t lds_u64[];
  uint64_t* wave_tmp = lds_u64;
  uint64_t* piece_start = lds_u64 + kWaveTmpWords;          // [src_chunks + 1]
  uint64_t* piece_src = piece_start + p.src_chunks + 1;     // [src_chunks]
  uint32_t* hist = reinterpret_cast<uint32_t*>(piece_src + p.src_chunks);
  const uint32_t F = 1u << p.radix_bits;
  const uint32_t S = p.subhists;
  const uint32_t srcF = 1u << p.src_radix_bits;
  const uint32_t src_pad = p.pad;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x, c = blockIdx.x, G = gridDim.x, B = blockDim.x;

  if (p.scan_state) {
    for (uint64_t i = static_cast<uint64_t>(c) * B + tid; i < p.scan_state_len;
         i += static_cast<uint64_t>(G) * B)
      p.scan_state[i] = 0;
  }
  for (uint32_t i = tid; i < S * F; i += B) hist[i] = 0;
  for (uint32_t j = tid; j < p.src_chunks; j += B) {
    const uint64_t ofi = static_cast<uint64_t>(j) * srcF + p.pid;
    const uint64_t b = p.src_offsets[ofi];
    const uint64_t e = (j + 1 == p.src_chunks && p.pid + 1 == srcF) ? p.src_padded_len
                                                                      : p.src_offsets[ofi + 1] - src_pad;
    piece_src[j] = b;
    piece_start[j] = e - b;
  }
  __syncthreads();
  const uint64_t total = block_exclusive_scan<uint64_t>(piece_start, p.src_chunks, wave_tmp);
  if (tid == 0) piece_start[p.src_chunks] = total;
  if (c == 0 && tid == 0 && total != p.n_expected && p.status) atomicOr(p.status, kStatusTransformLength);
  __syncthreads();

  const uint64_t clen = canonical_chunk_len<K>(total, G);
  const uint64_t lo = min(clen * c, total);
  const uint64_t hi = (c + 1 == G) ? total : min(lo + clen, total);
  const Tuple<K>* src = static_cast<const Tuple<K>*>(p.src_rel);
  K* dk = static_cast<K*>(p.dst_keys);
  K* dp = static_cast<K*>(p.dst_pays);
  uint32_t* my = hist + ((tid / kWave) % S) * F;

  // First piece containing this thread's first position (binary search),
  // then a forward cursor as positions advance by B.
  uint32_t j = 0;
  {
    const uint64_t x = lo + tid;
    uint32_t l = 0, r = p.src_chunks;  // invariant: piece_start[l] <= x < piece_start[r]
    while (r - l > 1) {
      const uint32_t m = (l + r) / 2;
      if (piece_start[m] <= x) l = m; else r = m;
    }
    j = l;
  }
  for (uint64_t x0 = lo; x0 < hi; x0 += B) {
    const uint64_t x = x0 + tid;
    const bool act = x < hi;
    uint32_t digit = 0;
    if (act) {
      while (piece_start[j + 1] <= x) ++j;
      const Tuple<K> t = load_tuple<K>(src + piece_src[j] + (x - piece_start[j]));
      dk[x] = t.key;
      dp[x] = t.value;
      digit = key_to_partition<K>(t.key, mask, p.ignore_bits);
    }
    wave_count_add(my, digit, act);
  }
  HistParams hp{};
  hp.counts = p.counts;
  finish_histogram<false>(hist, wave_tmp, F, S, hp, c, G);
}

// Decoupled look-back (Merrill & Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", 2016) exclusive scan over counts[F][G] in
// partition-major order.  Tiles take a dynamic ticket, so a tile only ever
// waits on tiles that started before it.  Each state word packs
// (value << 2) | flag with flag 1 = aggregate, 2 = inclusive prefix.
constexpr uint64_t kFlagAgg = 1, kFlagIncl = 2;

__global__ __launch_bounds__(kScanBlock) void offsets_scan_kernel(const uint32_t* counts, uint32_t F,
                                                                   uint32_t G, uint32_t pad,
                                                                   uint64_t* local_offsets,
                                                                   uint64_t* offsets, uint64_t* state) {
  __shared__ uint64_t wave_tmp[kWaveTmpWords];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t w = tid / kWave;
  if (tid == 0) s_tile = atomicAdd(reinterpret_cast<unsigned int*>(state), 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t n = static_cast<uint64_t>(F) * G;
  const uint64_t base = static_cast<uint64_t>(tile) * kScanTile + static_cast<uint64_t>(tid) * kScanItems;

  uint32_t v[kScanItems];
  uint64_t local = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? counts[base + k] : 0u;
    local += v[k];
  }
  const uint64_t incl = wave_inclusive_scan(local);
  if (lane == kWave - 1) wave_tmp[w] = incl;
  __syncthreads();
  if (tid == 0) {
    uint64_t run = 0;
    for (uint32_t i = 0; i < kScanBlock / kWave; ++i) {
      const uint64_t t = wave_tmp[i];
      wave_tmp[i] = run;
      run += t;
    }
    wave_tmp[kScanBlock / kWave] = run;
  }
  __syncthreads();
  const uint64_t aggregate = wave_tmp[kScanBlock / kWave];
  uint64_t* tile_state = state + 1;

  if (w == 0) {
    uint64_t prefix = 0;
    if (tile == 0) {
      if (lane == 0)
        __hip_atomic_store(&tile_state[0], (aggregate << 2) | kFlagIncl, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(&tile_state[tile], (aggregate << 2) | kFlagAgg, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
      int64_t j = static_cast<int64_t>(tile) - 1;
      for (;;) {
        const int64_t idx = j - static_cast<int64_t>(lane);
        uint64_t s = (idx >= 0) ? __hip_atomic_load(&tile_state[idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                                : kFlagIncl;  // virtual inclusive zero before tile 0
        const uint64_t flag = s & 3u;
        const uint64_t incl_mask = ballot(flag == kFlagIncl);
        const uint64_t invalid_mask = ballot(flag == 0);
        // lanes [0, first inclusive] must all be valid
        const uint64_t window = incl_mask ? ((incl_mask & (~incl_mask + 1)) << 1) - 1 : ~0ull;
        if (invalid_mask & window) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const uint64_t contrib = ((window >> lane) & 1ull) ? (s >> 2) : 0ull;
        prefix += wave_broadcast_last(wave_inclusive_scan(contrib));
        if (incl_mask) break;
        j -= kWave;
      }
      if (lane == 0)
        __hip_atomic_store(&tile_state[tile], ((prefix + aggregate) << 2) | kFlagIncl, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  uint64_t run = s_prefix + wave_tmp[w] + incl - local;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    const uint64_t e = base + k;
    if (e < n) {
      const uint32_t part = static_cast<uint32_t>(e / G);
      const uint32_t chunk = static_cast<uint32_t>(e - static_cast<uint64_t>(part) * G);
      const uint64_t off = run + static_cast<uint64_t>(part + 1) * pad;
      local_offsets[static_cast<uint64_t>(chunk) * F + part] = off;
      if (chunk == 0) offsets[part] = off;
    }
    run += v[k];
  }
}

uint32_t choose_subhists(uint32_t F, uint32_t block) {
  const uint32_t waves = block / kWave;
  uint32_t s = kHistLdsBudget / (F * 4u);
  if (s < 1) s = 1;
  if (s > waves) s = waves;
  return s;
}

template <typename K>
hipError_t hist_launch_t(const void* keys, const void* pays, void* dst_keys, void* dst_pays,
                         uint64_t n, uint32_t chunks, uint32_t radix_bits, uint32_t ignore_bits,
                         uint64_t* offsets, uint32_t* counts, uint64_t* scan_state,
                         uint64_t scan_state_len, uint32_t block, hipStream_t stream) {
  HistParams p{};
  p.keys = keys;
  p.pays = pays;
  p.dst_keys = dst_keys;
  p.dst_pays = dst_pays;
  p.n = n;
  p.chunk_len = canonical_chunk_len<K>(n, chunks);
  p.radix_bits = radix_bits;
  p.ignore_bits = ignore_bits;
  p.pad = padding_len<K>();
  const uint32_t F = 1u << radix_bits;
  p.subhists = choose_subhists(F, block);
  p.offsets = offsets;
  p.counts = counts;
  p.scan_state = scan_state;
  p.scan_state_len = scan_state_len;
  const size_t lds = kWaveTmpWords * 8 + static_cast<size_t>(p.subhists) * F * 4;
  if (offsets) {
    hipLaunchKernelGGL((hist_kernel<K, true, false>), dim3(chunks), dim3(block), lds, stream, p);
  } else if (dst_keys) {
    hipLaunchKernelGGL((hist_kernel<K, false, true>), dim3(chunks), dim3(block), lds, stream, p);
  } else {
    hipLaunchKernelGGL((hist_kernel<K, false, false>), dim3(chunks), dim3(block), lds, stream, p);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_hist_chunked(KeyType t, const void* keys, uint64_t n, uint32_t chunks,
                               uint32_t radix_bits, uint32_t ignore_bits, uint64_t* offsets,
                               uint32_t block, hipStream_t stream) {
  if (t == KeyType::I32)
    return hist_launch_t<int32_t>(keys, nullptr, nullptr, nullptr, n, chunks, radix_bits, ignore_bits,
                                  offsets, nullptr, nullptr, 0, block, stream);
  return hist_launch_t<int64_t>(keys, nullptr, nullptr, nullptr, n, chunks, radix_bits, ignore_bits,
                                offsets, nullptr, nullptr, 0, block, stream);
}

hipError_t launch_hist_counts(KeyType t, const void* keys, const void* pays, void* dst_keys,
                              void* dst_pays, uint64_t n, uint32_t chunks, uint32_t radix_bits,
                              uint32_t ignore_bits, uint32_t* counts, uint64_t* scan_state,
                              uint64_t scan_state_len, uint32_t block, hipStream_t stream) {
  if (t == KeyType::I32)
    return hist_launch_t<int32_t>(keys, pays, dst_keys, dst_pays, n, chunks, radix_bits, ignore_bits,
                                  nullptr, counts, scan_state, scan_state_len, block, stream);
  return hist_launch_t<int64_t>(keys, pays, dst_keys, dst_pays, n, chunks, radix_bits, ignore_bits,
                                nullptr, counts, scan_state, scan_state_len, block, stream);
}

hipError_t launch_transform_counts(KeyType t, const void* src_rel, const uint64_t* src_offsets,
                                   uint32_t src_chunks, uint32_t src_radix_bits,
                                   uint64_t src_padded_len, uint32_t pid, uint64_t n_expected,
                                   void* dst_keys, void* dst_pays, uint32_t chunks,
                                   uint32_t radix_bits, uint32_t ignore_bits, uint32_t* counts,
                                   uint64_t* scan_state, uint64_t scan_state_len,
                                   uint32_t* status, uint32_t block, hipStream_t stream) {
  TransformParams p{};
  p.src_rel = src_rel;
  p.src_offsets = src_offsets;
  p.src_chunks = src_chunks;
  p.src_radix_bits = src_radix_bits;
  p.src_padded_len = src_padded_len;
  p.pid = pid;
  p.n_expected = n_expected;
  p.dst_keys = dst_keys;
  p.dst_pays = dst_pays;
  p.radix_bits = radix_bits;
  p.ignore_bits = ignore_bits;
  p.pad = pad_tuples(t);
  const uint32_t F = 1u << radix_bits;
  p.subhists = choose_subhists(F, block);
  p.counts = counts;
  p.scan_state = scan_state;
  p.scan_state_len = scan_state_len;
  p.status = status;
  const size_t lds = kWaveTmpWords * 8 + (2ull * src_chunks + 1) * 8 + static_cast<size_t>(p.subhists) * F * 4;
  if (t == KeyType::I32)
    hipLaunchKernelGGL(transform_counts_kernel<int32_t>, dim3(chunks), dim3(block), lds, stream, p);
  else
    hipLaunchKernelGGL(transform_counts_kernel<int64_t>, dim3(chunks), dim3(block), lds, stream, p);
  return hipGetLastError();
}

hipError_t launch_hist_batched(KeyType t, const HistParams* jobs, const uint32_t* prefix, uint32_t njobs,
                               uint32_t blocks, uint32_t radix_bits, uint32_t block, hipStream_t stream) {
  if (blocks == 0) return hipSuccess;
  const uint32_t F = 1u << radix_bits;
  const size_t lds = kWaveTmpWords * 8 + static_cast<size_t>(choose_subhists(F, block)) * F * 4;
  if (t == KeyType::I32)
    hipLaunchKernelGGL(hist_batched_kernel<int32_t>, dim3(blocks), dim3(block), lds, stream, jobs, prefix, njobs);
  else
    hipLaunchKernelGGL(hist_batched_kernel<int64_t>, dim3(blocks), dim3(block), lds, stream, jobs, prefix, njobs);
  return hipGetLastError();
}

HistParams hist_params(KeyType t, const void* keys, uint64_t n, uint32_t chunks, uint32_t radix_bits,
                       uint32_t block, uint32_t* counts) {
  HistParams p{};
  p.keys = keys;
  p.n = n;
  p.chunk_len = t == KeyType::I32 ? canonical_chunk_len<int32_t>(n, chunks) : canonical_chunk_len<int64_t>(n, chunks);
  p.radix_bits = radix_bits;
  p.ignore_bits = 0;
  p.pad = pad_tuples(t);
  p.subhists = choose_subhists(1u << radix_bits, block);
  p.counts = counts;
  return p;
}

hipError_t launch_offsets_batched(const BatchScanJob* jobs, uint32_t njobs, hipStream_t stream) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(offsets_batched_kernel, dim3(njobs), dim3(1024), 0, stream, jobs);
  return hipGetLastError();
}

hipError_t launch_offsets_scan(const uint32_t* counts, uint32_t fanout, uint32_t chunks, uint32_t pad,
                               uint64_t* local_offsets, uint64_t* offsets, uint64_t* scan_state,
                               hipStream_t stream) {
  const uint64_t n = static_cast<uint64_t>(fanout) * chunks;
  const uint32_t tiles = static_cast<uint32_t>((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(offsets_scan_kernel, dim3(tiles), dim3(kScanBlock), 0, stream, counts, fanout, chunks,
                     pad, local_offsets, offsets, scan_state);
  return hipGetLastError();
}

// Scratch of the allocation-free contiguous prefix sum: counts[F][chunks]
// (u32, kNoAllocCounters of them) followed by the look-back scan state
// (u64, scan_state_words(kNoAllocCounters)).  One copy per device, held until
// the process ends; every call on a device shares it.
namespace {
constexpr uint64_t kNoAllocBytes = kNoAllocCounters * 4 + scan_state_words(kNoAllocCounters) * 8;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
void* g_scratch_ptr[kNoAllocDevices];
std::mutex g_scratch_mutex;
}  // namespace

hipError_t noalloc_scratch(uint32_t** counts, uint64_t** scan_state) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNoAllocDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    // the array's name is a host-side stand-in: the device address of this
    // device's copy comes from the runtime, with this device current
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = ptr;
  }
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  *counts = reinterpret_cast<uint32_t*>(base);
  *scan_state = reinterpret_cast<uint64_t*>(base + kNoAllocCounters * 4);
  return hipSuccess;
}

}  // namespace fij
