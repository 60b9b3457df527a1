// Radix join: per-partition LDS hash build and probe with aggregation.
//
// Semantics follow sql-ops/cudautils/radix_join.cu: partitions are assigned
// to blocks by the reference's greedy rule (:58-98); each block builds an LDS
// table for each of its partitions and probes it (:103-194 perfect hashing,
// :199-311 bucket chaining).  The reference aggregates SUM(probe.value) into
// per-thread partials; this library writes those partials unchanged (when
// asked) and additionally a per-block {match count, SUM(build.value),
// SUM(probe.value)} as wrapping 64-bit integers.
#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {

constexpr uint32_t kChainBuckets = 2048;  // sql-ops/build.rs:72
constexpr uint32_t kChainTail = 0xFFFFu;

struct JoinParams {
  const void* build_rel;
  const uint64_t* build_offsets;
  uint64_t build_len;
  const void* probe_rel;
  const uint64_t* probe_offsets;
  uint64_t probe_len;
  uint32_t radix_bits;
  uint32_t ignore_bits;
  uint32_t ht_entries;
  int64_t* result_set;
  JoinAggregate* aggregates;
  const uint32_t* task_assignments;
  uint32_t* status;
};

// Largest fanout whose task assignment runs by pointer doubling (its LDS:
// prefix sums and two jump arrays of F + 1 entries).
constexpr uint32_t kAssignDoublingMax = 8192;

// radix_join.cu:58-98.  One block: partition sizes are staged in LDS in
// parallel, then the reference's greedy cut is computed in parallel, with
// the same result.
__global__ __launch_bounds__(1024) void assign_tasks_kernel(const uint64_t* probe_offsets,
                                                            uint32_t probe_len, uint32_t pad,
                                                            uint32_t radix_bits, uint32_t grid,
                                                            uint32_t* ta) {
  extern __shared__ uint32_t sizes[];
  const uint32_t F = 1u << radix_bits;
  for (uint32_t p = threadIdx.x; p < F; p += blockDim.x) {
    const uint32_t upper = (p + 1 < F) ? static_cast<uint32_t>(probe_offsets[p + 1] - pad) : probe_len;
    sizes[p] = static_cast<uint32_t>(upper - probe_offsets[p]);
  }
  __syncthreads();
  // The greedy rule: a task ends at the first partition whose size brings
  // the task to at least avg probe tuples.
  if (F <= kAssignDoublingMax) {
    // With P the exclusive prefix sums of the sizes, the task starting at
    // partition s ends before J(s) = the first j > s with P[j] >= P[s] + avg
    // (F if none), so ta[k] = J^k(0): J by binary searches, then J^m
    // squared and applied to ta[0 .. m) in log2(grid) rounds.
    __shared__ uint32_t wave_tmp[1024 / kWave + 1];
    const uint32_t avg = (probe_len - pad * F + grid - 1u) / grid;
    uint32_t* P = sizes;                                       // [F + 1]
    uint16_t* cur = reinterpret_cast<uint16_t*>(sizes + F + 1);  // [F + 1]
    uint16_t* nxt = cur + F + 1;                               // [F + 1]
    const uint32_t total = block_exclusive_scan<uint32_t>(P, F, wave_tmp);
    if (threadIdx.x == 0) P[F] = total;
    __syncthreads();
    for (uint32_t s0 = threadIdx.x; s0 <= F; s0 += blockDim.x) {
      uint32_t lo = s0 + 1, hi = F + 1;
      if (s0 < F) {
        const uint64_t target = static_cast<uint64_t>(P[s0]) + avg;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) / 2;
          if (P[mid] >= target) hi = mid; else lo = mid + 1;
        }
      }
      cur[s0] = static_cast<uint16_t>(min(lo, F));
    }
    if (threadIdx.x == 0) {
      ta[0] = 0;
      ta[grid] = F;
    }
    __syncthreads();
    for (uint64_t m = 1; m < grid; m *= 2) {
      for (uint64_t i = threadIdx.x; i < m && m + i < grid; i += blockDim.x) ta[m + i] = cur[ta[i]];
      for (uint32_t s0 = threadIdx.x; s0 <= F; s0 += blockDim.x) nxt[s0] = cur[cur[s0]];
      __syncthreads();
      uint16_t* t = cur;
      cur = nxt;
      nxt = t;
    }
    return;
  }
  // Larger fanouts: taken 64 partitions at a time by wave 0: an inclusive
  // scan of the sizes gives every lane the running sum since the last cut,
  // a ballot finds the next cut; one step per cut.
  if (threadIdx.x < kWave) {
    const uint32_t lane = threadIdx.x;
    const uint32_t probe_tuples = probe_len - pad * F;
    const uint32_t avg = (probe_tuples + grid - 1u) / grid;
    if (lane == 0) ta[0] = 0;
    uint32_t task_id = 1, task_size = 0;  // wave-uniform
    for (uint32_t base = 0; base < F && task_id < grid; base += kWave) {
      const bool valid = base + lane < F;
      const uint32_t incl = wave_inclusive_scan(valid ? sizes[base + lane] : 0u);
      uint32_t carry = task_size, before = 0, from = 0;  // running sum at lane l: carry + incl[l] - before
      while (task_id < grid) {
        const uint64_t hit = ballot(lane >= from && valid && carry + incl - before >= avg);
        if (hit == 0) break;
        const uint32_t l = static_cast<uint32_t>(__builtin_ctzll(hit));
        if (lane == 0) ta[task_id] = base + l + 1;
        task_id += 1;
        carry = 0;
        before = __builtin_amdgcn_readlane(incl, l);
        from = l + 1;
      }
      task_size = carry + __builtin_amdgcn_readlane(incl, kWave - 1) - before;
    }
    for (uint32_t t = task_id + lane; t <= grid; t += kWave) ta[t] = F;
  }
}

// Probe tuples in flight per thread in the LDS joins.
constexpr uint32_t kJoinLoads = 8;

__device__ __forceinline__ void partition_range(const uint64_t* offsets, uint64_t padded_len, uint32_t F,
                                                uint32_t p, uint32_t pad, uint64_t& begin, uint64_t& end) {
  begin = offsets[p];
  end = (p + 1 < F) ? offsets[p + 1] - pad : padded_len;
}

template <typename K>
__device__ __forceinline__ uint32_t perfect_index(K key, uint32_t ignore_bits) {
  return static_cast<uint32_t>(static_cast<uint64_t>(static_cast<int64_t>(key)) >> ignore_bits);
}

// Bucket of the chaining table: the key's index above the ignored bits (a
// u32, as key_to_partition returns it) widened to the key type, then the
// key type's multiply-shift hash -- 64-bit for i64 keys
// (radix_join.cu:283-284,296-297: hash<K>(key_to_partition(...))).
template <typename K>
__device__ __forceinline__ uint32_t chain_bucket(K key, uint32_t ignore_bits, uint32_t log2_buckets) {
  return mult_shift_hash<K>(static_cast<K>(perfect_index<K>(key, ignore_bits)), log2_buckets);
}

// Block-wide sum of three u64 values into aggregates[blockIdx.x] (+=).
__device__ void block_aggregate(JoinAggregate* out, uint64_t c, uint64_t sb, uint64_t sp) {
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

template <typename K>
__global__ __launch_bounds__(1024) void join_perfect_kernel(JoinParams p) {
  extern __shared__ uint64_t lds_u64[];
  Tuple<K>* ht = reinterpret_cast<Tuple<K>*>(lds_u64);
  const uint32_t F = 1u << p.radix_bits;
  const uint32_t pad = padding_len<K>();
  const uint32_t E = p.ht_entries;
  const Tuple<K>* build = static_cast<const Tuple<K>*>(p.build_rel);
  const Tuple<K>* probe = static_cast<const Tuple<K>*>(p.probe_rel);
  uint64_t cnt = 0, sb = 0, sp = 0;
  int64_t ref_sum = 0;
  bool overflow = false;
  for (uint32_t part = p.task_assignments[blockIdx.x]; part < p.task_assignments[blockIdx.x + 1]; ++part) {
    uint64_t b0, b1, q0, q1;
    partition_range(p.build_offsets, p.build_len, F, part, pad, b0, b1);
    partition_range(p.probe_offsets, p.probe_len, F, part, pad, q0, q1);
    for (uint32_t i = threadIdx.x; i < E; i += blockDim.x) {
      ht[i].key = null_key<K>();
      ht[i].value = 0;
    }
    __syncthreads();
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
      const Tuple<K> t = load_tuple<K>(build + i);
      if (t.key == null_key<K>()) continue;  // null keys never join
      const uint32_t idx = perfect_index<K>(t.key, p.ignore_bits);
      if (idx < E) ht[idx] = t;
      else overflow = true;
    }
    __syncthreads();
    // thread i takes tuples q0 + i, q0 + i + blockDim, ... (kJoinLoads loads
    // in flight; lanes past the end re-read the last tuple and skip it)
    for (uint64_t i0 = q0; i0 < q1; i0 += kJoinLoads * blockDim.x) {
      Tuple<K> t[kJoinLoads];
#pragma unroll
      for (uint32_t k = 0; k < kJoinLoads; ++k)
        t[k] = load_tuple<K>(probe + min<uint64_t>(i0 + k * blockDim.x + threadIdx.x, q1 - 1));
#pragma unroll
      for (uint32_t k = 0; k < kJoinLoads; ++k) {
        const uint32_t idx = perfect_index<K>(t[k].key, p.ignore_bits);
        if (i0 + k * blockDim.x + threadIdx.x < q1 && t[k].key != null_key<K>() && idx < E &&
            ht[idx].key == t[k].key) {
          cnt += 1;
          sb += static_cast<uint64_t>(static_cast<int64_t>(ht[idx].value));
          sp += static_cast<uint64_t>(static_cast<int64_t>(t[k].value));
          ref_sum += static_cast<int64_t>(t[k].value);
        }
      }
    }
    __syncthreads();
  }
  if (overflow && p.status) atomicOr(p.status, kStatusHashTableFull);
  if (p.result_set) p.result_set[static_cast<uint64_t>(blockDim.x) * blockIdx.x + threadIdx.x] += ref_sum;
  if (p.aggregates) block_aggregate(p.aggregates, cnt, sb, sp);
}

template <typename K>
__global__ __launch_bounds__(1024) void join_chaining_kernel(JoinParams p) {
  extern __shared__ uint64_t lds_u64[];
  const uint32_t E = p.ht_entries;  // entries (<= 0xFFFF)
  K* keys = reinterpret_cast<K*>(lds_u64);
  K* vals = keys + E;
  uint32_t* heads = reinterpret_cast<uint32_t*>(vals + E);
  uint16_t* links = reinterpret_cast<uint16_t*>(heads + kChainBuckets);
  const uint32_t log2_buckets = 31u - __builtin_clz(kChainBuckets);
  const uint32_t F = 1u << p.radix_bits;
  const uint32_t pad = padding_len<K>();
  const Tuple<K>* build = static_cast<const Tuple<K>*>(p.build_rel);
  const Tuple<K>* probe = static_cast<const Tuple<K>*>(p.probe_rel);
  uint64_t cnt = 0, sb = 0, sp = 0;
  int64_t ref_sum = 0;
  bool overflow = false;
  for (uint32_t part = p.task_assignments[blockIdx.x]; part < p.task_assignments[blockIdx.x + 1]; ++part) {
    uint64_t b0, b1, q0, q1;
    partition_range(p.build_offsets, p.build_len, F, part, pad, b0, b1);
    partition_range(p.probe_offsets, p.probe_len, F, part, pad, q0, q1);
    const uint64_t nb = b1 - b0;
    if (nb > E) overflow = true;
    const uint32_t nbuild = nb > E ? E : static_cast<uint32_t>(nb);
    for (uint32_t i = threadIdx.x; i < kChainBuckets; i += blockDim.x) heads[i] = kChainTail;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbuild; i += blockDim.x) {
      const Tuple<K> t = load_tuple<K>(build + b0 + i);
      keys[i] = t.key;
      vals[i] = t.value;
      if (t.key == null_key<K>()) continue;  // stored but never linked: null keys never join
      const uint32_t bucket = chain_bucket<K>(t.key, p.ignore_bits, log2_buckets);
      links[i] = static_cast<uint16_t>(atomicExch(&heads[bucket], i));
    }
    __syncthreads();
    for (uint64_t i0 = q0; i0 < q1; i0 += kJoinLoads * blockDim.x) {
      Tuple<K> t[kJoinLoads];
#pragma unroll
      for (uint32_t k = 0; k < kJoinLoads; ++k)
        t[k] = load_tuple<K>(probe + min<uint64_t>(i0 + k * blockDim.x + threadIdx.x, q1 - 1));
#pragma unroll
      for (uint32_t k = 0; k < kJoinLoads; ++k) {
        if (i0 + k * blockDim.x + threadIdx.x >= q1 || t[k].key == null_key<K>()) continue;
        const uint32_t bucket = chain_bucket<K>(t[k].key, p.ignore_bits, log2_buckets);
        for (uint32_t e = heads[bucket]; e != kChainTail; e = links[e]) {
          if (keys[e] == t[k].key) {
            cnt += 1;
            sb += static_cast<uint64_t>(static_cast<int64_t>(vals[e]));
            sp += static_cast<uint64_t>(static_cast<int64_t>(t[k].value));
            ref_sum += static_cast<int64_t>(t[k].value);
          }
        }
      }
    }
    __syncthreads();
  }
  if (overflow && p.status) atomicOr(p.status, kStatusHashTableFull);
  if (p.result_set) p.result_set[static_cast<uint64_t>(blockDim.x) * blockIdx.x + threadIdx.x] += ref_sum;
  if (p.aggregates) block_aggregate(p.aggregates, cnt, sb, sp);
}

template <typename K>
hipError_t launch_join_t(const JoinLaunch& a, hipStream_t stream) {
  JoinParams p{};
  p.build_rel = a.build_rel;
  p.build_offsets = a.build_offsets;
  p.build_len = a.build_len;
  p.probe_rel = a.probe_rel;
  p.probe_offsets = a.probe_offsets;
  p.probe_len = a.probe_len;
  p.radix_bits = a.radix_bits;
  p.ignore_bits = a.ignore_bits;
  p.result_set = a.result_set;
  p.aggregates = a.aggregates;
  p.task_assignments = a.task_assignments;
  p.status = a.status;
  if (a.scheme == JoinScheme::Perfect) {
    auto kern = join_perfect_kernel<K>;
    const uint32_t lds = dynamic_lds_budget(reinterpret_cast<const void*>(kern), a.lds_bytes);
    if (lds < 2 * sizeof(Tuple<K>)) return hipErrorInvalidConfiguration;
    // cuda_radix_join.rs:234-237: entries = LDS / sizeof(entry) - 1.
    p.ht_entries = lds / sizeof(Tuple<K>) - 1;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds));
    hipLaunchKernelGGL(kern, dim3(a.grid), dim3(a.block), lds, stream, p);
  } else {
    auto kern = join_chaining_kernel<K>;
    const uint32_t lds = dynamic_lds_budget(reinterpret_cast<const void*>(kern), a.lds_bytes);
    if (lds < kChainBuckets * 4 + 64) return hipErrorInvalidConfiguration;
    // radix_join.cu:213-221: entries = (LDS - buckets*4) / (2 + k + v), u16 links.
    uint32_t entries = (lds - kChainBuckets * 4) / (2 + 2 * sizeof(K));
    if (entries > kChainTail) entries = kChainTail;
    p.ht_entries = entries;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds));
    hipLaunchKernelGGL(kern, dim3(a.grid), dim3(a.block), lds, stream, p);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_assign_tasks(const uint64_t* probe_offsets, uint64_t probe_len, uint32_t pad,
                               uint32_t radix_bits, uint32_t grid, uint32_t* task_assignments,
                               hipStream_t stream) {
  const uint32_t F = 1u << radix_bits;
  const size_t lds = F <= kAssignDoublingMax ? (F + 1) * 4 + 2 * (F + 1) * 2 : F * 4;
  hipLaunchKernelGGL(assign_tasks_kernel, dim3(1), dim3(1024), lds, stream, probe_offsets,
                     static_cast<uint32_t>(probe_len), pad, radix_bits, grid, task_assignments);
  return hipGetLastError();
}

hipError_t launch_join(const JoinLaunch& a, hipStream_t stream) {
  if (a.type == KeyType::I32) return launch_join_t<int32_t>(a, stream);
  return launch_join_t<int64_t>(a, stream);
}

}  // namespace fij
