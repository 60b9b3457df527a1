// Batched small joins: many independent joins whose build side fits one
// block's LDS hash table, joined without partitioning by one persistent
// launch.  Each block takes whole joins from a ticket counter, builds the
// build side into a block-wide linear-probing table (CAS inserts), streams
// the probe side against it and writes the join's {count, SUM(build.value),
// SUM(probe.value)}; a build side larger than the table is built in several
// fills, each followed by a full probe pass.  Null keys never join.
#include "batch.hpp"
#include "common.hpp"

namespace fij {
namespace {

template <typename K, uint32_t CAP, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void small_join_kernel(const SmallJob* jobs, uint32_t n_jobs, uint32_t* ticket) {
  constexpr uint32_t RCAP = CAP * 3 / 4;  // build tuples per table fill
  constexpr uint32_t lg = 31u - __builtin_clz(CAP);
  constexpr uint32_t U = 4;  // probe loads in flight per thread
  extern __shared__ uint64_t lds_u64[];
  K* keys = reinterpret_cast<K*>(lds_u64);
  K* vals = keys + CAP;
  __shared__ uint32_t s_job;
  __shared__ uint64_t red[3][BLOCK / 64];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < CAP; i += BLOCK) keys[i] = null_key<K>();
  for (;;) {
    if (tid == 0) s_job = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t j = s_job;
    if (j >= n_jobs) break;  // uniform
    const SmallJob job = jobs[j];
    const K* bk = static_cast<const K*>(job.build_keys);
    const K* bp = static_cast<const K*>(job.build_pays);
    const K* pk = static_cast<const K*>(job.probe_keys);
    const K* pp = static_cast<const K*>(job.probe_pays);
    uint64_t cnt = 0, sb = 0, sp = 0;
    for (uint64_t f0 = 0; f0 < job.build_len; f0 += RCAP) {
      const uint64_t f1 = min<uint64_t>(job.build_len, f0 + RCAP);
      for (uint64_t i = f0 + tid; i < f1; i += BLOCK) {
        const K key = bk[i];
        if (key == null_key<K>()) continue;
        uint32_t h = mult_shift_hash<K>(key, lg);
        for (;;) {
          K prev;
          if constexpr (sizeof(K) == 8)
            prev = static_cast<K>(atomicCAS(reinterpret_cast<unsigned long long*>(&keys[h]),
                                            static_cast<unsigned long long>(null_key<K>()),
                                            static_cast<unsigned long long>(key)));
          else
            prev = static_cast<K>(atomicCAS(reinterpret_cast<unsigned int*>(&keys[h]),
                                            static_cast<unsigned int>(null_key<K>()), static_cast<unsigned int>(key)));
          if (prev == null_key<K>()) {
            vals[h] = bp[i];
            break;
          }
          h = (h + 1) & (CAP - 1);
        }
      }
      __syncthreads();
      const uint64_t n = job.probe_len;
      for (uint64_t i0 = 0; i0 < n; i0 += static_cast<uint64_t>(U) * BLOCK) {
        K key[U], pay[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint64_t i = min(i0 + u * BLOCK + tid, n - 1);
          key[u] = pk[i];
          pay[u] = pp[i];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          if (i0 + u * BLOCK + tid >= n || key[u] == null_key<K>()) continue;
          uint32_t h = mult_shift_hash<K>(key[u], lg);
          K kk;
          while ((kk = keys[h]) != null_key<K>()) {
            if (kk == key[u]) {
              cnt += 1;
              sb += static_cast<uint64_t>(static_cast<int64_t>(vals[h]));
              sp += static_cast<uint64_t>(static_cast<int64_t>(pay[u]));
            }
            h = (h + 1) & (CAP - 1);
          }
        }
      }
      __syncthreads();
      for (uint32_t i = tid; i < CAP; i += BLOCK) keys[i] = null_key<K>();
      __syncthreads();
    }
    cnt = wave_broadcast_last(wave_inclusive_scan(cnt));
    sb = wave_broadcast_last(wave_inclusive_scan(sb));
    sp = wave_broadcast_last(wave_inclusive_scan(sp));
    if (lane_id() == 0) {
      red[0][tid / 64] = cnt;
      red[1][tid / 64] = sb;
      red[2][tid / 64] = sp;
    }
    __syncthreads();
    if (tid == 0) {
      uint64_t a = 0, b = 0, c = 0;
      for (uint32_t w = 0; w < BLOCK / 64; ++w) {
        a += red[0][w];
        b += red[1][w];
        c += red[2][w];
      }
      job.result->count = a;
      job.result->sum_build = b;
      job.result->sum_probe = c;
    }
    __syncthreads();  // s_job and red are reused
  }
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0; }

template <typename K, uint32_t CAP>
hipError_t launch_small_t(const SmallJob* jobs, uint32_t n_jobs, uint32_t* ticket, uint32_t grid,
                          hipStream_t stream) {
  constexpr uint32_t BLOCK = 1024;
  auto kern = small_join_kernel<K, CAP, BLOCK>;
  const size_t lds = static_cast<size_t>(CAP) * 2 * sizeof(K);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, ticket);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), lds, stream, jobs, n_jobs, ticket);
  return hipGetLastError();
}

}  // namespace

uint64_t small_join_table_tuples(KeyType t) { return t == KeyType::I32 ? 16384u * 3 / 4 : 8192u * 3 / 4; }

hipError_t launch_small_joins(KeyType t, const SmallJob* jobs, uint32_t n_jobs, uint32_t* ticket, uint32_t grid,
                              hipStream_t stream) {
  if (n_jobs == 0) return hipSuccess;
  if (t == KeyType::I32) return launch_small_t<int32_t, 16384>(jobs, n_jobs, ticket, grid, stream);
  return launch_small_t<int64_t, 8192>(jobs, n_jobs, ticket, grid, stream);
}

}  // namespace fij
