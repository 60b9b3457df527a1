// Batched radix-join pipeline: the execution method of the radix-join binary
// (radix-join/src/execution_methods/gpu_radix_join.rs:90-565) re-planned for
// one MI355X.  The reference runs pass 1 over both relations and then, for
// every pass-1 partition, six launches (two transforms, two pass-2 scatters,
// task assignment, join) on two streams (:422-513).  Here every phase is one
// launch over all partitions:
//
//   1. hist1    chunked pass-1 histograms of R and S      (hist.hip)
//   2. scatter1 LDS line-buffered pass-1 scatter           (partition.hip)
//   3. plan     partition lengths, chunk-piece tables, pass-2 tile map
//   4. sort2    pass 2 as an LDS counting sort of TILE2-tuple tiles of each
//               pass-1 partition: one coalesced read and one coalesced write
//               per tuple; each tile records where every pass-2 digit's run
//               starts inside it
//   5. join     one wave per (pass-1, pass-2) partition pair: gathers the
//               build runs into a wave-private LDS hash table (linear
//               probing) and probes the probe runs, aggregating match count,
//               SUM(build.value) and SUM(probe.value)
//
// With bits2 == 0 (single pass) the join reads the pass-1 chunk pieces
// directly.  Work items of the join are taken from an atomic counter so that
// skewed partitions do not serialise a fixed wave.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
namespace {

// --------------------------------------------------------------- plan ----

// One block per pass-1 partition p: piece c of p is chunk c's range of p.
// piece_start[p][c] = logical start of piece c within p; [p][G1] = |p|.
__global__ __launch_bounds__(256) void plan_pieces_kernel(const uint64_t* off1, uint32_t F1, uint32_t G1,
                                                          uint32_t pad, uint64_t padded_len, uint32_t* piece_start,
                                                          uint32_t* plen) {
  extern __shared__ uint32_t lens[];
  __shared__ uint32_t wave_tmp[32];
  const uint32_t p = blockIdx.x;
  const uint64_t olen = static_cast<uint64_t>(F1) * G1;
  for (uint32_t c = threadIdx.x; c < G1; c += blockDim.x) {
    const uint64_t ofi = static_cast<uint64_t>(c) * F1 + p;
    const uint64_t b = off1[ofi];
    const uint64_t e = (ofi + 1 < olen) ? off1[ofi + 1] - pad : padded_len;
    lens[c] = static_cast<uint32_t>(e - b);
  }
  __syncthreads();
  const uint32_t total = block_exclusive_scan<uint32_t>(lens, G1, wave_tmp);
  uint32_t* out = piece_start + static_cast<uint64_t>(p) * (G1 + 1);
  for (uint32_t c = threadIdx.x; c < G1; c += blockDim.x) out[c] = lens[c];
  if (threadIdx.x == 0) {
    out[G1] = total;
    plen[p] = total;
  }
}

// One block: partition starts in the concatenated pass-2 output and the tile
// map (tile_start[p] = first tile of p); also resets the join's work counter
// and aggregates.  blockIdx.x selects the relation.
__global__ __launch_bounds__(1024) void plan_tiles_kernel(PipelineDevice d) {
  extern __shared__ uint64_t buf[];
  __shared__ uint64_t wave_tmp[32];
  const uint32_t rel = blockIdx.x;
  const PipelineSide& s = rel == 0 ? d.build : d.probe;
  const uint32_t F1 = d.F1;
  uint64_t* starts = buf;              // [F1]
  uint64_t* tiles = buf + F1;          // [F1]
  for (uint32_t p = threadIdx.x; p < F1; p += blockDim.x) {
    const uint32_t n = s.plen[p];
    starts[p] = n;
    tiles[p] = d.bits2 ? (n + d.tile2 - 1) / d.tile2 : 0;
  }
  __syncthreads();
  const uint64_t total = block_exclusive_scan<uint64_t>(starts, F1, wave_tmp);
  const uint64_t total_tiles = block_exclusive_scan<uint64_t>(tiles, F1, wave_tmp);
  for (uint32_t p = threadIdx.x; p < F1; p += blockDim.x) {
    s.pstart[p] = starts[p];
    s.tile_start[p] = static_cast<uint32_t>(tiles[p]);
  }
  if (threadIdx.x == 0) {
    s.pstart[F1] = total;
    s.tile_start[F1] = static_cast<uint32_t>(total_tiles);
  }
  if (rel == 0) {
    if (threadIdx.x == 0) *d.work_counter = 0;
    for (uint32_t i = threadIdx.x; i < d.agg_len; i += blockDim.x) {
      d.aggregates[i].count = 0;
      d.aggregates[i].sum_build = 0;
      d.aggregates[i].sum_probe = 0;
    }
  }
}

// ------------------------------------------------------------- sort2 ----

template <typename K, uint32_t TILE2, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void sort2_kernel(PipelineDevice d, uint32_t rel) {
  constexpr uint32_t ITEMS = TILE2 / BLOCK;
  using T = Tuple<K>;
  extern __shared__ uint64_t lds_u64[];
  T* sorted = reinterpret_cast<T*>(lds_u64);                        // [TILE2]
  uint32_t* cnt = reinterpret_cast<uint32_t*>(sorted + TILE2);      // [F2]
  uint32_t* pstart_l = cnt + d.F2;                                   // [G1 + 1] local piece starts
  uint64_t* psrc_l = reinterpret_cast<uint64_t*>(pstart_l + d.G1 + 2);  // [G1 + 1] (8-aligned: G1 even)
  __shared__ uint32_t s_p, s_c0, s_np;
  __shared__ uint32_t wave_tmp[32];
  const PipelineSide& s = rel == 0 ? d.build : d.probe;
  const uint32_t F1 = d.F1, F2 = d.F2, G1 = d.G1, tid = threadIdx.x;
  const uint32_t b = blockIdx.x;
  if (b >= s.tile_start[F1]) return;  // uniform: grid is an upper bound
  if (tid == 0) {
    uint32_t lo = 0, hi = F1;  // last p with tile_start[p] <= b
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) / 2;
      if (s.tile_start[m] <= b) lo = m; else hi = m;
    }
    s_p = lo;
  }
  for (uint32_t i = tid; i < F2; i += BLOCK) cnt[i] = 0;
  __syncthreads();
  const uint32_t p = s_p;
  const uint32_t j = b - s.tile_start[p];
  const uint32_t n_p = s.plen[p];
  const uint32_t lo = j * TILE2;
  const uint32_t n = min(TILE2, n_p - lo);
  const uint32_t* ps = s.piece_start + static_cast<uint64_t>(p) * (G1 + 1);
  if (tid == 0) {
    // pieces overlapping [lo, lo + n): first c with ps[c+1] > lo
    uint32_t l = 0, h = G1;
    while (h - l > 1) {
      const uint32_t m = (l + h) / 2;
      if (ps[m] <= lo) l = m; else h = m;
    }
    uint32_t c1 = l;
    while (c1 < G1 && ps[c1] < lo + n) ++c1;
    s_c0 = l;
    s_np = c1 - l;
  }
  __syncthreads();
  const uint32_t c0 = s_c0, np = s_np;
  for (uint32_t i = tid; i <= np; i += BLOCK) {
    pstart_l[i] = ps[c0 + i];
    if (i < np) psrc_l[i] = s.off1[static_cast<uint64_t>(c0 + i) * F1 + p];
  }
  __syncthreads();
  const T* src = static_cast<const T*>(s.rel1);
  const uint64_t mask2 = radix_mask(d.bits2, d.bits1);
  K key[ITEMS], val[ITEMS];
  uint32_t dig[ITEMS], rank[ITEMS];
#pragma unroll
  for (uint32_t k = 0; k < ITEMS; ++k) {
    const uint32_t x = lo + k * BLOCK + tid;
    const bool act = x < lo + n;
    key[k] = 0;
    val[k] = 0;
    if (act) {
      uint32_t l = 0, h = np;  // piece i with pstart_l[i] <= x < pstart_l[i+1]
      while (h - l > 1) {
        const uint32_t m = (l + h) / 2;
        if (pstart_l[m] <= x) l = m; else h = m;
      }
      const T t = load_tuple_nt<K>(src + psrc_l[l] + (x - pstart_l[l]));
      key[k] = t.key;
      val[k] = t.value;
    }
    dig[k] = key_to_partition<K>(key[k], mask2, d.bits1);
    rank[k] = wave_rank_add(cnt, dig[k], act);
  }
  __syncthreads();
  block_exclusive_scan<uint32_t>(cnt, F2, wave_tmp);
  // per-tile run starts of every pass-2 digit (meta[tile][F2] = n)
  uint16_t* meta = s.meta + static_cast<uint64_t>(s.tile_start[p] + j) * (F2 + 1);
  for (uint32_t i = tid; i < F2; i += BLOCK) meta[i] = static_cast<uint16_t>(cnt[i]);
  if (tid == 0) meta[F2] = static_cast<uint16_t>(n);
#pragma unroll
  for (uint32_t k = 0; k < ITEMS; ++k) {
    const uint32_t x = lo + k * BLOCK + tid;
    if (x < lo + n) {
      T t;
      t.key = key[k];
      t.value = val[k];
      sorted[cnt[dig[k]] + rank[k]] = t;
    }
  }
  __syncthreads();
  T* out = static_cast<T*>(s.rel2) + s.pstart[p] + lo;
  for (uint32_t i = tid; i < n; i += BLOCK) {
    const T t = sorted[i];
    store_tuple<K>(out + i, t.key, t.value);
  }
}

// -------------------------------------------------------------- join ----

constexpr uint32_t kMaxRuns = 64;

// Run descriptor source for one side of a work item.
struct RunList {
  const void* base;   // Tuple array
  uint32_t count;     // number of runs
};

// Loads run `r` (start index into the side's tuple array, length).
__device__ __forceinline__ void run_of(const PipelineDevice& d, const PipelineSide& s, uint32_t p, uint32_t q,
                                       uint32_t r, uint64_t& start, uint32_t& len) {
  if (d.bits2) {
    const uint32_t tile = s.tile_start[p] + r;
    const uint16_t* m = s.meta + static_cast<uint64_t>(tile) * (d.F2 + 1);
    const uint32_t b = m[q], e = m[q + 1];
    start = s.pstart[p] + static_cast<uint64_t>(r) * d.tile2 + b;
    len = e - b;
  } else {
    // single pass: run r is chunk r's piece of partition p
    const uint32_t* ps = s.piece_start + static_cast<uint64_t>(p) * (d.G1 + 1);
    start = s.off1[static_cast<uint64_t>(r) * d.F1 + p];
    len = ps[r + 1] - ps[r];
  }
}

__device__ __forceinline__ uint32_t run_count(const PipelineDevice& d, const PipelineSide& s, uint32_t p) {
  return d.bits2 ? s.tile_start[p + 1] - s.tile_start[p] : d.G1;
}

template <typename K>
struct WaveTable {
  K* keys;
  K* vals;
  uint32_t cap;  // power of two
};

// Per-wave scratch: run starts / exclusive prefix of the current run batch.
struct WaveRuns {
  uint64_t* start;   // [kMaxRuns]
  uint32_t* prefix;  // [kMaxRuns + 1]
};

// Fills the wave's run batch [r0, r0 + 64) of one side and returns the
// number of tuples in it.
__device__ uint32_t load_run_batch(const PipelineDevice& d, const PipelineSide& s, uint32_t p, uint32_t q,
                                   uint32_t r0, uint32_t nruns, WaveRuns& wr) {
  const uint32_t lane = lane_id();
  const uint32_t r = r0 + lane;
  uint64_t start = 0;
  uint32_t len = 0;
  if (r < nruns) run_of(d, s, p, q, r, start, len);
  const uint32_t incl = wave_inclusive_scan(len);
  wr.start[lane] = start;
  wr.prefix[lane] = incl - len;
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  if (lane == 0) wr.prefix[kMaxRuns] = total;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return total;
}

// Index of element e of the current batch: run with prefix[i] <= e < prefix[i+1].
__device__ __forceinline__ uint64_t batch_element(const WaveRuns& wr, uint32_t e) {
  uint32_t l = 0, h = kMaxRuns;
  while (h - l > 1) {
    const uint32_t m = (l + h) / 2;
    if (wr.prefix[m] <= e) l = m; else h = m;
  }
  return wr.start[l] + (e - wr.prefix[l]);
}

template <typename K, uint32_t CAP, uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void join_kernel(PipelineDevice d) {
  using T = Tuple<K>;
  constexpr uint32_t RCAP = CAP * 3 / 4;       // build tuples per table fill (load <= 3/4)
  constexpr uint32_t BPL = 12;                 // build loads per lane per round
  constexpr uint32_t SPL = 8;                  // probe loads per lane per round
  extern __shared__ uint64_t lds_u64[];
  const uint32_t w = threadIdx.x / 64;
  const uint32_t lane = lane_id();
  // per-wave LDS: keys[CAP], vals[CAP], run starts, run prefix
  const size_t wave_bytes = CAP * 2 * sizeof(K) + kMaxRuns * 8 + (kMaxRuns + 1) * 4 + 4;
  uint8_t* mine = reinterpret_cast<uint8_t*>(lds_u64) + w * ((wave_bytes + 15) & ~size_t(15));
  WaveTable<K> tab{reinterpret_cast<K*>(mine), reinterpret_cast<K*>(mine) + CAP, CAP};
  WaveRuns wr{reinterpret_cast<uint64_t*>(mine + CAP * 2 * sizeof(K)),
              reinterpret_cast<uint32_t*>(mine + CAP * 2 * sizeof(K) + kMaxRuns * 8)};
  const uint32_t lg = 31u - __builtin_clz(CAP);
  for (uint32_t i = lane; i < CAP; i += 64) tab.keys[i] = null_key<K>();
  const T* rel_b = static_cast<const T*>(d.bits2 ? d.build.rel2 : d.build.rel1);
  const T* rel_p = static_cast<const T*>(d.bits2 ? d.probe.rel2 : d.probe.rel1);
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const uint32_t items = d.F1 * F2;
  uint64_t cnt = 0, sb = 0, sp = 0;
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(d.work_counter, 1u);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= items) break;
    const uint32_t p = item / F2, q = item % F2;
    const uint32_t nrb = run_count(d, d.build, p);
    const uint32_t nrp = run_count(d, d.probe, p);
    for (uint32_t rb0 = 0; rb0 < nrb; rb0 += kMaxRuns) {
      const uint32_t tb = load_run_batch(d, d.build, p, q, rb0, nrb, wr);
      for (uint32_t e0 = 0; e0 < tb; e0 += RCAP) {
        // build: up to RCAP tuples per table fill, loaded BPL per lane at a
        // time (independent loads), then inserted
        const uint32_t e_end = min(tb, e0 + RCAP);
        for (uint32_t e1 = e0; e1 < e_end; e1 += BPL * 64) {
        K bk[BPL], bv[BPL];
#pragma unroll
        for (uint32_t k = 0; k < BPL; ++k) {
          const uint32_t e = e1 + k * 64 + lane;
          if (e < e_end) {
            const T t = load_tuple<K>(rel_b + batch_element(wr, e));
            bk[k] = t.key;
            bv[k] = t.value;
          } else {
            bk[k] = null_key<K>();
          }
        }
#pragma unroll
        for (uint32_t k = 0; k < BPL; ++k) {
          if (bk[k] != null_key<K>()) {
            uint32_t h = mult_shift_hash<K>(bk[k], lg);
            for (;;) {
              K prev;
              if constexpr (sizeof(K) == 8)
                prev = static_cast<K>(atomicCAS(reinterpret_cast<unsigned long long*>(&tab.keys[h]),
                                                static_cast<unsigned long long>(null_key<K>()),
                                                static_cast<unsigned long long>(bk[k])));
              else
                prev = static_cast<K>(atomicCAS(reinterpret_cast<unsigned int*>(&tab.keys[h]),
                                                static_cast<unsigned int>(null_key<K>()),
                                                static_cast<unsigned int>(bk[k])));
              if (prev == null_key<K>()) {
                tab.vals[h] = bv[k];
                break;
              }
              h = (h + 1) & (CAP - 1);
            }
          }
        }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // probe every probe run of the item against this table fill
        for (uint32_t rp0 = 0; rp0 < nrp; rp0 += kMaxRuns) {
          // the run scratch is shared with the build side: save nothing, the
          // build batch is reloaded below if another fill follows
          const uint32_t tp = load_run_batch(d, d.probe, p, q, rp0, nrp, wr);
          for (uint32_t f0 = 0; f0 < tp; f0 += SPL * 64) {
            K pk[SPL], pv[SPL];
#pragma unroll
            for (uint32_t k = 0; k < SPL; ++k) {
              const uint32_t e = f0 + k * 64 + lane;
              if (e < tp) {
                const T t = load_tuple<K>(rel_p + batch_element(wr, e));
                pk[k] = t.key;
                pv[k] = t.value;
              } else {
                pk[k] = null_key<K>();
                pv[k] = 0;
              }
            }
#pragma unroll
            for (uint32_t k = 0; k < SPL; ++k) {
              if (pk[k] != null_key<K>()) {
                uint32_t h = mult_shift_hash<K>(pk[k], lg);
                K kk;
                while ((kk = tab.keys[h]) != null_key<K>()) {
                  if (kk == pk[k]) {
                    cnt += 1;
                    sb += static_cast<uint64_t>(static_cast<int64_t>(tab.vals[h]));
                    sp += static_cast<uint64_t>(static_cast<int64_t>(pv[k]));
                  }
                  h = (h + 1) & (CAP - 1);
                }
              }
            }
          }
        }
        // clear the table and restore the build run batch for the next fill
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < CAP; i += 64) tab.keys[i] = null_key<K>();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (e0 + RCAP < tb || nrp > kMaxRuns) load_run_batch(d, d.build, p, q, rb0, nrb, wr);
      }
    }
  }
  // block reduction into aggregates[blockIdx.x]
  __shared__ uint64_t red[3][16];
  const uint64_t wc = wave_broadcast_last(wave_inclusive_scan(cnt));
  const uint64_t wb = wave_broadcast_last(wave_inclusive_scan(sb));
  const uint64_t wp = wave_broadcast_last(wave_inclusive_scan(sp));
  if (lane == 0) {
    red[0][w] = wc;
    red[1][w] = wb;
    red[2][w] = wp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, bb = 0, c = 0;
    for (uint32_t i = 0; i < WAVES; ++i) {
      a += red[0][i];
      bb += red[1][i];
      c += red[2][i];
    }
    d.aggregates[blockIdx.x].count += a;
    d.aggregates[blockIdx.x].sum_build += bb;
    d.aggregates[blockIdx.x].sum_probe += c;
  }
}

template <typename K, uint32_t CAP, uint32_t WAVES>
size_t join_lds_bytes() {
  const size_t wave_bytes = CAP * 2 * sizeof(K) + kMaxRuns * 8 + (kMaxRuns + 1) * 4 + 4;
  return WAVES * ((wave_bytes + 15) & ~size_t(15));
}

template <typename K, uint32_t TILE2, uint32_t BLOCK>
size_t sort2_lds_bytes(uint32_t F2, uint32_t G1) {
  return TILE2 * sizeof(Tuple<K>) + F2 * 4 + (G1 + 2) * 4 + (G1 + 2) * 8 + 16;
}

}  // namespace

// ------------------------------------------------------------- launches ----

uint32_t pipeline_tile2(KeyType t) { (void)t; return 8192u; }

hipError_t launch_pipeline_plan(const PipelineDevice& d, hipStream_t stream) {
  for (int rel = 0; rel < 2; ++rel) {
    const PipelineSide& s = rel == 0 ? d.build : d.probe;
    hipLaunchKernelGGL(plan_pieces_kernel, dim3(d.F1), dim3(256), d.G1 * 4 + 16, stream, s.off1, d.F1, d.G1,
                       d.pad, s.padded_len1, s.piece_start, s.plen);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(plan_tiles_kernel, dim3(2), dim3(1024), 2 * d.F1 * 8, stream, d);
  return hipGetLastError();
}

template <typename K>
hipError_t launch_sort2_t(const PipelineDevice& d, uint32_t rel, uint32_t max_tiles, hipStream_t stream) {
  constexpr uint32_t TILE2 = 8192;
  constexpr uint32_t BLOCK = 1024;
  auto kern = sort2_kernel<K, TILE2, BLOCK>;
  const size_t lds = sort2_lds_bytes<K, TILE2, BLOCK>(d.F2, d.G1);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(max_tiles), dim3(BLOCK), lds, stream, d, rel);
  return hipGetLastError();
}

hipError_t launch_pipeline_sort2(const PipelineDevice& d, KeyType t, uint32_t rel, uint32_t max_tiles,
                                 hipStream_t stream) {
  if (max_tiles == 0) return hipSuccess;
  if (t == KeyType::I32) return launch_sort2_t<int32_t>(d, rel, max_tiles, stream);
  return launch_sort2_t<int64_t>(d, rel, max_tiles, stream);
}

size_t pipeline_sort2_lds(KeyType t, uint32_t F2, uint32_t G1) {
  if (t == KeyType::I32) return sort2_lds_bytes<int32_t, 8192, 1024>(F2, G1);
  return sort2_lds_bytes<int64_t, 8192, 1024>(F2, G1);
}

template <typename K, uint32_t CAP, uint32_t WAVES>
hipError_t launch_join_t(const PipelineDevice& d, uint32_t grid, hipStream_t stream) {
  auto kern = join_kernel<K, CAP, WAVES>;
  const size_t lds = join_lds_bytes<K, CAP, WAVES>();
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(WAVES * 64), lds, stream, d);
  return hipGetLastError();
}

uint32_t pipeline_join_blocks_per_cu(KeyType t, bool large) {
  (void)t;
  (void)large;
  return 2;
}

hipError_t launch_pipeline_join(const PipelineDevice& d, KeyType t, bool large_tables, uint32_t grid,
                                 hipStream_t stream) {
  if (t == KeyType::I32) {
    if (large_tables) return launch_join_t<int32_t, 8192, 2>(d, grid, stream);
    return launch_join_t<int32_t, 1024, 4>(d, grid, stream);
  }
  if (large_tables) return launch_join_t<int64_t, 4096, 2>(d, grid, stream);
  return launch_join_t<int64_t, 1024, 4>(d, grid, stream);
}

}  // namespace fij
