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
// skewed partitions do not serialise a fixed wave; pairs too large for a
// wave's registers are deferred, split into work units and joined from
// block-shared (join_block_kernel) or wave LDS tables (join_general_kernel).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
namespace {

// --------------------------------------------------------------- plan ----

// Partition lengths, starts in the concatenated pass-2 output and the tile
// map (tile_start[p] = first tile of p) of one relation per block (the sides
// in d.sides, build first); block 0 also resets the join's work counters and
// aggregates.  Pass-1 partition p of the contiguous layout is the range
// [off1[p], off1[p + 1] - pad), the last one ends at padded_len1.
__global__ __launch_bounds__(1024) void plan_kernel(PipelineDevice d) {
  extern __shared__ uint64_t buf[];
  __shared__ uint64_t wave_tmp[32];
  const uint32_t rel = d.sides == 2 ? 1u : blockIdx.x;
  const PipelineSide& s = rel == 0 ? d.build : d.probe;
  const uint32_t F1 = d.F1;
  uint64_t* starts = buf;              // [F1]
  uint64_t* tiles = buf + F1;          // [F1]
  for (uint32_t p = threadIdx.x; p < F1; p += blockDim.x) {
    const uint64_t b = s.off1[p];
    const uint64_t e = p + 1 < F1 ? s.off1[p + 1] - d.pad : s.padded_len1;
    const uint32_t n = static_cast<uint32_t>(e - b);
    s.plen[p] = n;
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
  if (blockIdx.x == 0) {
    if (threadIdx.x < kJoinCounterWords / kJoinCounterStride) {
      d.work_counter[threadIdx.x * kJoinCounterStride] = 0;
      d.work_counter[threadIdx.x * kJoinCounterStride + 1] = 0;  // the high half of a 64-bit counter
    }
    for (uint32_t i = threadIdx.x; i < d.agg_len; i += blockDim.x) {
      d.aggregates[i].count = 0;
      d.aggregates[i].sum_build = 0;
      d.aggregates[i].sum_probe = 0;
    }
  }
}

// Batched plans: one block per join side (sides[2 * join + side]) computes
// the lengths, pass-2 starts and tile map of that side's F1j partitions
// (partitions part_base .. part_base + F1j - 1 of the plan, contiguous layout
// up to rel1_end); the tile map is relative to the side's first tile, whose
// count goes to side_tiles[join side] (tile_bases_kernel and
// tile_offset_kernel then make the tiles of all joins one range).  Block 0
// also writes the pass-2 totals and resets the join counters, each join's
// side-0 block clears its result.
__global__ __launch_bounds__(1024) void plan_batched_kernel(PipelineDevice d, const BatchSide* sides, uint32_t F1j,
                                                            uint32_t* side_tiles, uint64_t rel2_total_b,
                                                            uint64_t rel2_total_p) {
  extern __shared__ uint64_t buf[];
  __shared__ uint64_t wave_tmp[32];
  const uint32_t js = blockIdx.x, side = js & 1u;
  const PipelineSide& s = side ? d.probe : d.build;
  const BatchSide bs = sides[js];
  uint64_t* starts = buf;       // [F1j]
  uint64_t* tiles = buf + F1j;  // [F1j]
  for (uint32_t p = threadIdx.x; p < F1j; p += blockDim.x) {
    const uint32_t vp = bs.part_base + p;
    const uint64_t b = s.off1[vp];
    const uint64_t e = p + 1 < F1j ? s.off1[vp + 1] - d.pad : bs.rel1_end;
    const uint32_t n = static_cast<uint32_t>(e - b);
    s.plen[vp] = n;
    starts[p] = n;
    tiles[p] = d.bits2 ? (n + d.tile2 - 1) / d.tile2 : 0;
  }
  __syncthreads();
  block_exclusive_scan<uint64_t>(starts, F1j, wave_tmp);
  const uint64_t total_tiles = block_exclusive_scan<uint64_t>(tiles, F1j, wave_tmp);
  for (uint32_t p = threadIdx.x; p < F1j; p += blockDim.x) {
    s.pstart[bs.part_base + p] = bs.rel2_base + starts[p];
    s.tile_start[bs.part_base + p] = static_cast<uint32_t>(tiles[p]);
  }
  if (threadIdx.x == 0) side_tiles[js] = static_cast<uint32_t>(total_tiles);
  if (js == 0) {
    if (threadIdx.x == 0) {
      d.build.pstart[d.F1] = rel2_total_b;
      d.probe.pstart[d.F1] = rel2_total_p;
    }
    if (threadIdx.x < kJoinCounterWords / kJoinCounterStride) {
      d.work_counter[threadIdx.x * kJoinCounterStride] = 0;
      d.work_counter[threadIdx.x * kJoinCounterStride + 1] = 0;  // the high half of a 64-bit counter
    }
  }
  if (side == 0 && threadIdx.x == 0) {
    JoinAggregate* r = d.job_results + d.job_slot[js >> 1];
    r->count = 0;
    r->sum_build = 0;
    r->sum_probe = 0;
  }
}

// One block: side_tiles[2 * join + side] becomes the exclusive scan, over
// the joins, of that side's tile counts (in place), and tile_start[F1] of
// each side the total.
__global__ __launch_bounds__(1024) void tile_bases_kernel(PipelineDevice d, uint32_t* side_tiles, uint32_t joins) {
  __shared__ uint32_t wave_tmp[16];
  const uint32_t tid = threadIdx.x, w = tid / kWave;
  for (uint32_t side = 0; side < 2; ++side) {
    uint32_t carry = 0;
    for (uint32_t j0 = 0; j0 < joins; j0 += 1024) {
      const uint32_t j = j0 + tid;
      const uint32_t x = j < joins ? side_tiles[2 * j + side] : 0u;
      const uint32_t incl = wave_inclusive_scan(x);
      if (lane_id() == kWave - 1) wave_tmp[w] = incl;
      __syncthreads();
      uint32_t before = 0, total = 0;
#pragma unroll
      for (uint32_t i = 0; i < 16; ++i) {
        before += i < w ? wave_tmp[i] : 0u;
        total += wave_tmp[i];
      }
      if (j < joins) side_tiles[2 * j + side] = carry + before + incl - x;
      carry += total;
      __syncthreads();
    }
    if (tid == 0) (side ? d.probe : d.build).tile_start[d.F1] = carry;
  }
}

// tile_start[p] += the first tile of p's join side (all partitions of both
// sides; F1j partitions per join).
__global__ __launch_bounds__(256) void tile_offset_kernel(PipelineDevice d, const uint32_t* side_tiles, uint32_t F1j) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * d.F1) return;
  const uint32_t side = i >= d.F1 ? 1u : 0u;
  const uint32_t p = i - side * d.F1;
  uint32_t* ts = side ? d.probe.tile_start : d.build.tile_start;
  ts[p] += side_tiles[2 * (p / F1j) + side];
}

// Wave-uniform read of data that no launch in flight writes, through the
// scalar data cache: it is counted apart from the vector loads, so waiting
// for it does not wait for loads already in flight.
template <typename U>
__device__ __forceinline__ U uniform_load(const U* ptr) {
  return *(const __attribute__((address_space(4))) U*)(ptr);
}

// ------------------------------------------------------------- sort2 ----

// Slots of the block-shared LDS table of a deferred item (64 KiB of keys and
// values; twice that in single-pass plans with big_block_tables); such an
// item's build side may fill 3/4 of them.
template <typename K>
constexpr uint32_t kBlockTableSlots = 65536 / (2 * sizeof(K));

// Tuples per pass-2 tile: 128 KiB of LDS (8,192 16-byte tuples; 16,382
// 8-byte tuples: their pairs are loaded from the 16-byte aligned tuple at
// or before the tile's start, and 8,192 pairs must cover the tile from
// either alignment).  A tile holds one run per pass-2 digit, and the join's
// gathers of a run cost by the cache lines it touches: longer runs, fewer
// partial lines.
template <typename K>
constexpr uint32_t kTile2 = sizeof(K) == 4 ? 131072 / sizeof(Tuple<K>) - 2 : 131072 / sizeof(Tuple<K>);

// Raw buffer resource over `bytes` bytes at `base` (uniform): loads past the
// end return zeros.  Third-dword flags of a raw 32-bit buffer on gfx9-family
// parts.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}


// One pass-2 tile: relation, pass-1 partition p, tile j of p, its length and
// its first tuple in rel1.
template <typename K>
struct TileRef {
  uint32_t rel, p, j, n;
  const Tuple<K>* src;
  uint32_t a;  // 8-byte tuples: 1 when src is not 16-byte aligned
};

// Pass 2 of both relations (the sides in d.sides) as one persistent launch:
// block b sorts tiles b, b + grid, ... of the concatenated tile lists (build
// tiles first).  A tile
// is an LDS counting sort of up to TILE2 tuples of one pass-1 partition by
// the pass-2 digit: one coalesced read, one coalesced write, and per digit
// its run bounds in meta.  The next tile's tuples are loaded while the
// current tile is ranked, sorted and written.
template <typename K, uint32_t TILE2, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void sort2_kernel(PipelineDevice d) {
  constexpr uint32_t ITEMS = (TILE2 + BLOCK - 1) / BLOCK;  // tuple slots per thread
  using T = Tuple<K>;
  extern __shared__ uint64_t lds_u64[];
  T* sorted = reinterpret_cast<T*>(lds_u64);                    // [TILE2]
  uint32_t* cnt = reinterpret_cast<uint32_t*>(sorted + TILE2);  // [F2]
  __shared__ uint32_t wave_tmp[32];
  const uint32_t F1 = d.F1, F2 = d.F2, tid = threadIdx.x;
  const uint32_t tiles_b = (d.sides & 1) ? uniform_load(d.build.tile_start + F1) : 0u;
  const uint32_t total = tiles_b + ((d.sides & 2) ? uniform_load(d.probe.tile_start + F1) : 0u);
  const uint64_t mask2 = radix_mask(d.bits2, d.bits1);
  uint32_t t = blockIdx.x;
  if (t >= total) return;  // uniform
  for (uint32_t i = tid; i < F2; i += BLOCK) cnt[i] = 0;
  auto locate = [&](uint32_t tile) {
    TileRef<K> r;
    r.rel = tile >= tiles_b;
    const PipelineSide& s = r.rel ? d.probe : d.build;
    const uint32_t u = r.rel ? tile - tiles_b : tile;
    uint32_t lo = 0, hi = F1;  // last p with tile_start[p] <= u (skips empty partitions)
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) / 2;
      if (uniform_load(s.tile_start + m) <= u) lo = m; else hi = m;
    }
    r.p = lo;
    r.j = u - uniform_load(s.tile_start + lo);
    r.n = min(TILE2, uniform_load(s.plen + lo) - r.j * TILE2);
    r.src = static_cast<const T*>(s.rel1) + uniform_load(s.off1 + lo) + static_cast<uint64_t>(r.j) * TILE2;
    r.a = sizeof(K) == 4 ? static_cast<uint32_t>((reinterpret_cast<uintptr_t>(r.src) / 8) & 1u) : 0u;
    return r;
  };
  // Tile index of a thread's tuple slot k.  8-byte tuples are loaded in
  // pairs from the 16-byte aligned tuple at or before the tile's start (a
  // window that starts a tuples early): slot k holds window position
  // 2 * ((k / 2) * BLOCK + tid) + k % 2.  16-byte tuples: k * BLOCK + tid.
  auto slot_pos = [&](uint32_t k) -> uint32_t {
    if constexpr (sizeof(K) == 4) return 2 * ((k / 2) * BLOCK + tid) + (k & 1u);
    return k * BLOCK + tid;
  };
  // Whether slot k holds a tuple of the tile (window positions a .. a + n - 1
  // for 8-byte tuples)
  auto slot_live = [&](const TileRef<K>& r, uint32_t k) -> bool {
    const uint32_t w = slot_pos(k);
    if constexpr (sizeof(K) == 4) return w >= r.a && w - r.a < r.n;
    return w < r.n;
  };
  // Non-temporal loads.  8-byte tuples (16 per thread) use buffer loads over
  // the tile: one offset register for all of them (the step goes in the
  // scalar offset), which keeps two 16-tuple buffers within 128 VGPRs; lanes
  // past the tile's end read zeros.  16-byte tuples (8 per thread) use
  // global loads; lanes past the end re-read its last tuple.  Either way
  // such lanes count nothing.
  auto load = [&](const TileRef<K>& r, T (&v)[ITEMS]) {
    if constexpr (sizeof(K) == 4) {
      // the pairs that hold the tile's tuples (a pair past them reads zeros).
      // The first pair may start one tuple before the tile (padding in front
      // of a partition, or the previous tile) and the last end one tuple
      // after it (the next tile, the next partition's padding, or the one
      // spare tuple every pass-1 buffer has after its end).
      const __amdgpu_buffer_rsrc_t rs = tile_rsrc(r.src - r.a, ((r.n + r.a + 1) / 2) * 16u);
#pragma unroll
      for (uint32_t k = 0; k < ITEMS; k += 2) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(tid * 16u),
                                                             static_cast<int>((k / 2) * BLOCK * 16u), 2);
        v[k].key = static_cast<K>(x[0]);
        v[k].value = static_cast<K>(x[1]);
        v[k + 1].key = static_cast<K>(x[2]);
        v[k + 1].value = static_cast<K>(x[3]);
      }
    } else {
#pragma unroll
      for (uint32_t k = 0; k < ITEMS; ++k) v[k] = load_tuple_nt<K>(r.src + min(k * BLOCK + tid, r.n - 1));
    }
  };
  auto process = [&](const TileRef<K>& r, const T (&v)[ITEMS]) {
    const PipelineSide& s = r.rel ? d.probe : d.build;
    // digit | rank << 16 per tuple (F2 <= 2^16, rank < TILE2 <= 2^16)
    uint32_t dr[ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) {
      const bool act = slot_live(r, k);
      const uint32_t dig = key_to_partition<K>(v[k].key, mask2, d.bits1);
      // 8-byte tuples: one LDS atomic per tuple is faster; 16-byte tuples
      // (half as many per tile) gain from aggregating lanes of a digit
      uint32_t rank;
      if constexpr (sizeof(K) == 4)
        rank = act ? atomicAdd(&cnt[dig], 1u) : 0u;
      else
        rank = wave_rank_add(cnt, dig, act);
      dr[k] = dig | (rank << 16);
    }
    __syncthreads();
    block_exclusive_scan<uint32_t>(cnt, F2, wave_tmp);
    // run bounds of every pass-2 digit in this tile: start | end << 16 (TILE2 <= 2^16)
    uint32_t* meta = s.meta + static_cast<uint64_t>(uniform_load(s.tile_start + r.p) + r.j) * F2;
    for (uint32_t i = tid; i < F2; i += BLOCK) meta[i] = cnt[i] | ((i + 1 < F2 ? cnt[i + 1] : r.n) << 16);
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k)
      if (slot_live(r, k)) sorted[cnt[dr[k] & 0xffffu] + (dr[k] >> 16)] = v[k];
    __syncthreads();
    T* out = static_cast<T*>(s.rel2) + uniform_load(s.pstart + r.p) + static_cast<uint64_t>(r.j) * TILE2;
    if constexpr (sizeof(K) == 4) {
      // 8-byte tuples: two per lane and 16-byte store, from the first
      // 16-byte aligned tuple of the tile on (a = 0 or 1 tuple before it,
      // and a last odd tuple, stored alone)
      const uint32_t a = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(out) / 8) & 1u);
      const uint32_t pairs = (r.n - min(a, r.n)) / 2;
      for (uint32_t i = tid; i < pairs; i += BLOCK) {
        const T x = sorted[a + 2 * i], y = sorted[a + 2 * i + 1];
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        u32x4 v;
        v.x = static_cast<uint32_t>(x.key);
        v.y = static_cast<uint32_t>(x.value);
        v.z = static_cast<uint32_t>(y.key);
        v.w = static_cast<uint32_t>(y.value);
        __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(out + a + 2 * i));
      }
      if (tid == 0 && a && r.n) store_tuple_nt<K>(out, sorted[0].key, sorted[0].value);
      if (tid == 1 && a + 2 * pairs < r.n) {
        const uint32_t e = a + 2 * pairs;
        store_tuple_nt<K>(out + e, sorted[e].key, sorted[e].value);
      }
    } else {
      for (uint32_t i = tid; i < r.n; i += BLOCK) {
        const T x = sorted[i];
        store_tuple_nt<K>(out + i, x.key, x.value);
      }
    }
    for (uint32_t i = tid; i < F2; i += BLOCK) cnt[i] = 0;
    __syncthreads();
  };
  TileRef<K> cur = locate(t);
  T va[ITEMS], vb[ITEMS];
  load(cur, va);
  __syncthreads();
  // one step: prefetch the next tile into `nbuf`, sort the current one from
  // `cbuf`; false when this block has no tile left
  auto step = [&](T (&cbuf)[ITEMS], T (&nbuf)[ITEMS]) -> bool {
    const uint32_t tn = t + gridDim.x;
    const bool more = tn < total;
    const TileRef<K> nx = more ? locate(tn) : cur;
    load(nx, nbuf);
    process(cur, cbuf);
    t = tn;
    cur = nx;
    return more;
  };
  for (;;) {
    if (!step(va, vb)) break;
    if (!step(vb, va)) break;
  }
}

// -------------------------------------------------------------- join ----

constexpr uint32_t kMaxRuns = 64;

// Runs of one side of item (p, q): with two passes, one per pass-2 tile of
// pass-1 partition p; with one pass, partition p cut into tile2-tuple pieces.
__device__ __forceinline__ uint32_t run_count(const PipelineDevice& d, const PipelineSide& s, uint32_t p) {
  if (d.bits2) return uniform_load(s.tile_start + p + 1) - uniform_load(s.tile_start + p);
  return (uniform_load(s.plen + p) + d.tile2 - 1) / d.tile2;
}

// Loaded, not yet combined, descriptor of run r0 + lane of item (p, q).  With
// two passes run r is pass-2 digit q's range inside tile r of pass-1
// partition p: a = first tuple of the tile, b0 = its meta word (start | end
// << 16).  With one pass run r is piece r of partition p: a = its first
// tuple, [b0, b1) = [0, its length).  Every lane loads from an in-bounds address
// (lanes past the last run repeat it) and combining is left to the consumer:
// no branch and no use follows the loads, so a prefetch does not make the
// gathers issued before it wait.
struct RawRun {
  uint64_t a;
  uint32_t b0, b1;
};

// Descriptor of run r (in range) of item (p, q).
__device__ __forceinline__ RawRun raw_run_at(const PipelineDevice& d, const PipelineSide& s, uint32_t p, uint32_t q,
                                             uint32_t r) {
  RawRun x;
  if (d.bits2) {
    // nruns == 0 reads row tile_start[p] <= tiles: the meta has a spare row
    const uint32_t tile = uniform_load(s.tile_start + p) + r;
    x.a = uniform_load(s.pstart + p) + static_cast<uint64_t>(r) * d.tile2;
    x.b0 = s.meta[static_cast<uint64_t>(tile) * d.F2 + q];
    x.b1 = 0;
  } else {
    const uint32_t plen = uniform_load(s.plen + p);
    x.a = uniform_load(s.off1 + p) + static_cast<uint64_t>(r) * d.tile2;
    x.b0 = 0;
    x.b1 = min(d.tile2, plen - min(plen, r * d.tile2));
  }
  return x;
}

__device__ __forceinline__ RawRun raw_run(const PipelineDevice& d, const PipelineSide& s, uint32_t p, uint32_t q,
                                          uint32_t r0, uint32_t nruns) {
  const uint32_t last = nruns ? nruns - 1 : 0;
  return raw_run_at(d, s, p, q, min(r0 + lane_id(), last));
}

// Lane-held descriptor of run r0 + lane (length 0 past the last run).
struct LaneRun {
  uint64_t start;
  uint32_t len;
};

__device__ __forceinline__ LaneRun finish_run_at(const PipelineDevice& d, const RawRun& x, bool live) {
  if (d.bits2) {
    const uint32_t b = x.b0 & 0xffffu, e = x.b0 >> 16;
    return LaneRun{x.a + b, live ? e - b : 0u};
  }
  return LaneRun{x.a, live ? x.b1 - x.b0 : 0u};
}

__device__ __forceinline__ LaneRun finish_run(const PipelineDevice& d, const RawRun& x, uint32_t r0,
                                              uint32_t nruns) {
  return finish_run_at(d, x, r0 + lane_id() < nruns);
}

__device__ __forceinline__ LaneRun lane_run(const PipelineDevice& d, const PipelineSide& s, uint32_t p, uint32_t q,
                                            uint32_t r0, uint32_t nruns) {
  return finish_run(d, raw_run(d, s, p, q, r0, nruns), r0, nruns);
}

// Per-wave LDS copy of one side's run batch: run starts and the exclusive
// prefix of the run lengths (prefix[kMaxRuns] = batch total).
struct WaveRuns {
  uint64_t* start;   // [kMaxRuns]
  uint32_t* prefix;  // [kMaxRuns + 1]
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Publishes a lane-held run batch to the wave's LDS; returns its tuple count.
__device__ __forceinline__ uint32_t publish_runs(const LaneRun& x, const WaveRuns& wr) {
  const uint32_t lane = lane_id();
  const uint32_t incl = wave_inclusive_scan(x.len);
  wr.start[lane] = x.start;
  wr.prefix[lane] = incl - x.len;
  const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
  if (lane == 0) wr.prefix[kMaxRuns] = total;
  return total;
}

// Tuple index of element e of the published batch: the run i with
// prefix[i] <= e < prefix[i + 1] (empty runs are skipped by taking the last
// such i).
__device__ __forceinline__ uint64_t batch_element(const WaveRuns& wr, uint32_t e) {
  // branch-free and fully unrolled: no loop between the gathers that follow,
  // so their loads stay in flight together (prefix[0] = 0 <= e)
  uint32_t l = 0;
#pragma unroll
  for (uint32_t step = kMaxRuns / 2; step > 0; step >>= 1)
    l = wr.prefix[l + step] <= e ? l + step : l;
  return wr.start[l] + (e - wr.prefix[l]);
}

// Gather indices of elements k * 64 + lane (k < N) of a published batch of
// at most N * 64 tuples, without a search: every non-empty run writes its
// index at its first element's position of the wave's `marks` (all zero
// between calls), and the run of element e is the maximum mark at or before
// e -- a wave max-scan per 64 elements, carried across rounds.  Elements
// past the batch end repeat its last element.
template <uint32_t N>
__device__ __forceinline__ void run_indices(const WaveRuns& wr, uint8_t* marks, const LaneRun& x, uint32_t total,
                                            uint64_t (&idx)[N]) {
  const uint32_t lane = lane_id();
  const uint32_t first = wave_inclusive_scan(x.len) - x.len;
  if (x.len) marks[first] = static_cast<uint8_t>(lane);
  wave_lds_sync();
  uint32_t carry = 0;
#pragma unroll
  for (uint32_t k = 0; k < N; ++k) {
    const uint32_t e = min(k * 64 + lane, total - 1);
    const uint32_t r = max(carry, wave_inclusive_max(marks[e]));
    carry = __builtin_amdgcn_readlane(r, 63);
    idx[k] = wr.start[r] + (e - wr.prefix[r]);
  }
  wave_lds_sync();
  if (x.len) marks[first] = 0;
}

// Wave-private LDS hash table: linear probing from the multiply-shift hash
// over CAP (power of two) slots.  KT is the stored key type; a free slot holds
// all ones.  Values are stored beside the keys.
template <typename KT, typename V, uint32_t CAP>
struct WaveTable {
  KT* keys;
  V* vals;
  static constexpr uint32_t lg = 31u - __builtin_clz(CAP);
  static constexpr KT kEmpty = static_cast<KT>(~static_cast<KT>(0));

  __device__ __forceinline__ void insert(KT key, V val) const {
    uint32_t h = mult_shift_hash<KT>(key, lg);
    for (;;) {
      KT prev;
      if constexpr (sizeof(KT) == 8)
        prev = static_cast<KT>(atomicCAS(reinterpret_cast<unsigned long long*>(&keys[h]),
                                         static_cast<unsigned long long>(kEmpty),
                                         static_cast<unsigned long long>(key)));
      else
        prev = static_cast<KT>(atomicCAS(reinterpret_cast<unsigned int*>(&keys[h]), static_cast<unsigned int>(kEmpty),
                                         static_cast<unsigned int>(key)));
      if (prev == kEmpty) {
        vals[h] = val;
        return;
      }
      h = (h + 1) & (CAP - 1);
    }
  }

  // Every build tuple with this key matches (probing continues to the first
  // free slot, so duplicate build keys all count).
  __device__ __forceinline__ void probe(KT key, V pay, uint64_t& cnt, uint64_t& sb, uint64_t& sp) const {
    uint32_t h = mult_shift_hash<KT>(key, lg);
    KT kk;
    while ((kk = keys[h]) != kEmpty) {
      if (kk == key) {
        cnt += 1;
        sb += static_cast<uint64_t>(static_cast<int64_t>(vals[h]));
        sp += static_cast<uint64_t>(static_cast<int64_t>(pay));
      }
      h = (h + 1) & (CAP - 1);
    }
  }

  // 16-byte stores of all ones.
  __device__ __forceinline__ void clear() const {
    uint4* k4 = reinterpret_cast<uint4*>(keys);
    constexpr uint32_t n4 = CAP * sizeof(KT) / 16;
    for (uint32_t i = lane_id(); i < n4; i += 64) k4[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
  }
};

constexpr uint32_t kMarkBytes = 640;  // run marks per side: the register-resident batch (10 * 64 tuples)

template <typename K, uint32_t CAP>
constexpr size_t join_wave_bytes() {
  return (CAP * 2 * sizeof(K) + 2 * kMaxRuns * 8 + 2 * (kMaxRuns + 1) * 4 + 2 * kMarkBytes + 15) & ~size_t(15);
}

// Any item, any run count and size: run batches of 64 per side, table fills
// of RCAP build tuples, every probe batch streamed against every fill.  The
// table is clear on entry and on return.  bt / pt are register scratch (the
// caller's idle pipeline buffers).
template <typename K, uint32_t CAP, uint32_t RCAP, uint32_t BPL, uint32_t SPL>
__device__ __forceinline__ void join_item_general(const PipelineDevice& d, uint32_t p, uint32_t q, uint32_t pr0,
                                                  uint32_t pr1, K* table_keys, const WaveRuns& rb,
                                                  const WaveRuns& rp, const Tuple<K>* rel_b, const Tuple<K>* rel_p,
                                                  Tuple<K> (&bt)[BPL], Tuple<K> (&pt)[SPL], uint64_t& cnt,
                                                  uint64_t& sb, uint64_t& sp) {
  // every build tuple of the item joins the probe runs [pr0, pr1)
  const WaveTable<K, K, CAP> tab{table_keys, table_keys + CAP};
  const uint32_t lane = lane_id();
  const uint32_t nb = run_count(d, d.build, p);
  for (uint32_t rb0 = 0; rb0 < nb; rb0 += kMaxRuns) {
    const uint32_t tb = publish_runs(lane_run(d, d.build, p, q, rb0, nb), rb);
    wave_lds_sync();
    for (uint32_t e0 = 0; e0 < tb; e0 += RCAP) {
      const uint32_t e_end = min(tb, e0 + RCAP);
      for (uint32_t e1 = e0; e1 < e_end; e1 += BPL * 64) {
#pragma unroll
        for (uint32_t k = 0; k < BPL; ++k) {
          const uint32_t e = e1 + k * 64 + lane;
          bt[k].key = null_key<K>();
          if (e < e_end) bt[k] = load_tuple<K>(rel_b + batch_element(rb, e));
        }
#pragma unroll
        for (uint32_t k = 0; k < BPL; ++k)
          if (bt[k].key != null_key<K>()) tab.insert(bt[k].key, bt[k].value);
      }
      wave_lds_sync();
      for (uint32_t rp0 = pr0; rp0 < pr1; rp0 += kMaxRuns) {
        const uint32_t tp = publish_runs(lane_run(d, d.probe, p, q, rp0, pr1), rp);
        wave_lds_sync();
        for (uint32_t f0 = 0; f0 < tp; f0 += SPL * 64) {
#pragma unroll
          for (uint32_t k = 0; k < SPL; ++k) {
            const uint32_t e = f0 + k * 64 + lane;
            pt[k].key = null_key<K>();
            if (e < tp) pt[k] = load_tuple<K>(rel_p + batch_element(rp, e));
          }
#pragma unroll
          for (uint32_t k = 0; k < SPL; ++k)
            if (pt[k].key != null_key<K>()) tab.probe(pt[k].key, pt[k].value, cnt, sb, sp);
        }
        wave_lds_sync();
      }
      tab.clear();
      wave_lds_sync();
    }
  }
}

// Build and probe of one item whose tuples are in registers (at most BPL * 64
// build and SPL * 64 probe tuples; lanes past tb / tp hold copies and are
// ignored).  All keys of the item share their low `shift` bits (its radix
// digits), so two keys are equal iff their values >> shift are.  When every
// build key's upper part fits 32 bits (below all ones), the table stores
// that narrow key: half the LDS traffic of 64-bit keys.
template <typename K, uint32_t CAP, uint32_t BPL, uint32_t SPL, bool NARROW>
__device__ __forceinline__ void join_in_registers(K* table_keys, const Tuple<K> (&bt)[BPL],
                                                  const Tuple<K> (&pt)[SPL], uint32_t tb, uint32_t tp,
                                                  uint32_t shift, uint64_t& cnt, uint64_t& sb, uint64_t& sp) {
  const uint32_t lane = lane_id();
  bool narrow = NARROW;
  if constexpr (NARROW && sizeof(K) == 8) {
    bool wide = false;
#pragma unroll
    for (uint32_t k = 0; k < BPL; ++k)
      wide |= k * 64 + lane < tb && bt[k].key != null_key<K>() &&
              (static_cast<uint64_t>(bt[k].key) >> shift) >= 0xffffffffull;
    narrow = __builtin_amdgcn_ballot_w64(wide) == 0;
  }
  if (narrow) {
    const WaveTable<uint32_t, K, CAP> tab{reinterpret_cast<uint32_t*>(table_keys), table_keys + CAP};
#pragma unroll
    for (uint32_t k = 0; k < BPL; ++k)
      if (k * 64 + lane < tb && bt[k].key != null_key<K>())
        tab.insert(static_cast<uint32_t>(static_cast<uint64_t>(static_cast<std::make_unsigned_t<K>>(bt[k].key)) >> shift),
                   bt[k].value);
    wave_lds_sync();
#pragma unroll
    for (uint32_t k = 0; k < SPL; ++k) {
      const uint64_t hi = static_cast<uint64_t>(static_cast<std::make_unsigned_t<K>>(pt[k].key)) >> shift;
      // a probe key whose upper part does not fit matches no build key
      if (k * 64 + lane < tp && pt[k].key != null_key<K>() && hi < 0xffffffffull)
        tab.probe(static_cast<uint32_t>(hi), pt[k].value, cnt, sb, sp);
    }
    wave_lds_sync();
    tab.clear();
  } else {
    const WaveTable<K, K, CAP> tab{table_keys, table_keys + CAP};
#pragma unroll
    for (uint32_t k = 0; k < BPL; ++k)
      if (k * 64 + lane < tb && bt[k].key != null_key<K>()) tab.insert(bt[k].key, bt[k].value);
    wave_lds_sync();
#pragma unroll
    for (uint32_t k = 0; k < SPL; ++k)
      if (k * 64 + lane < tp && pt[k].key != null_key<K>()) tab.probe(pt[k].key, pt[k].value, cnt, sb, sp);
    wave_lds_sync();
    tab.clear();
  }
  wave_lds_sync();
}

// Per-block {count, SUM(build.value), SUM(probe.value)} added to
// aggregates[blockIdx.x].
template <uint32_t WAVES>
__device__ __forceinline__ void join_block_aggregate(const PipelineDevice& d, uint64_t cnt, uint64_t sb, uint64_t sp) {
  __shared__ uint64_t red[3][WAVES];
  const uint32_t w = threadIdx.x / 64;
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
    uint64_t a = 0, b = 0, c = 0;
    for (uint32_t i = 0; i < WAVES; ++i) {
      a += red[0][i];
      b += red[1][i];
      c += red[2][i];
    }
    d.aggregates[blockIdx.x].count += a;
    d.aggregates[blockIdx.x].sum_build += b;
    d.aggregates[blockIdx.x].sum_probe += c;
  }
}

// Batched plans: the wave's {count, SUM(build.value), SUM(probe.value)}
// added to the result of join `job`, then cleared.  Called by whole waves.
__device__ __forceinline__ void wave_job_aggregate(const PipelineDevice& d, uint32_t job, uint64_t& cnt, uint64_t& sb,
                                                   uint64_t& sp) {
  const uint64_t wc = wave_broadcast_last(wave_inclusive_scan(cnt));
  const uint64_t wb = wave_broadcast_last(wave_inclusive_scan(sb));
  const uint64_t wp = wave_broadcast_last(wave_inclusive_scan(sp));
  if (lane_id() == 0 && (wc | wb | wp)) {
    JoinAggregate* r = d.job_results + d.job_slot[job];
    atomicAdd(reinterpret_cast<unsigned long long*>(&r->count), static_cast<unsigned long long>(wc));
    atomicAdd(reinterpret_cast<unsigned long long*>(&r->sum_build), static_cast<unsigned long long>(wb));
    atomicAdd(reinterpret_cast<unsigned long long*>(&r->sum_probe), static_cast<unsigned long long>(wp));
  }
  cnt = 0;
  sb = 0;
  sp = 0;
}

// The per-wave LDS layout shared by both join kernels: the table (keys, then
// values) and the two sides' run batches.
template <typename K, uint32_t CAP>
struct WaveLds {
  K* table_keys;
  WaveRuns rb, rp;
  uint8_t *mb, *mp;  // run marks of the two sides
  __device__ __forceinline__ explicit WaveLds(uint64_t* lds) {
    uint8_t* mine = reinterpret_cast<uint8_t*>(lds) + (threadIdx.x / 64) * join_wave_bytes<K, CAP>();
    table_keys = reinterpret_cast<K*>(mine);
    uint8_t* runs = mine + CAP * 2 * sizeof(K);
    rb = WaveRuns{reinterpret_cast<uint64_t*>(runs), reinterpret_cast<uint32_t*>(runs + 2 * kMaxRuns * 8)};
    rp = WaveRuns{reinterpret_cast<uint64_t*>(runs + kMaxRuns * 8),
                  reinterpret_cast<uint32_t*>(runs + 2 * kMaxRuns * 8 + (kMaxRuns + 1) * 4)};
    mb = runs + 2 * kMaxRuns * 8 + 2 * (kMaxRuns + 1) * 4;
    mp = mb + kMarkBytes;
  }
};

// A deferred item's probe runs are split into work units of about
// kUnitTuples probe tuples (at most one unit per run), so that a partition
// that holds a heavy key is joined by many waves; every unit builds the
// item's whole build side.
constexpr uint64_t kUnitTuples = 65536;

// Turns the items the register-resident join deferred (more than 64 runs or
// more tuples than its registers hold on a side: it set overflow[item] = 1
// and the counter word kOverflowAny) into work units {item, first run, end
// run, path}: an item whose build side fits one fill of a block-shared table
// (block_rcap tuples) is joined by join_block_kernel (path 0), a larger one
// by join_general_kernel's wave tables with refills (path 1).  Blocks take
// tickets of 64 items, one item per lane; the waves of a block split the
// items' run rows between them, so the partition of a heavy key (thousands
// of pass-2 tiles, one run row each) is summed by the whole block.  Flags
// found are cleared, so they are all zero again for the next join.
constexpr uint32_t kUnitsWaves = 8;
constexpr uint32_t kUnitsRows = 32;  // rows in flight per lane

// Runs of one side of item (p, q) for the calling lane (p and q may differ
// between lanes), and the sum of the lengths of its runs slice, slice +
// kUnitsWaves, ... (in groups of kUnitsRows rows).
__device__ __forceinline__ void lane_item_side(const PipelineDevice& d, const PipelineSide& s, uint32_t p,
                                               uint32_t q, bool live, uint32_t slice, uint32_t& runs,
                                               uint64_t& tuples) {
  runs = 0;
  tuples = 0;
  if (!d.bits2) {
    const uint32_t plen = live ? s.plen[p] : 0u;
    runs = (plen + d.tile2 - 1) / d.tile2;
    tuples = slice == 0 ? plen : 0u;
    return;
  }
  const uint32_t t0 = s.tile_start[p];
  runs = live ? s.tile_start[p + 1] - t0 : 0u;
  uint32_t most = runs;
#pragma unroll
  for (uint32_t off = 32; off > 0; off >>= 1) most = max(most, static_cast<uint32_t>(__shfl_xor(static_cast<int>(most), off)));
  // coalesced across lanes where they share p: lane l reads word q of each row
  for (uint32_t r = slice * kUnitsRows; r < most; r += kUnitsWaves * kUnitsRows) {
    uint32_t w[kUnitsRows];
#pragma unroll
    for (uint32_t k = 0; k < kUnitsRows; ++k) {
      const uint32_t rr = min(r + k, runs ? runs - 1 : 0u);  // rows of an empty item: its spare row
      w[k] = s.meta[static_cast<uint64_t>(t0 + rr) * d.F2 + q];
    }
#pragma unroll
    for (uint32_t k = 0; k < kUnitsRows; ++k)
      if (r + k < runs) tuples += (w[k] >> 16) - (w[k] & 0xffffu);
  }
}

__global__ __launch_bounds__(kUnitsWaves * 64) void join_units_kernel(PipelineDevice d, uint32_t block_rcap) {
  if (uniform_load(d.work_counter + kOverflowAny) == 0) return;
  __shared__ uint32_t s_ticket;
  __shared__ uint64_t part[2][kUnitsWaves][64];
  const uint32_t lane = lane_id(), w = threadIdx.x / 64;
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const uint32_t items = d.F1 * F2;
  for (;;) {
    if (threadIdx.x == 0) s_ticket = atomicAdd(d.work_counter + kOverflowTicket, 1u);
    __syncthreads();
    const uint32_t t = s_ticket;
    __syncthreads();  // every thread has read s_ticket
    if (static_cast<uint64_t>(t) * 64 >= items) break;
    // one item per lane, the same in every wave
    const uint32_t item = min(t * 64 + lane, items - 1);
    const bool flagged = t * 64 + lane < items && d.overflow[item] != 0;
    if (__builtin_amdgcn_ballot_w64(flagged) == 0) continue;
    const uint32_t p = item / F2, q = item % F2;
    uint32_t np, nb;
    uint64_t tuples, btuples;
    lane_item_side(d, d.probe, p, q, flagged, w, np, tuples);
    lane_item_side(d, d.build, p, q, flagged, w, nb, btuples);
    part[0][w][lane] = tuples;
    part[1][w][lane] = btuples;
    __syncthreads();  // partial sums published; every wave has read its flag
    if (w != 0) continue;
    for (uint32_t i = 1; i < kUnitsWaves; ++i) {
      tuples += part[0][i][lane];
      btuples += part[1][i][lane];
    }
    if (flagged) d.overflow[item] = 0;
    uint32_t path = btuples <= block_rcap ? 0u : 1u;
    // A larger build side gets a hash table in the HBM arena when the arena
    // has room for it (at least twice its tuples, a power of two): path 2.
    uint32_t big_idx = 0, tslot = 0, tlg = 0;
    if (flagged && path == 1 && d.hbm_cap && btuples <= (d.hbm_cap >> 1)) {
      tlg = max(8u, 64u - static_cast<uint32_t>(__builtin_clzll(2 * btuples - 1)));
      const uint64_t size = 1ull << tlg;
      const uint64_t first =
          atomicAdd(reinterpret_cast<unsigned long long*>(d.work_counter + kHbmSlots), static_cast<unsigned long long>(size));
      if (first + size <= d.hbm_cap) {
        big_idx = atomicAdd(d.work_counter + kBigCount, 1u);
        if (big_idx < d.big_cap) {
          path = 2;
          tslot = static_cast<uint32_t>(first);
          d.big[big_idx] = make_uint4(item, tslot, tlg, 0u);
        }
      }
    }
    uint32_t units = 0, bunits = 0;
    if (flagged) units = max(1u, static_cast<uint32_t>(min<uint64_t>((tuples + kUnitTuples - 1) / kUnitTuples, np)));
    if (path == 2) bunits = max(1u, static_cast<uint32_t>(min<uint64_t>((btuples + kUnitTuples - 1) / kUnitTuples, nb)));
    // block-table units go to the front of `units`, wave-table units to the
    // back (unit_cap bounds both together), so each join kernel visits only
    // its own; HBM-table build units to the front of `hbm_units`, their probe
    // units to its back
    const uint32_t ub = path == 0 ? units : 0u, uw = path == 1 ? units : 0u;
    const uint32_t hb = path == 2 ? bunits : 0u, hp = path == 2 ? units : 0u;
    const uint32_t incl_b = wave_inclusive_scan(ub), incl_w = wave_inclusive_scan(uw);
    const uint32_t incl_hb = wave_inclusive_scan(hb), incl_hp = wave_inclusive_scan(hp);
    uint32_t base_b = 0, base_w = 0, base_hb = 0, base_hp = 0;
    if (lane == 63) {
      if (incl_b) base_b = atomicAdd(d.work_counter + kUnitCount, incl_b);
      if (incl_w) base_w = atomicAdd(d.work_counter + kWaveUnitCount, incl_w);
      if (incl_hb) base_hb = atomicAdd(d.work_counter + kHbmBuildUnits, incl_hb);
      if (incl_hp) base_hp = atomicAdd(d.work_counter + kHbmProbeUnits, incl_hp);
    }
    base_b = __builtin_amdgcn_readlane(base_b, 63) + incl_b - ub;
    base_w = __builtin_amdgcn_readlane(base_w, 63) + incl_w - uw;
    base_hb = __builtin_amdgcn_readlane(base_hb, 63) + incl_hb - hb;
    base_hp = __builtin_amdgcn_readlane(base_hp, 63) + incl_hp - hp;
    for (uint32_t u = 0; u < units; ++u) {
      const uint32_t r0 = static_cast<uint32_t>(static_cast<uint64_t>(u) * np / units);
      const uint32_t r1 = static_cast<uint32_t>(static_cast<uint64_t>(u + 1) * np / units);
      if (path == 2)
        d.hbm_units[d.hbm_unit_cap - 1 - (base_hp + u)] = make_uint4(big_idx, r0, r1, 0u);
      else
        d.units[path == 0 ? base_b + u : d.unit_cap - 1 - (base_w + u)] = make_uint4(t * 64 + lane, r0, r1, path);
    }
    for (uint32_t u = 0; u < bunits; ++u) {
      const uint32_t r0 = static_cast<uint32_t>(static_cast<uint64_t>(u) * nb / bunits);
      const uint32_t r1 = static_cast<uint32_t>(static_cast<uint64_t>(u + 1) * nb / bunits);
      d.hbm_units[base_hb + u] = make_uint4(big_idx, r0, r1, 0u);
    }
  }
}

// Joins the wave-table work units (path 1, stored from the back of `units`;
// one wave per unit, tickets of up to 8 units).  When no item was deferred
// every block only adds a zero aggregate.
template <typename K, uint32_t CAP, uint32_t WAVES, uint32_t BPL, uint32_t SPL>
__global__ __launch_bounds__(WAVES * 64) void join_general_kernel(PipelineDevice d) {
  using T = Tuple<K>;
  constexpr uint32_t RCAP = CAP * 3 / 4;
  extern __shared__ uint64_t lds_u64[];
  const WaveLds<K, CAP> l(lds_u64);
  const uint32_t lane = lane_id();
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const T* rel_b = static_cast<const T*>(d.bits2 ? d.build.rel2 : d.build.rel1);
  const T* rel_p = static_cast<const T*>(d.bits2 ? d.probe.rel2 : d.probe.rel1);
  uint64_t cnt = 0, sb = 0, sp = 0;
  const uint32_t n = uniform_load(d.work_counter + kOverflowAny) ? uniform_load(d.work_counter + kWaveUnitCount) : 0;
  if (n) {
    WaveTable<K, K, CAP>{l.table_keys, l.table_keys + CAP}.clear();
    wave_lds_sync();
    const uint32_t per_wave = n / (gridDim.x * WAVES * 16);
    const uint32_t tk = per_wave < 1 ? 1 : (per_wave > 8 ? 8 : per_wave);
    T bt[BPL], pt[SPL];
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(d.work_counter + kUnitTicket, tk);
      t = __builtin_amdgcn_readfirstlane(t);
      if (t >= n) break;
      const uint32_t e = min(t + tk, n);
      for (uint32_t u = t; u < e; ++u) {
        const uint4 unit = d.units[d.unit_cap - 1 - u];
        const uint32_t item = __builtin_amdgcn_readfirstlane(unit.x);
        join_item_general<K, CAP, RCAP, BPL, SPL>(d, item / F2, item % F2, __builtin_amdgcn_readfirstlane(unit.y),
                                                  __builtin_amdgcn_readfirstlane(unit.z), l.table_keys, l.rb, l.rp,
                                                  rel_b, rel_p, bt, pt, cnt, sb, sp);
        if (d.job_results) wave_job_aggregate(d, (item / F2) >> d.job_shift, cnt, sb, sp);
      }
    }
  }
  if (!d.job_results) join_block_aggregate<WAVES>(d, cnt, sb, sp);
}

// Block-wide batch of up to BLOCK runs of one side of an item, one run per
// thread: run starts and the exclusive prefix of their lengths, and a bucket
// table: bucket c holds the run of element c << shift (at most 2 * BLOCK
// buckets), the slot after the last bucket holds the last run.
struct BlockRuns {
  uint64_t* start;   // [BLOCK]
  uint32_t* prefix;  // [BLOCK]
  uint16_t* bucket;  // [2 * BLOCK + 1]
};

template <uint32_t BLOCK>
constexpr size_t block_runs_bytes() {
  return BLOCK * (sizeof(uint64_t) + sizeof(uint32_t)) + (2 * BLOCK + 4) * sizeof(uint16_t);
}

// Publishes runs r0 .. r0 + BLOCK - 1 (those below nruns > r0) of item
// (p, q) and returns their tuple count.  Ends with a barrier.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t publish_block_runs(const PipelineDevice& d, const PipelineSide& s, uint32_t p,
                                                       uint32_t q, uint32_t r0, uint32_t nruns, const BlockRuns& br,
                                                       uint32_t* wave_sums) {
  const uint32_t tid = threadIdx.x, w = tid / 64;
  const uint32_t r = r0 + tid;
  const LaneRun x = finish_run_at(d, raw_run_at(d, s, p, q, min(r, nruns - 1)), r < nruns);
  const uint32_t incl = wave_inclusive_scan(x.len);
  if (lane_id() == 63) wave_sums[w] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (uint32_t i = 0; i < BLOCK / 64; ++i) {
    const uint32_t v = wave_sums[i];
    before += i < w ? v : 0u;
    total += v;
  }
  br.start[tid] = x.start;
  br.prefix[tid] = before + incl - x.len;
  __syncthreads();
  return total;
}

// First step of a binary search over m runs (1 <= m <= BLOCK): the largest
// power of two below m, 0 for m = 1.  The steps sum to at most 2 * step - 1
// < BLOCK, and every slot of a batch is written (slots at or past m carry the
// batch total, above any element), so the search needs no bound check.
__device__ __forceinline__ uint32_t block_search_step(uint32_t m) {
  return m > 1 ? 1u << (31 - __builtin_clz(m - 1)) : 0u;
}

// Fills the bucket table of a published batch of m runs holding n tuples and
// returns its shift (buckets of 2^shift elements, at most 2 * BLOCK of them).
// Each bucket is a binary search over the runs; the two buckets of a thread
// search together.  Ends with a barrier.
template <uint32_t BLOCK>
__device__ __forceinline__ uint32_t publish_block_buckets(const BlockRuns& br, uint32_t m, uint32_t n) {
  constexpr uint32_t kLogBuckets = __builtin_ctz(2 * BLOCK);
  static_assert((2 * BLOCK & (2 * BLOCK - 1)) == 0, "power-of-two block");
  const uint32_t tid = threadIdx.x;
  const uint32_t shift = n > 2 * BLOCK ? 32 - __builtin_clz(n - 1) - kLogBuckets : 0u;
  const uint32_t nbk = (n + (1u << shift) - 1) >> shift;
  const uint32_t step0 = block_search_step(m);
  uint32_t pos[2], l[2] = {0, 0};
#pragma unroll
  for (uint32_t k = 0; k < 2; ++k) pos[k] = (tid + k * BLOCK) << shift;
  for (uint32_t step = step0; step > 0; step >>= 1) {
#pragma unroll
    for (uint32_t k = 0; k < 2; ++k) l[k] = br.prefix[l[k] + step] <= pos[k] ? l[k] + step : l[k];
  }
#pragma unroll
  for (uint32_t k = 0; k < 2; ++k)
    if (tid + k * BLOCK < nbk) br.bucket[tid + k * BLOCK] = static_cast<uint16_t>(l[k]);
  if (tid == 0) br.bucket[nbk] = static_cast<uint16_t>(m - 1);
  __syncthreads();
  return shift;
}

// Tuple indices of elements e[0..N-1] (each below the batch's tuple count) of
// a published block batch: for each, the last run i with prefix[i] <= e
// (empty runs share their successor's prefix and are passed).  The run lies
// between the runs of the element's bucket and of the next bucket, and a
// binary search over that range (usually zero to two runs wide) finds it.
// The N searches take the same steps (sized for the widest range; a
// candidate past a search's range is clamped to its end, which is then
// accepted only if it is the run), and each step issues all N LDS reads
// before it uses any.
template <uint32_t N>
__device__ __forceinline__ void block_elements(const BlockRuns& br, const uint32_t (&e)[N], uint32_t shift,
                                               uint64_t (&idx)[N]) {
  uint32_t l[N], hi[N];
  uint32_t range = 0;
#pragma unroll
  for (uint32_t k = 0; k < N; ++k) {
    const uint32_t b = e[k] >> shift;
    l[k] = br.bucket[b];
    hi[k] = br.bucket[b + 1];
  }
#pragma unroll
  for (uint32_t k = 0; k < N; ++k) range = max(range, hi[k] - l[k]);
  for (uint32_t st = range ? 1u << (31 - __builtin_clz(range)) : 0u; st > 0; st >>= 1) {
    uint32_t c[N], v[N];
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) c[k] = min(l[k] + st, hi[k]);
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) v[k] = br.prefix[c[k]];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (uint32_t k = 0; k < N; ++k) l[k] = v[k] <= e[k] ? c[k] : l[k];
  }
#pragma unroll
  for (uint32_t k = 0; k < N; ++k) idx[k] = br.start[l[k]] + (e[k] - br.prefix[l[k]]);
}

// Joins the path-0 work units, one block per unit: the item's whole build
// side (at most CAP * 3 / 4 tuples) goes into a block-shared LDS table, then
// the unit's probe tuples are streamed against it.  Work is split by tuples,
// not runs: a batch of up to BLOCK runs is published block-wide and every
// thread takes elements tid, tid + BLOCK, ... of it, SPL loads in flight, so
// a unit of a few long runs (a heavy key, a small build side against a long
// probe side) keeps every wave busy.
template <typename K, uint32_t CAP, uint32_t BLOCK, uint32_t SPL>
__global__ __launch_bounds__(BLOCK) void join_block_kernel(PipelineDevice d) {
  using T = Tuple<K>;
  extern __shared__ uint64_t lds_u64[];
  __shared__ uint32_t s_unit;
  __shared__ uint32_t wave_sums[BLOCK / 64];
  K* const tkeys = reinterpret_cast<K*>(lds_u64);
  const WaveTable<K, K, CAP> tab{tkeys, tkeys + CAP};
  uint64_t* const run_start = reinterpret_cast<uint64_t*>(tkeys + 2 * CAP);
  uint32_t* const run_prefix = reinterpret_cast<uint32_t*>(run_start + BLOCK);
  const BlockRuns br{run_start, run_prefix, reinterpret_cast<uint16_t*>(run_prefix + BLOCK)};
  const uint32_t tid = threadIdx.x;
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const T* rel_b = static_cast<const T*>(d.bits2 ? d.build.rel2 : d.build.rel1);
  const T* rel_p = static_cast<const T*>(d.bits2 ? d.probe.rel2 : d.probe.rel1);
  uint64_t cnt = 0, sb = 0, sp = 0;
  const uint32_t n = uniform_load(d.work_counter + kOverflowAny) ? uniform_load(d.work_counter + kUnitCount) : 0;
  auto clear = [&]() {
    uint4* k4 = reinterpret_cast<uint4*>(tkeys);
    for (uint32_t i = tid; i < CAP * sizeof(K) / 16; i += BLOCK) k4[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
  };
  if (n) {
    clear();
    for (;;) {
      if (tid == 0) s_unit = atomicAdd(d.work_counter + kBlockTicket, 1u);
      __syncthreads();  // also orders the clear before the inserts
      const uint32_t u = s_unit;
      __syncthreads();  // every thread has read s_unit
      if (u >= n) break;
      const uint4 unit = d.units[u];
      const uint32_t item = __builtin_amdgcn_readfirstlane(unit.x);
      const uint32_t p = item / F2, q = item % F2;
      const uint32_t pr0 = __builtin_amdgcn_readfirstlane(unit.y), pr1 = __builtin_amdgcn_readfirstlane(unit.z);
      const uint32_t nb = run_count(d, d.build, p);
      for (uint32_t r0 = 0; r0 < nb; r0 += BLOCK) {
        const uint32_t tb = publish_block_runs<BLOCK>(d, d.build, p, q, r0, nb, br, wave_sums);
        const uint32_t s0 = publish_block_buckets<BLOCK>(br, min(nb - r0, BLOCK), tb);
        for (uint32_t e0 = 0; e0 < tb; e0 += 2 * BLOCK) {
          uint32_t e[2];
          uint64_t idx[2];
          T t[2];
#pragma unroll
          for (uint32_t k = 0; k < 2; ++k) e[k] = min(e0 + k * BLOCK + tid, tb - 1);
          block_elements(br, e, s0, idx);
#pragma unroll
          for (uint32_t k = 0; k < 2; ++k) t[k] = load_tuple<K>(rel_b + idx[k]);
#pragma unroll
          for (uint32_t k = 0; k < 2; ++k)
            if (e0 + k * BLOCK + tid < tb && t[k].key != null_key<K>()) tab.insert(t[k].key, t[k].value);
        }
        __syncthreads();  // inserts done; the batch may be overwritten
      }
      for (uint32_t r0 = pr0; r0 < pr1; r0 += BLOCK) {
        const uint32_t tp = publish_block_runs<BLOCK>(d, d.probe, p, q, r0, pr1, br, wave_sums);
        const uint32_t s0 = publish_block_buckets<BLOCK>(br, min(pr1 - r0, BLOCK), tp);
        for (uint32_t e0 = 0; e0 < tp; e0 += SPL * BLOCK) {
          uint32_t e[SPL];
          uint64_t idx[SPL];
          T t[SPL];
#pragma unroll
          for (uint32_t k = 0; k < SPL; ++k) e[k] = min(e0 + k * BLOCK + tid, tp - 1);
          block_elements(br, e, s0, idx);
#pragma unroll
          for (uint32_t k = 0; k < SPL; ++k) t[k] = load_tuple<K>(rel_p + idx[k]);
#pragma unroll
          for (uint32_t k = 0; k < SPL; ++k)
            if (e0 + k * BLOCK + tid < tp && t[k].key != null_key<K>()) tab.probe(t[k].key, t[k].value, cnt, sb, sp);
        }
        __syncthreads();  // probes done; the batch (and then the table) may be overwritten
      }
      if (d.job_results) wave_job_aggregate(d, p >> d.job_shift, cnt, sb, sp);
      clear();
    }
  }
  if (!d.job_results) join_block_aggregate<BLOCK / 64>(d, cnt, sb, sp);
}

// ----------------------------------------------------- HBM hash tables ----
// A deferred item whose build side exceeds a block's LDS table gets a
// linear-probing table of 2^lg slots (at least twice its build tuples) in
// the plan's HBM arena: its build units insert the build side once, then its
// probe units probe it, so the cost is linear in both sides (the LDS path
// refills a wave table once per RCAP build tuples and streams the unit's
// probe runs against every fill).  Hashing and probing follow the
// no-partitioning join (no_partitioning_join.cu:39-249): multiply-shift from
// the key, linear probing, every match counted; i32 entries are installed
// by one 64-bit CAS of key and payload against an all-ones entry, i64 by a
// CAS on the key and then the payload store.  The arena is all ones between
// joins: the clear kernel restores the slots a join reserved.

template <typename K>
__device__ __forceinline__ void hbm_insert(Tuple<K>* tab, uint32_t lg, K key, K val) {
  const uint64_t mask = (1ull << lg) - 1;
  uint64_t h = mult_shift_hash<K>(key, lg);
  for (uint64_t a = 0; a <= mask; ++a, h = (h + 1) & mask) {
    if constexpr (sizeof(K) == 4) {
      const unsigned long long want = static_cast<unsigned long long>(static_cast<uint32_t>(key)) |
                                      (static_cast<unsigned long long>(static_cast<uint32_t>(val)) << 32);
      if (atomicCAS(reinterpret_cast<unsigned long long*>(tab + h), ~0ull, want) == ~0ull) return;
    } else {
      if (atomicCAS(reinterpret_cast<unsigned long long*>(&tab[h].key), ~0ull, static_cast<unsigned long long>(key)) ==
          ~0ull) {
        tab[h].value = val;
        return;
      }
    }
  }
}

// Every build tuple with this key matches; probing ends at the first free
// slot.  (The i64 payload of a slot whose key was just installed may still
// be in flight only while the build kernel runs; probes run in a later
// launch.)
template <typename K>
__device__ __forceinline__ void hbm_probe(const Tuple<K>* tab, uint32_t lg, K key, K pay, uint64_t& cnt, uint64_t& sb,
                                          uint64_t& sp) {
  const uint64_t mask = (1ull << lg) - 1;
  uint64_t h = mult_shift_hash<K>(key, lg);
  for (uint64_t a = 0; a <= mask; ++a, h = (h + 1) & mask) {
    const Tuple<K> e = load_tuple<K>(tab + h);
    if (e.key == null_key<K>()) return;
    if (e.key == key) {
      cnt += 1;
      sb += static_cast<uint64_t>(static_cast<int64_t>(e.value));
      sp += static_cast<uint64_t>(static_cast<int64_t>(pay));
    }
  }
}

// One block per HBM work unit (tickets over build units when BUILD, else over
// probe units): the unit's runs of its item's build (probe) side, published
// block-wide in batches of BLOCK runs, SPL gathers in flight per thread.
template <typename K, uint32_t BLOCK, uint32_t SPL, bool BUILD>
__global__ __launch_bounds__(BLOCK) void join_hbm_kernel(PipelineDevice d) {
  using T = Tuple<K>;
  __shared__ uint32_t s_unit;
  __shared__ uint32_t wave_sums[BLOCK / 64];
  __shared__ uint64_t run_start[BLOCK];
  __shared__ uint32_t run_prefix[BLOCK];
  __shared__ uint16_t run_bucket[2 * BLOCK + 4];
  const BlockRuns br{run_start, run_prefix, run_bucket};
  const uint32_t tid = threadIdx.x;
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const PipelineSide& side = BUILD ? d.build : d.probe;
  const T* rel = static_cast<const T*>(d.bits2 ? side.rel2 : side.rel1);
  T* const arena = static_cast<T*>(d.hbm_table);
  uint64_t cnt = 0, sb = 0, sp = 0;
  const uint32_t n = uniform_load(d.work_counter + kOverflowAny)
                         ? uniform_load(d.work_counter + (BUILD ? kHbmBuildUnits : kHbmProbeUnits))
                         : 0u;
  for (uint32_t u = 0; n;) {
    if (tid == 0) s_unit = atomicAdd(d.work_counter + (BUILD ? kHbmBuildTicket : kHbmProbeTicket), 1u);
    __syncthreads();
    u = s_unit;
    __syncthreads();  // every thread has read s_unit
    if (u >= n) break;
    const uint4 unit = d.hbm_units[BUILD ? u : d.hbm_unit_cap - 1 - u];
    const uint4 big = d.big[__builtin_amdgcn_readfirstlane(unit.x)];
    const uint32_t item = __builtin_amdgcn_readfirstlane(big.x);
    const uint32_t p = item / F2, q = item % F2;
    T* const tab = arena + __builtin_amdgcn_readfirstlane(big.y);
    const uint32_t lg = __builtin_amdgcn_readfirstlane(big.z);
    const uint32_t r_end = __builtin_amdgcn_readfirstlane(unit.z);
    for (uint32_t r0 = __builtin_amdgcn_readfirstlane(unit.y); r0 < r_end; r0 += BLOCK) {
      const uint32_t tn = publish_block_runs<BLOCK>(d, side, p, q, r0, r_end, br, wave_sums);
      const uint32_t s0 = publish_block_buckets<BLOCK>(br, min(r_end - r0, BLOCK), tn);
      for (uint32_t e0 = 0; e0 < tn; e0 += SPL * BLOCK) {
        uint32_t e[SPL];
        uint64_t idx[SPL];
        T t[SPL];
#pragma unroll
        for (uint32_t k = 0; k < SPL; ++k) e[k] = min(e0 + k * BLOCK + tid, tn - 1);
        block_elements(br, e, s0, idx);
#pragma unroll
        for (uint32_t k = 0; k < SPL; ++k) t[k] = load_tuple<K>(rel + idx[k]);
#pragma unroll
        for (uint32_t k = 0; k < SPL; ++k) {
          if (e0 + k * BLOCK + tid >= tn || t[k].key == null_key<K>()) continue;  // null keys never join
          if constexpr (BUILD)
            hbm_insert<K>(tab, lg, t[k].key, t[k].value);
          else
            hbm_probe<K>(tab, lg, t[k].key, t[k].value, cnt, sb, sp);
        }
      }
      __syncthreads();  // the batch may be overwritten
    }
    if (!BUILD && d.job_results) wave_job_aggregate(d, p >> d.job_shift, cnt, sb, sp);
  }
  if (!BUILD && !d.job_results) join_block_aggregate<BLOCK / 64>(d, cnt, sb, sp);
}

// Restores the slots reserved in this join to all ones (16-byte stores).
__global__ __launch_bounds__(256) void join_hbm_clear_kernel(PipelineDevice d, uint32_t tuple_bytes) {
  if (uniform_load(d.work_counter + kOverflowAny) == 0) return;
  const uint64_t reserved = *reinterpret_cast<const volatile uint64_t*>(d.work_counter + kHbmSlots);
  const uint64_t slots = min<uint64_t>(reserved, d.hbm_cap);
  uint4* w = static_cast<uint4*>(d.hbm_table);
  const uint64_t n16 = slots * tuple_bytes / 16;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    w[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
}

// One wave per work item (p, q).  Items are handed out in tickets of up to 8
// consecutive items.  With XCD = true, work group g draws first from the
// items of pass-1 partitions p = g mod 8 (its XCD's share: neighbouring runs
// of consecutive q share cache lines, which then stay in that XCD's L2) and
// then from the other shares.  The common item (at most 64 runs per side, at
// most BPL * 64 build and SPL * 64 probe tuples) is software-pipelined: its
// gathers are issued, then the previous item is built and probed from
// registers while they are in flight, and the next item's run descriptors
// are fetched behind them.
template <typename K, uint32_t CAP, uint32_t WAVES, uint32_t BPL, uint32_t SPL, bool NARROW, bool XCD>
__global__ __launch_bounds__(WAVES * 64) void join_kernel(PipelineDevice d) {
  using T = Tuple<K>;
  constexpr uint32_t RCAP = CAP * 3 / 4;  // build tuples per table fill (load factor <= 3/4)
  constexpr uint32_t FAST_B = RCAP < BPL * 64 ? RCAP : BPL * 64;
  static_assert(WAVES <= 16, "aggregate scratch");
  extern __shared__ uint64_t lds_u64[];
  const WaveLds<K, CAP> l(lds_u64);
  K* const table_keys = l.table_keys;
  const WaveRuns& rb = l.rb;
  const WaveRuns& rp = l.rp;
  const uint32_t lane = lane_id();
  static_assert(BPL * 64 <= kMarkBytes && SPL * 64 <= kMarkBytes, "run marks");
  WaveTable<K, K, CAP>{table_keys, table_keys + CAP}.clear();
  for (uint32_t i = lane; i < 2 * kMarkBytes / 4; i += 64) reinterpret_cast<uint32_t*>(l.mb)[i] = 0;
  wave_lds_sync();
  const T* rel_b = static_cast<const T*>(d.bits2 ? d.build.rel2 : d.build.rel1);
  const T* rel_p = static_cast<const T*>(d.bits2 ? d.probe.rel2 : d.probe.rel1);
  const uint32_t F1 = d.F1;
  const uint32_t F2 = d.bits2 ? d.F2 : 1;
  const uint32_t shift = d.bits1 + d.bits2;
  const uint32_t items = F1 * F2;
  const T* const dummy = reinterpret_cast<const T*>(d.aggregates);  // 16 readable bytes, value unused
  const uint32_t per_wave = items / (gridDim.x * WAVES * 16);
  const uint32_t tk = per_wave < 1 ? 1 : (per_wave > 8 ? 8 : per_wave);
  // Tickets: group g's share is a list of items; ticket t of it holds the tk
  // items j0 + R * s (s < tk), j0 = (t / R) * R * tk + t % R, R = the waves
  // drawing from one share.  So the R waves of a share work on R adjacent
  // items at a time, and the cache lines two adjacent items' runs share are
  // read once into that XCD's L2 rather than once per item.
  constexpr uint32_t NG = XCD ? kJoinCounters : 1;
  const uint32_t g0 = XCD ? blockIdx.x % kJoinCounters : 0;
  const uint32_t R = max(1u, gridDim.x * WAVES / NG);
  uint32_t gi = 0, grp = g0, tbase = 0, ts = tk, tlen = 0;
  auto next_item = [&]() -> uint32_t {
    for (;;) {
      if (ts < tk) {
        const uint32_t j = tbase + R * ts++;
        if (j < tlen) return XCD ? (grp + kJoinCounters * (j / F2)) * F2 + j % F2 : j;
        ts = tk;  // later items of this ticket are past the share's end too
      }
      if (gi >= NG) return items;
      const uint32_t g = XCD ? (g0 + gi) % kJoinCounters : 0;
      const uint32_t len = XCD ? (g < F1 ? (F1 - g + kJoinCounters - 1) / kJoinCounters * F2 : 0) : items;
      uint32_t t = 0;
      if (lane == 0 && len) t = atomicAdd(d.work_counter + g * kJoinCounterStride, 1u);
      t = __builtin_amdgcn_readfirstlane(t);
      const uint64_t j0 = static_cast<uint64_t>(t / R) * R * tk + t % R;
      if (len && j0 < len) {
        grp = g;
        tbase = static_cast<uint32_t>(j0);
        ts = 0;
        tlen = len;
      } else {
        ++gi;
      }
    }
  };
  uint64_t cnt = 0, sb = 0, sp = 0;
  uint32_t nb = 0, np = 0;
  RawRun xb, xp;
  // Run descriptors of item i (i < items).  Branch-free: issued behind the
  // current item's gathers without making them wait.
  auto fetch = [&](uint32_t i) {
    const uint32_t p = i / F2, q = i % F2;
    nb = run_count(d, d.build, p);
    np = run_count(d, d.probe, p);
    xb = raw_run(d, d.build, p, q, 0, nb);
    xp = raw_run(d, d.probe, p, q, 0, np);
  };
  uint32_t cur = next_item();
  if (cur < items) fetch(cur);
  bool pend = false;  // an item's tuples are in flight in the work buffer
  uint32_t pend_tb = 0, pend_tp = 0;
  // batched plans: the join whose result cnt / sb / sp belong to, and that
  // of the pending item
  const bool batched = d.job_results != nullptr;
  uint32_t acc_job = 0, pend_job = 0;
  // One pipeline step: stage item `cur` into (sbuf, spbuf) and finish the
  // pending item held in (wbuf, wpbuf).  Returns false when no items are
  // left (the pending item, if any, is finished first).  The build/probe has
  // a single call site, which keeps the register allocation of the two
  // buffers simple.
  auto step = [&](T (&sbuf)[BPL], T (&spbuf)[SPL], T (&wbuf)[BPL], T (&wpbuf)[SPL]) -> bool {
    if (batched && pend && pend_job != acc_job) {
      wave_job_aggregate(d, acc_job, cnt, sb, sp);
      acc_job = pend_job;
    }
    const bool live = cur < items;
    bool staged = false;
    uint32_t tb = 0, tp = 0, nxt = items;
    uint64_t ib[BPL], ip[SPL];
    const T* gb = dummy;
    const T* gp = dummy;
#pragma unroll
    for (uint32_t k = 0; k < BPL; ++k) ib[k] = 0;
#pragma unroll
    for (uint32_t k = 0; k < SPL; ++k) ip[k] = 0;
    if (live) {
      const LaneRun lb = finish_run(d, xb, 0, nb), lp = finish_run(d, xp, 0, np);
      tb = publish_runs(lb, rb);
      tp = publish_runs(lp, rp);
      wave_lds_sync();
      nxt = next_item();
      const bool empty = (tb == 0 && nb <= kMaxRuns) || (tp == 0 && np <= kMaxRuns);  // tb, tp: first 64 runs
      const bool fast = nb <= kMaxRuns && np <= kMaxRuns && tb <= FAST_B && tp <= SPL * 64;
      if (!empty && fast) {
        // gather indices (lanes past the end repeat the last tuple)
        run_indices<BPL>(rb, l.mb, lb, tb, ib);
        run_indices<SPL>(rp, l.mp, lp, tp, ip);
        gb = rel_b;
        gp = rel_p;
        staged = true;
      } else if (!empty && lane == 0) {
        // does not fit the registers: deferred to join_units_kernel, which
        // hands it to join_block_kernel or join_general_kernel (an empty
        // item has no match)
        d.overflow[cur] = 1;
        d.work_counter[kOverflowAny] = 1;
      }
    }
    // Every step issues the same gathers back to back (a step that stages
    // nothing reads a fixed in-bounds word), so the wait before the pending
    // item's build counts exactly this step's loads on every path: an LDS
    // read, a branch or a skipped load between gathers would make the build
    // wait for this step's gathers too.
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (uint32_t k = 0; k < BPL; ++k) sbuf[k] = load_tuple<K>(gb + ib[k]);
#pragma unroll
    for (uint32_t k = 0; k < SPL; ++k) spbuf[k] = load_tuple<K>(gp + ip[k]);
    if (live) fetch(min(nxt, items - 1));
    if (pend) join_in_registers<K, CAP, BPL, SPL, NARROW>(table_keys, wbuf, wpbuf, pend_tb, pend_tp, shift, cnt, sb, sp);
    pend = staged;
    pend_tb = tb;
    pend_tp = tp;
    if (batched && staged) pend_job = (cur / F2) >> d.job_shift;
    cur = nxt;
    return live;
  };
  T ab[BPL], ap[SPL], bb[BPL], bp[SPL];
  for (;;) {
    if (!step(ab, ap, bb, bp)) break;
    if (!step(bb, bp, ab, ap)) break;
  }
  if (batched)
    wave_job_aggregate(d, acc_job, cnt, sb, sp);
  else
    join_block_aggregate<WAVES>(d, cnt, sb, sp);
}

template <typename K, uint32_t CAP, uint32_t WAVES>
size_t join_lds_bytes() {
  return WAVES * join_wave_bytes<K, CAP>();
}

template <typename K, uint32_t TILE2>
size_t sort2_lds_bytes(uint32_t F2) {
  return TILE2 * sizeof(Tuple<K>) + F2 * 4 + 16;
}

// dst = the sum of src[0..n) (one block; wrapping 64-bit sums).
__global__ __launch_bounds__(256) void sum_aggregates_kernel(const JoinAggregate* src, uint32_t n,
                                                             JoinAggregate* dst) {
  __shared__ uint64_t red[3][4];
  uint64_t c = 0, b = 0, p = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    c += src[i].count;
    b += src[i].sum_build;
    p += src[i].sum_probe;
  }
  c = wave_broadcast_last(wave_inclusive_scan(c));
  b = wave_broadcast_last(wave_inclusive_scan(b));
  p = wave_broadcast_last(wave_inclusive_scan(p));
  const uint32_t w = threadIdx.x / 64;
  if (lane_id() == 0) {
    red[0][w] = c;
    red[1][w] = b;
    red[2][w] = p;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    dst->count = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    dst->sum_build = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    dst->sum_probe = red[2][0] + red[2][1] + red[2][2] + red[2][3];
  }
}

}  // namespace

// ------------------------------------------------------------- launches ----

hipError_t launch_sum_aggregates(const JoinAggregate* src, uint32_t n, JoinAggregate* dst, hipStream_t stream) {
  hipLaunchKernelGGL(sum_aggregates_kernel, dim3(1), dim3(256), 0, stream, src, n, dst);
  return hipGetLastError();
}

uint32_t pipeline_block_table_tuples(KeyType t, bool big) {
  return (t == KeyType::I32 ? kBlockTableSlots<int32_t> : kBlockTableSlots<int64_t>) * (big ? 2 : 1) * 3 / 4;
}

uint32_t pipeline_tile2(KeyType t) { return t == KeyType::I32 ? kTile2<int32_t> : kTile2<int64_t>; }

hipError_t launch_pipeline_plan(const PipelineDevice& d, hipStream_t stream) {
  const uint32_t blocks = d.sides == 3 ? 2u : 1u;
  hipLaunchKernelGGL(plan_kernel, dim3(blocks), dim3(1024), 2 * d.F1 * 8, stream, d);
  return hipGetLastError();
}

hipError_t launch_pipeline_plan_batched(const PipelineDevice& d, const BatchSide* sides, uint32_t joins,
                                        uint32_t* side_tiles, uint64_t rel2_total_b, uint64_t rel2_total_p,
                                        hipStream_t stream) {
  if (joins == 0) return hipSuccess;
  const uint32_t F1j = d.F1 / joins;
  hipLaunchKernelGGL(plan_batched_kernel, dim3(2 * joins), dim3(1024), 2 * F1j * 8, stream, d, sides, F1j,
                     side_tiles, rel2_total_b, rel2_total_p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(tile_bases_kernel, dim3(1), dim3(1024), 0, stream, d, side_tiles, joins);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(tile_offset_kernel, dim3((2 * d.F1 + 255) / 256), dim3(256), 0, stream, d,
                     static_cast<const uint32_t*>(side_tiles), F1j);
  return hipGetLastError();
}

template <typename K>
hipError_t launch_sort2_t(const PipelineDevice& d, uint32_t grid, hipStream_t stream) {
  constexpr uint32_t TILE2 = kTile2<K>;
  constexpr uint32_t BLOCK = 1024;
  auto kern = sort2_kernel<K, TILE2, BLOCK>;
  const size_t lds = sort2_lds_bytes<K, TILE2>(d.F2);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), lds, stream, d);
  return hipGetLastError();
}

hipError_t launch_pipeline_sort2(const PipelineDevice& d, KeyType t, uint32_t grid, hipStream_t stream) {
  if (grid == 0) return hipSuccess;
  if (t == KeyType::I32) return launch_sort2_t<int32_t>(d, grid, stream);
  return launch_sort2_t<int64_t>(d, grid, stream);
}

size_t pipeline_sort2_lds(KeyType t, uint32_t F2) {
  if (t == KeyType::I32) return sort2_lds_bytes<int32_t, kTile2<int32_t>>(F2);
  return sort2_lds_bytes<int64_t, kTile2<int64_t>>(F2);
}

template <typename K, uint32_t CAP, uint32_t WAVES, uint32_t BPL, uint32_t SPL>
hipError_t launch_register_join(const PipelineDevice& d, uint32_t grid, hipStream_t stream) {
  const size_t lds = join_lds_bytes<K, CAP, WAVES>();
  auto kern = join_kernel<K, CAP, WAVES, BPL, SPL, true, true>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(WAVES * 64), lds, stream, d);
  return hipGetLastError();
}

// Tuples per lane of one side of the register join (its gathers per item):
// every step issues that many gathers whether the item fills them or not,
// so a side whose items are small gets fewer.  The need is twice the
// average item side (joins of one size class differ in size by up to 2x)
// plus four standard deviations; 10 per lane (640 tuples) at most, which
// with two buffers per side keeps 2 waves per SIMD at i64.  An item beyond
// its side's capacity is deferred like any other large item.
static uint32_t register_join_lanes(uint64_t side_tuples, uint64_t items) {
  if (items == 0) return 10;
  const double avg = 2.0 * static_cast<double>(side_tuples) / static_cast<double>(items);
  const double need = avg + 4.0 * std::sqrt(avg) + 16.0;
  return need <= 128.0 ? 2u : need <= 256.0 ? 4u : 10u;
}

// The wave table holds a whole build side of the register join (at most
// BPL * 64 tuples) at a load factor of at most 3/4: 256 slots for 2 tuples
// per lane, 512 for 4, CAP (1024) for 10.  A smaller table leaves LDS for
// more waves per CU.
template <typename K, uint32_t CAP, uint32_t WAVES>
hipError_t launch_register_join_sized(const PipelineDevice& d, uint32_t grid, hipStream_t stream) {
  const uint64_t items = static_cast<uint64_t>(d.F1) * (d.bits2 ? d.F2 : 1u);
  const uint32_t bpl = register_join_lanes(d.build.n, items), spl = register_join_lanes(d.probe.n, items);
  if (bpl == 10 && spl == 10) return launch_register_join<K, CAP, WAVES, 10, 10>(d, grid, stream);
  if (spl == 10) {
    if (bpl == 2) return launch_register_join<K, CAP / 4, WAVES, 2, 10>(d, grid, stream);
    return launch_register_join<K, CAP / 2, WAVES, 4, 10>(d, grid, stream);
  }
  if (bpl == 10) {
    if (spl == 2) return launch_register_join<K, CAP, WAVES, 10, 2>(d, grid, stream);
    return launch_register_join<K, CAP, WAVES, 10, 4>(d, grid, stream);
  }
  return launch_register_join<K, CAP / 2, WAVES, 4, 4>(d, grid, stream);
}

// The deferred items' joins with block-shared tables of BCAP slots: 64 KiB
// (two blocks of 512 threads per CU) or, for single-pass plans whose build
// partitions need it, 128 KiB (one block per CU).
template <typename K, uint32_t BCAP, uint32_t GCAP, uint32_t GWAVES>
hipError_t launch_deferred_t(const PipelineDevice& d, uint32_t grid, hipStream_t stream) {
  constexpr uint32_t BBLOCK = 512;
  hipLaunchKernelGGL(join_units_kernel, dim3(grid), dim3(kUnitsWaves * 64), 0, stream, d, BCAP * 3 / 4);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  {
    auto blk = join_block_kernel<K, BCAP, BBLOCK, 16>;
    const size_t lds = BCAP * 2 * sizeof(K) + block_runs_bytes<BBLOCK>();
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(blk), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds));
    hipLaunchKernelGGL(blk, dim3(grid), dim3(BBLOCK), lds, stream, d);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // deferred items are large: bigger tables, fewer waves (grid <= aggregates)
  const size_t lds = join_lds_bytes<K, GCAP, GWAVES>();
  auto gen = join_general_kernel<K, GCAP, GWAVES, 12, 8>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(gen), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(gen, dim3(grid), dim3(GWAVES * 64), lds, stream, d);
  e = hipGetLastError();
  if (e != hipSuccess || d.hbm_cap == 0) return e;
  // items with an HBM table: build, probe, then restore the arena
  hipLaunchKernelGGL((join_hbm_kernel<K, 512, 8, true>), dim3(grid), dim3(512), 0, stream, d);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((join_hbm_kernel<K, 512, 8, false>), dim3(grid), dim3(512), 0, stream, d);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(join_hbm_clear_kernel, dim3(grid), dim3(256), 0, stream, d,
                     static_cast<uint32_t>(sizeof(Tuple<K>)));
  return hipGetLastError();
}

template <typename K, uint32_t CAP, uint32_t WAVES, uint32_t GCAP, uint32_t GWAVES>
hipError_t launch_join_t(const PipelineDevice& d, bool deferred, uint32_t grid, hipStream_t stream) {
  if (!deferred) {
    // large tables (single-pass plans): one shape
    if constexpr (CAP > 1024)
      return launch_register_join<K, CAP, WAVES, 10, 10>(d, grid, stream);
    else
      return launch_register_join_sized<K, CAP, WAVES>(d, grid, stream);
  }
  if constexpr (CAP > 1024) {
    if (d.big_block_tables) return launch_deferred_t<K, 2 * kBlockTableSlots<K>, GCAP, GWAVES>(d, grid, stream);
  }
  return launch_deferred_t<K, kBlockTableSlots<K>, GCAP, GWAVES>(d, grid, stream);
}

hipError_t launch_pipeline_join(const PipelineDevice& d, KeyType t, bool large_tables, bool deferred, uint32_t grid,
                                hipStream_t stream) {
  if (t == KeyType::I32) {
    if (large_tables) return launch_join_t<int32_t, 8192, 2, 8192, 2>(d, deferred, grid, stream);
    return launch_join_t<int32_t, 1024, 4, 4096, 4>(d, deferred, grid, stream);
  }
  if (large_tables) return launch_join_t<int64_t, 4096, 2, 4096, 2>(d, deferred, grid, stream);
  return launch_join_t<int64_t, 1024, 4, 2048, 4>(d, deferred, grid, stream);
}

}  // namespace fij
