// Radix scatter kernels.
//
// Semantics follow the reference's chunked scatters
// (sql-ops/cudautils/radix_partition.cu:472-525 NC, :688-920 SSWWCv2): each
// block owns one input chunk and writes its tuples, in row layout, into the
// partition ranges that start at chunk_offsets[chunk][p].  Order within a
// partition is not defined (as in the reference).  Padding in front of each
// partition is never written.
//
// The buffered variant is designed for wave64 / 160 KB LDS.  A block's waves
// have one of two roles, so that no wave both loads and stores global
// memory: on gfx950 one counter (vmcnt) tracks a wave's loads and stores in
// issue order, and a wave that waits for a prefetched load would also wait
// for every store it issued before that load.
//   loaders:  load the chunk's key and payload columns (16 bytes per lane
//             and column) eight steps ahead, take each tuple's position
//             with one 64-bit LDS atomic on its partition's {fill, limit}
//             word, and write the tuple into the partition's LDS ring of R
//             aligned 128-byte lines; a tuple that completes a line queues
//             the line (one queue reservation per wave and step);
//   flushers: write the lines queued in the previous step, 64/L lines per
//             non-temporal store instruction, each a full aligned 128-byte
//             line.
// Steps are separated by two barriers (a step's placement and the flush of
// the previous step's lines, then a commit of the per-partition marks).  A
// line is read by the flushers one step after it was completed, so the ring
// slots a step may fill end RL positions after the first line that was
// still unflushed when the previous step began; a tuple beyond that limit is
// stored directly (rare unless a partition receives more than about RL/2
// tuples per step) and positions below a partition's low-water mark are
// never written from the ring.
#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {


// Destination of partition d's tuples: the relation or, from spill_from on,
// the spill region (both indexed by the relation's tuple index).
template <bool SPILL, typename T>
__device__ __forceinline__ T* part_dst(const ScatterParams& p, T* rel, uint64_t cbase, uint32_t d) {
  if constexpr (SPILL) {
    if (d >= p.spill_from) return reinterpret_cast<T*>(p.spill_addr) + cbase;
  }
  return rel;
}

template <typename K>
__device__ __forceinline__ void chunk_range(const ScatterParams& p, uint32_t chunk, uint32_t G, uint64_t& begin,
                                            uint64_t& len) {
  begin = min(p.chunk_len * chunk, p.n);
  len = (chunk + 1 == G) ? p.n - begin : min(p.chunk_len, p.n - begin);
}

// Direct scatter (reference NC): per-chunk LDS cursors, wave-aggregated ranks.
template <typename K, bool SPILL>
__global__ __launch_bounds__(1024) void scatter_direct_kernel(ScatterParams p) {
  extern __shared__ uint32_t cursors[];
  const uint32_t F = 1u << p.radix_bits;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x, B = blockDim.x;
  const uint64_t* off = p.chunk_offsets + static_cast<uint64_t>(blockIdx.x) * F;
  const uint64_t cbase = off[0];
  for (uint32_t i = tid; i < F; i += B) cursors[i] = static_cast<uint32_t>(off[i] - cbase);
  __syncthreads();
  uint64_t begin, len;
  chunk_range<K>(p, blockIdx.x, gridDim.x, begin, len);
  const K* keys = static_cast<const K*>(p.keys) + begin;
  const K* pays = static_cast<const K*>(p.pays) + begin;
  Tuple<K>* rel = static_cast<Tuple<K>*>(p.relation) + cbase;
  for (uint64_t i0 = 0; i0 < len; i0 += B) {
    const uint64_t i = i0 + tid;
    const bool act = i < len;
    K key = 0, pay = 0;
    if (act) {
      key = keys[i];
      pay = pays[i];
    }
    const uint32_t d = key_to_partition<K>(key, mask, p.ignore_bits);
    const uint32_t r = wave_rank_add(cursors, d, act);
    if (act) store_tuple<K>(part_dst<SPILL>(p, rel, cbase, d) + r, key, pay);
  }
}

// High-fanout scatter (fanouts whose line rings do not fit LDS): the chunk
// is read in tiles of ITEMS * BLOCK tuples; a tile is counting-sorted by
// digit in LDS (ranks from LDS atomics on the tile's counters, a block scan)
// and written out in sorted order, so that the lanes of a store write runs
// of consecutive addresses -- one run per digit of the tile -- instead of
// one address per lane.  Partial lines are merged in the L2 and the memory
// side cache before they reach HBM.  The next tile is loaded while the
// current one is sorted and written.
template <typename K, uint32_t ITEMS, uint32_t BLOCK, bool SPILL>
__global__ __launch_bounds__(BLOCK) void scatter_tile_kernel(ScatterParams p) {
  using T = Tuple<K>;
  constexpr uint32_t TILE = ITEMS * BLOCK;
  extern __shared__ __attribute__((aligned(16))) uint64_t lds_u64[];
  __shared__ uint32_t wave_tmp[BLOCK / kWave + 1];
  const uint32_t F = 1u << p.radix_bits;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x;
  T* sorted = reinterpret_cast<T*>(lds_u64);               // [TILE]
  uint32_t* start = reinterpret_cast<uint32_t*>(sorted + TILE);  // [F] tile counts, then run starts
  uint32_t* cursor = start + F;                             // [F] next position, relative to cbase
  const uint64_t* off = p.chunk_offsets + static_cast<uint64_t>(blockIdx.x) * F;
  const uint64_t cbase = off[0];
  for (uint32_t i = tid; i < F; i += BLOCK) {
    cursor[i] = static_cast<uint32_t>(off[i] - cbase);
    start[i] = 0;
  }
  uint64_t begin, len;
  chunk_range<K>(p, blockIdx.x, gridDim.x, begin, len);
  const K* keys = static_cast<const K*>(p.keys) + begin;
  const K* pays = static_cast<const K*>(p.pays) + begin;
  T* rel = static_cast<T*>(p.relation) + cbase;
  __syncthreads();
  if (len == 0) return;  // uniform
  // lanes past the chunk's end re-read its last tuple and are ignored
  auto load = [&](uint64_t t0, T (&v)[ITEMS]) {
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) {
      const uint64_t e = min(t0 + k * BLOCK + tid, len - 1);
      v[k].key = __builtin_nontemporal_load(keys + e);
      v[k].value = __builtin_nontemporal_load(pays + e);
    }
  };
  auto process = [&](uint64_t t0, const T (&v)[ITEMS]) {
    const uint32_t n = static_cast<uint32_t>(min<uint64_t>(TILE, len - t0));
    uint32_t dig[ITEMS], rank[ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k) {
      const bool act = k * BLOCK + tid < n;
      dig[k] = key_to_partition<K>(v[k].key, mask, p.ignore_bits);
      rank[k] = wave_rank_add(start, dig[k], act);
    }
    __syncthreads();
    block_exclusive_scan<uint32_t>(start, F, wave_tmp);
#pragma unroll
    for (uint32_t k = 0; k < ITEMS; ++k)
      if (k * BLOCK + tid < n) sorted[start[dig[k]] + rank[k]] = v[k];
    __syncthreads();
    // element i of the sorted tile is element i - start[d] of digit d's run
    for (uint32_t i = tid; i < n; i += BLOCK) {
      const T x = sorted[i];
      const uint32_t d = key_to_partition<K>(x.key, mask, p.ignore_bits);
      const uint32_t pos = cursor[d] + (i - start[d]);
      store_tuple<K>(as_global(part_dst<SPILL>(p, rel, cbase, d) + pos), x.key, x.value);
    }
    __syncthreads();
    for (uint32_t i = tid; i < F; i += BLOCK) {
      const uint32_t end = i + 1 < F ? start[i + 1] : n;
      cursor[i] += end - start[i];
    }
    __syncthreads();
    for (uint32_t i = tid; i < F; i += BLOCK) start[i] = 0;
    // (the next tile's ranks are taken after the barrier inside process)
  };
  T va[ITEMS], vb[ITEMS];
  uint64_t t0 = 0;
  load(t0, va);
  auto step = [&](T (&cbuf)[ITEMS], T (&nbuf)[ITEMS]) -> bool {
    const uint64_t tn = t0 + TILE;
    const bool more = tn < len;
    load(more ? tn : t0, nbuf);
    process(t0, cbuf);
    __syncthreads();
    t0 = tn;
    return more;
  };
  for (;;) {
    if (!step(va, vb)) break;
    if (!step(vb, va)) break;
  }
}

template <typename K, uint32_t ITEMS, uint32_t BLOCK>
constexpr uint32_t tile_scatter_lds(uint32_t fanout) {
  return ITEMS * BLOCK * sizeof(Tuple<K>) + 2 * 4 * fanout;
}

// Steps whose loads a loader has in flight.
constexpr uint32_t kScatterPrefetch = 8;

// V keys (and V payloads) per loader lane and step: one 16-byte load each.
// The type is only as aligned as its elements: a column is aligned to its
// element size, and gfx950 loads 16 bytes from any 4-byte aligned address.
template <typename K>
struct Vec16 {
  typedef K type __attribute__((ext_vector_type(16 / sizeof(K)), aligned(sizeof(K))));
};

// Buffered scatter of chunk `chunk` of G (one block).
template <typename K, int BLOCK, bool SPILL>
__device__ __forceinline__ void scatter_swwc_body(const ScatterParams& p, uint32_t chunk, uint32_t G) {
  using T = Tuple<K>;
  using VK = typename Vec16<K>::type;
  constexpr uint32_t V = 16 / sizeof(K);          // tuples per loader lane and step
  constexpr uint32_t L = kAlignBytes / sizeof(T);  // tuples per 128-byte line
  constexpr uint32_t NW = BLOCK / kWave;
  constexpr uint32_t NLW = NW / 2;                 // loader waves
  constexpr uint32_t NFW = NW - NLW;               // flusher waves
  constexpr uint32_t NLT = NLW * kWave;            // loader lanes
  constexpr uint32_t LPW = kWave / L;              // lines per flusher store
  constexpr uint32_t UF = L / 4;                   // flusher stores per round
  constexpr uint32_t D = kScatterPrefetch;         // steps loaded ahead
  extern __shared__ __attribute__((aligned(16))) uint64_t lds_u64[];
  __shared__ uint32_t qcount[2];

  const uint32_t F = 1u << p.radix_bits;
  const uint32_t R = p.ring_lines;
  const uint32_t RL = R * L;
  const uint32_t QL = p.queue_len;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x;
  const uint32_t w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const uint32_t lane = lane_id();

  // Positions are tuple indices from `org`, a 128-byte aligned address at or
  // below the chunk's first partition: a position is line aligned when it is
  // a multiple of L, and a partition's ring slot for position P is P mod RL.
  T* ring = reinterpret_cast<T*>(lds_u64);  // [F][RL]
  uint2* queue = reinterpret_cast<uint2*>(ring + static_cast<size_t>(F) * RL);  // [2][QL] {partition, line}
  uint64_t* fl = reinterpret_cast<uint64_t*>(queue + 2 * static_cast<size_t>(QL));  // {fill, lim}
  uint32_t* fprev = reinterpret_cast<uint32_t*>(fl + F);  // fill at the end of the previous step
  uint32_t* ring_lo = fprev + F;       // positions below are never written from the ring
  uint32_t* ring_lo_prev = ring_lo + F;  // ring_lo during the previous step

  const uint64_t* off = p.chunk_offsets + static_cast<uint64_t>(chunk) * F;
  const uint64_t cbase = off[0];
  T* const rel = static_cast<T*>(p.relation) + cbase;
  const uint32_t a = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(rel) % kAlignBytes) / sizeof(T));
  // Destination of partition d's position P: org(d) + P.
  T* const org_rel = rel - a;
  T* const org_spill = SPILL ? reinterpret_cast<T*>(p.spill_addr) + cbase - a : org_rel;
  auto org = [&](uint32_t d) -> T* {
    if constexpr (SPILL) return d >= p.spill_from ? org_spill : org_rel;
    return org_rel;
  };
  for (uint32_t i = tid; i < F; i += BLOCK) {
    const uint32_t s0 = static_cast<uint32_t>(off[i] - cbase) + a;
    fl[i] = static_cast<uint64_t>(s0) | (static_cast<uint64_t>((s0 & ~(L - 1)) + RL) << 32);
    fprev[i] = s0;
    ring_lo[i] = s0;
    ring_lo_prev[i] = s0;
  }
  if (tid < 2) qcount[tid] = 0;
  __syncthreads();

  uint64_t begin, len;
  chunk_range<K>(p, chunk, G, begin, len);
  const K* keys = static_cast<const K*>(p.keys) + begin;
  const K* pays = static_cast<const K*>(p.pays) + begin;
  const uint64_t nvec = len / V;

  // Loader: places up to V tuples of one lane (act: bit k = tuple k is live)
  // and queues the lines they complete into queue buffer qb.  One LDS round
  // trip: the 64-bit atomic returns the position and the step's limit.
  auto place = [&](const VK& key, const VK& pay, uint32_t act, uint32_t qb) {
    uint32_t d[V], pos[V], lm[V];
#pragma unroll
    for (uint32_t k = 0; k < V; ++k) {
      d[k] = key_to_partition<K>(key[k], mask, p.ignore_bits);
      pos[k] = 0;
      lm[k] = 0;
      if ((act >> k) & 1u) {
        const uint64_t o = atomicAdd(reinterpret_cast<unsigned long long*>(&fl[d[k]]), 1ull);
        pos[k] = static_cast<uint32_t>(o);
        lm[k] = static_cast<uint32_t>(o >> 32);
      }
    }
    uint64_t cm[V];
    uint32_t total = 0;
#pragma unroll
    for (uint32_t k = 0; k < V; ++k) {
      bool complete = false;
      if ((act >> k) & 1u) {
        if (pos[k] < lm[k]) {
          T t;
          t.key = key[k];
          t.value = pay[k];
          ring[static_cast<size_t>(d[k]) * RL + (pos[k] & (RL - 1))] = t;
          complete = (pos[k] & (L - 1)) == L - 1;
        } else {
          store_tuple<K>(as_global(org(d[k]) + pos[k]), key[k], pay[k]);
        }
      }
      cm[k] = ballot(complete);
      total += static_cast<uint32_t>(__popcll(cm[k]));
    }
    if (total) {
      uint32_t q0 = 0;
      if (lane == 0) q0 = atomicAdd(&qcount[qb], total);
      q0 = __builtin_amdgcn_readfirstlane(q0) + qb * QL;
#pragma unroll
      for (uint32_t k = 0; k < V; ++k) {
        if ((cm[k] >> lane) & 1u) queue[q0 + mbcnt(cm[k])] = make_uint2(d[k], pos[k] & ~(L - 1));
        q0 += static_cast<uint32_t>(__popcll(cm[k]));
      }
    }
  };
  // Flusher: writes the lines of queue buffer qb, UF stores per round.
  auto flush = [&](uint32_t qb) {
    const uint32_t nq = qcount[qb];
    const uint32_t j = lane % L;
    for (uint32_t q0 = (w - NLW) * LPW * UF; q0 < nq; q0 += NFW * LPW * UF) {
      uint2 e[UF];
      bool live[UF];
#pragma unroll
      for (uint32_t u = 0; u < UF; ++u) {
        const uint32_t q = q0 + u * LPW + lane / L;
        live[u] = q < nq;
        e[u] = live[u] ? queue[qb * QL + q] : make_uint2(0, 0);
      }
      T t[UF];
      uint32_t lo[UF];
#pragma unroll
      for (uint32_t u = 0; u < UF; ++u) {
        lo[u] = ring_lo_prev[e[u].x];
        t[u] = ring[static_cast<size_t>(e[u].x) * RL + ((e[u].y + j) & (RL - 1))];
      }
#pragma unroll
      for (uint32_t u = 0; u < UF; ++u)
        if (live[u] && e[u].y + j >= lo[u])
          store_tuple_nt<K>(as_global(org(e[u].x) + e[u].y + j), t[u].key, t[u].value);
    }
  };
  // Between steps: the limit of the next step, the low-water mark after a
  // direct store, and the flushed queue buffer emptied for reuse.
  auto commit = [&](uint32_t qb_flushed) {
    for (uint32_t i = tid; i < F; i += BLOCK) {
      const uint64_t v = fl[i];
      const uint32_t f = static_cast<uint32_t>(v);
      const uint32_t fp = fprev[i];
      const uint32_t lo = ring_lo[i];
      ring_lo_prev[i] = lo;
      if (f > static_cast<uint32_t>(v >> 32)) ring_lo[i] = f;
      reinterpret_cast<uint32_t*>(&fl[i])[1] = (fp & ~(L - 1)) + RL;
      fprev[i] = f;
    }
    if (tid == 0) qcount[qb_flushed] = 0;
  };

  // The roles run separate loops with the same barriers (two per step,
  // three at the end), so the loaders' load waits are counted over loads
  // only.
  const uint64_t steps = (nvec + NLT - 1) / NLT;
  const uint64_t steps_run = (steps + D - 1) / D * D;  // whole rounds of D
  const uint32_t qlast = static_cast<uint32_t>(steps_run & 1);
  if (w < NLW) {
    const uint32_t lt = tid;  // loader lane
    if (steps) {
      // D register buffers, each step's loads issued D steps ahead; loads
      // past the last full vector re-read it, so every step issues the
      // same loads.
      VK kv[D], pv[D];
      auto load = [&](VK& kb, VK& pb, uint64_t v) {
        kb = __builtin_nontemporal_load(as_global(reinterpret_cast<const VK*>(keys)) + v);
        pb = __builtin_nontemporal_load(as_global(reinterpret_cast<const VK*>(pays)) + v);
      };
#pragma unroll
      for (uint32_t b = 0; b < D; ++b) {
        load(kv[b], pv[b], min(static_cast<uint64_t>(b) * NLT + lt, nvec - 1));
        // issued in buffer order: the loop's waits count on it
        __builtin_amdgcn_sched_barrier(0);
      }
      for (uint64_t s0 = 0; s0 < steps_run; s0 += D) {
#pragma unroll
        for (uint32_t b = 0; b < D; ++b) {
          const uint64_t s = s0 + b;
          place(kv[b], pv[b], s * NLT + lt < nvec ? (1u << V) - 1 : 0u, static_cast<uint32_t>(s & 1));
          // The buffer is reloaded only after its last use, so the loaded
          // registers stay in place across the loop's back edge (a copy of a
          // register with a load in flight would wait for that load).
          __builtin_amdgcn_sched_barrier(0);
          load(kv[b], pv[b], min((s + D) * NLT + lt, nvec - 1));
          __syncthreads();
          commit(static_cast<uint32_t>(s & 1) ^ 1u);
          __syncthreads();
        }
      }
    }
    // The last len % V tuples (loader lane 0).
    VK key, pay;
    uint32_t act = 0;
#pragma unroll
    for (uint32_t k = 0; k < V; ++k) {
      const uint64_t e = nvec * V + k;
      key[k] = 0;
      pay[k] = 0;
      if (lt == 0 && e < len) {
        key[k] = keys[e];
        pay[k] = pays[e];
        act |= 1u << k;
      }
    }
    place(key, pay, act, qlast);
    __syncthreads();
    commit(qlast ^ 1u);
    __syncthreads();
    __syncthreads();
  } else {
    for (uint64_t s = 0; s < steps_run; ++s) {
      flush(static_cast<uint32_t>(s & 1) ^ 1u);
      __syncthreads();
      commit(static_cast<uint32_t>(s & 1) ^ 1u);
      __syncthreads();
    }
    flush(qlast ^ 1u);
    __syncthreads();
    commit(qlast ^ 1u);
    __syncthreads();
    flush(qlast);
    __syncthreads();
  }
  // Write the partially filled last line of every partition.
  for (uint32_t idx = tid; idx < F * L; idx += BLOCK) {
    const uint32_t i = idx / L;
    const uint32_t f = static_cast<uint32_t>(fl[i]);
    const uint32_t pos = (f & ~(L - 1)) + (idx % L);
    if (pos < f && pos >= ring_lo[i]) {
      const T t = ring[static_cast<size_t>(i) * RL + (pos & (RL - 1))];
      store_tuple<K>(as_global(org(i) + pos), t.key, t.value);
    }
  }
}

template <typename K, int BLOCK, bool SPILL>
__global__ __launch_bounds__(BLOCK) void scatter_swwc_kernel(ScatterParams p) {
  scatter_swwc_body<K, BLOCK, SPILL>(p, blockIdx.x, gridDim.x);
}

// Many relations in one launch: block b scatters chunk b - prefix[j] of
// relation j (batched joins; every relation has the same fanout).
template <typename K, int BLOCK>
__global__ __launch_bounds__(BLOCK) void scatter_swwc_batched_kernel(const ScatterParams* __restrict__ jobs,
                                                                     const uint32_t* __restrict__ prefix,
                                                                     uint32_t njobs) {
  uint32_t lo = 0, hi = njobs;  // the last j with prefix[j] <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) / 2;
    if (scalar_load(prefix + m) <= blockIdx.x) lo = m; else hi = m;
  }
  // (read-only, unaliased descriptors: the compiler treats the pointers
  // loaded from them as global)
  const ScatterParams p = jobs[lo];
  const uint32_t c0 = scalar_load(prefix + lo);
  scatter_swwc_body<K, BLOCK, false>(p, blockIdx.x - c0, scalar_load(prefix + lo + 1) - c0);
}

// Queue entries per step buffer: a step completes at most one line per
// placed tuple and at most one line per ring line.
template <typename K>
uint32_t swwc_queue_len(uint32_t fanout, uint32_t ring_lines, uint32_t block) {
  const uint64_t tile = static_cast<uint64_t>(block / 2) * (16 / sizeof(K)) + 1;
  const uint64_t ring = static_cast<uint64_t>(fanout) * ring_lines;
  return static_cast<uint32_t>(tile < ring ? tile : ring);
}

uint32_t swwc_bytes(uint32_t fanout, uint32_t ring_lines, uint32_t queue_len) {
  const uint64_t bytes = static_cast<uint64_t>(fanout) * ring_lines * kAlignBytes  // rings
                         + static_cast<uint64_t>(queue_len) * 2 * 8                // line queues
                         + static_cast<uint64_t>(fanout) * 4 * 5;                  // per-partition words
  return bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(bytes);
}

// Launches the buffered scatter with the deepest ring (up to 256 lines per
// partition, so that low fanouts buffer many tuples per partition) whose LDS
// fits next to the kernel's static LDS; returns hipErrorNotSupported
// without launching when even one line does not fit.
template <typename K, int BLOCK, bool SPILL>
hipError_t launch_swwc_t(ScatterParams p, uint32_t fanout, uint32_t chunks, uint32_t lds_total,
                         hipStream_t stream) {
  auto kern = scatter_swwc_kernel<K, BLOCK, SPILL>;
  const uint32_t budget = dynamic_lds_budget(reinterpret_cast<const void*>(kern), lds_total);
  uint32_t ring = 0;
  for (uint32_t r = 256; r >= 1; r /= 2) {
    if (swwc_bytes(fanout, r, swwc_queue_len<K>(fanout, r, BLOCK)) <= budget) {
      ring = r;
      break;
    }
  }
  if (ring == 0) return hipErrorNotSupported;
  p.ring_lines = ring;
  p.queue_len = swwc_queue_len<K>(fanout, ring, BLOCK);
  const uint32_t lds = swwc_bytes(fanout, ring, p.queue_len);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(chunks), dim3(BLOCK), lds, stream, p);
  return hipGetLastError();
}

// Launches the tile-sorting scatter with the largest tile (4096 or 2048
// tuples of 16 bytes, 8192 or 4096 of 8 bytes) whose LDS fits; returns
// hipErrorNotSupported without launching when none does.
template <typename K, uint32_t ITEMS, bool SPILL>
hipError_t launch_tile_items(const ScatterParams& p, uint32_t fanout, uint32_t chunks, uint32_t lds_total,
                             hipStream_t stream) {
  auto kern = scatter_tile_kernel<K, ITEMS, 1024, SPILL>;
  const uint32_t budget = dynamic_lds_budget(reinterpret_cast<const void*>(kern), lds_total);
  const uint32_t lds = tile_scatter_lds<K, ITEMS, 1024>(fanout);
  if (lds > budget) return hipErrorNotSupported;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(chunks), dim3(1024), lds, stream, p);
  return hipGetLastError();
}

template <typename K>
hipError_t launch_tile_t(const ScatterParams& p, uint32_t fanout, uint32_t chunks, uint32_t lds_total, bool spill,
                         hipStream_t stream) {
  constexpr uint32_t I = sizeof(K) == 8 ? 4 : 8;
  hipError_t e = spill ? launch_tile_items<K, I, true>(p, fanout, chunks, lds_total, stream)
                       : launch_tile_items<K, I, false>(p, fanout, chunks, lds_total, stream);
  if (e != hipErrorNotSupported) return e;
  return spill ? launch_tile_items<K, I / 2, true>(p, fanout, chunks, lds_total, stream)
               : launch_tile_items<K, I / 2, false>(p, fanout, chunks, lds_total, stream);
}

template <typename K>
hipError_t launch_partition_t(const PartitionLaunch& a, hipStream_t stream) {
  ScatterParams p{};
  p.keys = a.keys;
  p.pays = a.pays;
  p.n = a.n;
  p.chunk_len = canonical_chunk_len<K>(a.n, a.chunks);
  p.radix_bits = a.radix_bits;
  p.ignore_bits = a.ignore_bits;
  p.chunk_offsets = a.chunk_offsets;
  p.relation = a.relation;
  const uint32_t F = 1u << a.radix_bits;
  const bool spill = a.spill != nullptr && a.spill_from < F;
  p.spill_from = spill ? a.spill_from : F;
  p.spill_addr = reinterpret_cast<uintptr_t>(a.spill) - static_cast<uintptr_t>(a.spill_base * sizeof(Tuple<K>));
  if (a.buffered) {
    hipError_t e;
    if (spill)
      e = launch_swwc_t<K, 1024, true>(p, F, a.chunks, a.lds_bytes, stream);
    else if (a.block >= 1024)
      e = launch_swwc_t<K, 1024, false>(p, F, a.chunks, a.lds_bytes, stream);
    else if (a.block >= 512)
      e = launch_swwc_t<K, 512, false>(p, F, a.chunks, a.lds_bytes, stream);
    else
      e = launch_swwc_t<K, 256, false>(p, F, a.chunks, a.lds_bytes, stream);
    if (e != hipErrorNotSupported) return e;
    // no line ring fits: the tile-sorting scatter
    e = launch_tile_t<K>(p, F, a.chunks, a.lds_bytes, spill, stream);
    if (e != hipErrorNotSupported) return e;
  }
  // Direct scatter: per-partition cursors only (4 bytes per partition).
  const uint32_t block = a.block > 1024 ? 1024 : a.block;
  if (spill)
    hipLaunchKernelGGL((scatter_direct_kernel<K, true>), dim3(a.chunks), dim3(block), F * 4, stream, p);
  else
    hipLaunchKernelGGL((scatter_direct_kernel<K, false>), dim3(a.chunks), dim3(block), F * 4, stream, p);
  return hipGetLastError();
}

}  // namespace

uint32_t swwc_lds_bytes(uint32_t fanout, uint32_t tuple_bytes, uint32_t ring_lines) {
  const uint32_t q = tuple_bytes == 8 ? swwc_queue_len<int32_t>(fanout, ring_lines, 1024)
                                      : swwc_queue_len<int64_t>(fanout, ring_lines, 1024);
  return swwc_bytes(fanout, ring_lines, q);
}

hipError_t launch_partition(const PartitionLaunch& a, hipStream_t stream) {
  if (a.type == KeyType::I32) return launch_partition_t<int32_t>(a, stream);
  return launch_partition_t<int64_t>(a, stream);
}

namespace {

template <typename K>
const void* batched_scatter_kernel(uint32_t block) {
  return block == 512 ? reinterpret_cast<const void*>(scatter_swwc_batched_kernel<K, 512>)
                      : reinterpret_cast<const void*>(scatter_swwc_batched_kernel<K, 1024>);
}

// Ring of the batched scatter for this fanout and block: the deepest that
// fits LDS, at most ring_cap lines when ring_cap is set (0: none fits).
template <typename K>
uint32_t batched_ring(uint32_t fanout, uint32_t lds_total, uint32_t block, uint32_t ring_cap) {
  const uint32_t budget = dynamic_lds_budget(batched_scatter_kernel<K>(block), lds_total);
  const uint32_t top = ring_cap ? (ring_cap < 256 ? ring_cap : 256) : 256;
  for (uint32_t r = top; r >= 1; r /= 2)
    if (swwc_bytes(fanout, r, swwc_queue_len<K>(fanout, r, block)) <= budget) return r;
  return 0;
}

template <typename K>
hipError_t scatter_params_batched_t(const void* keys, const void* pays, uint64_t n, uint32_t chunks,
                                    uint32_t radix_bits, const uint64_t* chunk_offsets, void* relation,
                                    uint32_t lds_total, uint32_t block, uint32_t ring_cap, ScatterParams* out) {
  const uint32_t F = 1u << radix_bits;
  const uint32_t ring = batched_ring<K>(F, lds_total, block, ring_cap);
  if (ring == 0) return hipErrorNotSupported;
  ScatterParams p{};
  p.keys = keys;
  p.pays = pays;
  p.n = n;
  p.chunk_len = canonical_chunk_len<K>(n, chunks);
  p.radix_bits = radix_bits;
  p.ignore_bits = 0;
  p.chunk_offsets = chunk_offsets;
  p.relation = relation;
  p.ring_lines = ring;
  p.queue_len = swwc_queue_len<K>(F, ring, block);
  p.spill_from = F;
  *out = p;
  return hipSuccess;
}

template <typename K>
hipError_t launch_partition_batched_t(const ScatterParams* jobs, const uint32_t* prefix, uint32_t njobs,
                                      uint32_t blocks, uint32_t radix_bits, uint32_t lds_total, uint32_t block,
                                      uint32_t ring_cap, hipStream_t stream) {
  const uint32_t F = 1u << radix_bits;
  const uint32_t ring = batched_ring<K>(F, lds_total, block, ring_cap);
  if (ring == 0) return hipErrorNotSupported;
  const uint32_t lds = swwc_bytes(F, ring, swwc_queue_len<K>(F, ring, block));
  (void)hipFuncSetAttribute(batched_scatter_kernel<K>(block), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  if (block == 512)
    hipLaunchKernelGGL((scatter_swwc_batched_kernel<K, 512>), dim3(blocks), dim3(512), lds, stream, jobs, prefix, njobs);
  else
    hipLaunchKernelGGL((scatter_swwc_batched_kernel<K, 1024>), dim3(blocks), dim3(1024), lds, stream, jobs, prefix,
                       njobs);
  return hipGetLastError();
}

}  // namespace

hipError_t scatter_params_batched(KeyType t, const void* keys, const void* pays, uint64_t n, uint32_t chunks,
                                  uint32_t radix_bits, const uint64_t* chunk_offsets, void* relation,
                                  uint32_t lds_total, uint32_t block, uint32_t ring_cap, ScatterParams* out) {
  if (block != 512 && block != 1024) return hipErrorInvalidValue;
  if (t == KeyType::I32)
    return scatter_params_batched_t<int32_t>(keys, pays, n, chunks, radix_bits, chunk_offsets, relation, lds_total,
                                             block, ring_cap, out);
  return scatter_params_batched_t<int64_t>(keys, pays, n, chunks, radix_bits, chunk_offsets, relation, lds_total,
                                           block, ring_cap, out);
}

hipError_t launch_partition_batched(KeyType t, const ScatterParams* jobs, const uint32_t* prefix, uint32_t njobs,
                                    uint32_t blocks, uint32_t radix_bits, uint32_t lds_total, uint32_t block,
                                    uint32_t ring_cap, hipStream_t stream) {
  if (blocks == 0) return hipSuccess;
  if (block != 512 && block != 1024) return hipErrorInvalidValue;
  if (t == KeyType::I32)
    return launch_partition_batched_t<int32_t>(jobs, prefix, njobs, blocks, radix_bits, lds_total, block, ring_cap,
                                               stream);
  return launch_partition_batched_t<int64_t>(jobs, prefix, njobs, blocks, radix_bits, lds_total, block, ring_cap,
                                             stream);
}

}  // namespace fij
