// Radix scatter kernels.
//
// Semantics follow the reference's chunked scatters
// (sql-ops/cudautils/radix_partition.cu:472-525 NC, :688-920 SSWWCv2): each
// block owns one input chunk and writes its tuples, in row layout, into the
// partition ranges that start at chunk_offsets[chunk][p].  Order within a
// partition is not defined (as in the reference).  Padding in front of each
// partition is never written.
//
// The buffered variant is designed for wave64 / 160 KB LDS: instead of the
// reference's per-buffer spin locks it processes the chunk in block-wide
// tiles with three barrier-separated phases, so no wave ever waits on
// another wave's progress inside a phase:
//   A  rank:   wave-aggregated LDS atomics give each tuple its position in
//              its partition's output sequence (ballot + mbcnt);
//   B  place:  the tuple goes into its partition's LDS ring of R aligned
//              128-byte lines; a tuple that completes a line enqueues it
//              (with the partition's low-water mark at that time).  A tuple
//              beyond the ring's capacity is stored directly;
//   C  flush + commit: waves write queued lines, 64/L lines per store
//              instruction, each a full aligned 128-byte line, while the
//              per-partition fill / flushed / low-water marks advance.
// Positions below a partition's low-water mark (front alignment slots and
// slots already stored directly) are never written from the ring.  The next
// tile's loads are in flight while a tile goes through the phases.
#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {

struct ScatterParams {
  const void* keys;
  const void* pays;
  uint64_t n;
  uint64_t chunk_len;
  uint32_t radix_bits;
  uint32_t ignore_bits;
  const uint64_t* chunk_offsets;
  void* relation;
  uint32_t ring_lines;
  uintptr_t spill_addr;  // address tuple index 0 would have in the spill region
  uint32_t spill_from;   // first partition written to the spill region
};

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
__device__ __forceinline__ void chunk_range(const ScatterParams& p, uint64_t& begin, uint64_t& len) {
  begin = min(p.chunk_len * blockIdx.x, p.n);
  len = (blockIdx.x + 1 == gridDim.x) ? p.n - begin : min(p.chunk_len, p.n - begin);
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
  chunk_range<K>(p, begin, len);
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

template <typename K, int BLOCK, int V, bool SPILL>
__global__ __launch_bounds__(BLOCK) void scatter_swwc_kernel(ScatterParams p) {
  using T = Tuple<K>;
  constexpr uint32_t L = kAlignBytes / sizeof(T);  // tuples per 128-byte line
  constexpr uint32_t TILE = BLOCK * V;
  constexpr uint32_t NW = BLOCK / kWave;
  constexpr uint32_t LPW = kWave / L;              // lines per wave store
  extern __shared__ uint64_t lds_u64[];
  __shared__ uint32_t qcount;

  const uint32_t F = 1u << p.radix_bits;
  const uint32_t R = p.ring_lines;
  const uint32_t RL = R * L;
  const uint64_t mask = radix_mask(p.radix_bits, p.ignore_bits);
  const uint32_t tid = threadIdx.x;

  T* ring = reinterpret_cast<T*>(lds_u64);            // [F][RL]
  uint32_t* base = reinterpret_cast<uint32_t*>(ring + static_cast<size_t>(F) * RL);
  uint32_t* fill = base + F;      // next position to assign (counts front alignment)
  uint32_t* flushed = fill + F;   // first position not yet written out (line aligned)
  uint32_t* ring_lo = flushed + F;  // positions below are never written from the ring
  uint32_t* cnt = ring_lo + F;    // per-tile arrivals
  uint32_t* qpart = cnt + F;      // [F * R] queued line: partition
  uint32_t* qline = qpart + static_cast<size_t>(F) * R;  // [F * R] queued line: line number
  uint32_t* qlo = qline + static_cast<size_t>(F) * R;    // [F * R] queued line: partition's low-water mark

  const uint64_t* off = p.chunk_offsets + static_cast<uint64_t>(blockIdx.x) * F;
  // The first partition starts at least one padding (= one line) into the
  // chunk's range, so cbase keeps every aligned line start non-negative.
  const uint64_t cbase = off[0] - L;
  T* rel = static_cast<T*>(p.relation) + cbase;
  for (uint32_t i = tid; i < F; i += BLOCK) {
    const uint32_t o = static_cast<uint32_t>(off[i] - cbase);
    const uint32_t m = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(rel + o) % kAlignBytes) / sizeof(T));
    base[i] = o - m;
    fill[i] = m;
    flushed[i] = 0;
    ring_lo[i] = m;
    cnt[i] = 0;
  }
  if (tid == 0) qcount = 0;
  __syncthreads();

  uint64_t begin, len;
  chunk_range<K>(p, begin, len);
  const K* keys = static_cast<const K*>(p.keys) + begin;
  const K* pays = static_cast<const K*>(p.pays) + begin;

  // Tile t0's keys and payloads; lanes past the chunk's end re-read its last
  // element, so every tile issues the same unconditional loads and the next
  // tile's loads stay in flight while this tile is ranked, placed and flushed.
  auto load_tile = [&](K (&key)[V], K (&pay)[V], uint64_t t0) {
#pragma unroll
    for (int k = 0; k < V; ++k) {
      const uint64_t e = min(t0 + static_cast<uint64_t>(k) * BLOCK + tid, len - 1);
      key[k] = __builtin_nontemporal_load(keys + e);
      pay[k] = __builtin_nontemporal_load(pays + e);
    }
  };
  auto process = [&](const K (&key)[V], const K (&pay)[V], uint64_t t0) {
    uint32_t d[V], r[V];
    bool act[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
      act[k] = t0 + static_cast<uint64_t>(k) * BLOCK + tid < len;
      d[k] = key_to_partition<K>(key[k], mask, p.ignore_bits);
      // The payloads are used on some paths only; using them here on every
      // path retires their loads in this tile's one wait, so the next
      // tile's loads (issued into the same registers later) never wait for
      // the loads in flight behind them.
      asm volatile("" ::"v"(pay[k]));
    }
    // A: rank (one LDS atomic per tuple: with 2^9 partitions the lanes of a
    // wave rarely share a digit, and a shared one only serialises that
    // address)
#pragma unroll
    for (int k = 0; k < V; ++k) r[k] = act[k] ? atomicAdd(&cnt[d[k]], 1u) : 0u;
    if (tid == 0) qcount = 0;
    __syncthreads();
    // B: place
#pragma unroll
    for (int k = 0; k < V; ++k) {
      bool complete = false;
      uint32_t pos = 0;
      if (act[k]) {
        const uint32_t f = fill[d[k]];
        const uint32_t fl = flushed[d[k]];
        pos = f + r[k];
        if (pos < fl + RL) {
          T t;
          t.key = key[k];
          t.value = pay[k];
          ring[static_cast<size_t>(d[k]) * RL + (pos & (RL - 1))] = t;
          complete = (pos & (L - 1)) == L - 1;
        } else {
          store_tuple<K>(part_dst<SPILL>(p, rel, cbase, d[k]) + base[d[k]] + pos, key[k], pay[k]);
        }
      }
      const uint64_t cm = ballot(complete);
      if (cm) {
        const int leader = __ffsll(static_cast<unsigned long long>(cm)) - 1;
        uint32_t qb = 0;
        if (static_cast<int>(lane_id()) == leader) qb = atomicAdd(&qcount, static_cast<uint32_t>(__popcll(cm)));
        qb = __builtin_amdgcn_readlane(qb, leader);
        if (complete) {
          const uint32_t q = qb + mbcnt(cm);
          qpart[q] = d[k];
          qline[q] = pos / L;
          qlo[q] = ring_lo[d[k]];
        }
      }
    }
    __syncthreads();
    // C: flush full lines (reads ring, the queue and base only) and commit
    // (writes fill, flushed, ring_lo and cnt only)
    {
      const uint32_t nq = qcount;
      const uint32_t w = tid / kWave;
      const uint32_t lane = lane_id();
      for (uint32_t q0 = w * LPW; q0 < nq; q0 += NW * LPW) {
        const uint32_t q = q0 + lane / L;
        if (q < nq) {
          const uint32_t pp = qpart[q];
          const uint32_t pos = qline[q] * L + (lane % L);
          if (pos >= qlo[q]) {
            const T t = ring[static_cast<size_t>(pp) * RL + (pos & (RL - 1))];
            store_tuple<K>(part_dst<SPILL>(p, rel, cbase, pp) + base[pp] + pos, t.key, t.value);
          }
        }
      }
    }
    for (uint32_t i = tid; i < F; i += BLOCK) {
      const uint32_t c = cnt[i];
      if (c) {
        const uint32_t nf = fill[i] + c;
        if (nf > flushed[i] + RL) ring_lo[i] = nf;
        flushed[i] = nf & ~(L - 1);
        fill[i] = nf;
        cnt[i] = 0;
      }
    }
    __syncthreads();
  };

  if (len) {
    // Three register buffers, loads issued two tiles ahead, unrolled by
    // three with no exit inside the loop: every path into the loop head has
    // consumed buffer c, so issuing its next loads never waits for the
    // loads in flight in a and b.  Loads past the end are clamped to the
    // last element.
    const uint64_t tiles = (len + TILE - 1) / TILE;
    K ka[V], pa[V], kb[V], pb[V], kc[V], pc[V];
    load_tile(ka, pa, 0);
    load_tile(kb, pb, TILE);
    uint64_t t = 0;
    for (; t + 3 <= tiles; t += 3) {
      load_tile(kc, pc, (t + 2) * TILE);
      process(ka, pa, t * TILE);
      load_tile(ka, pa, (t + 3) * TILE);
      process(kb, pb, (t + 1) * TILE);
      load_tile(kb, pb, (t + 4) * TILE);
      process(kc, pc, (t + 2) * TILE);
    }
    if (t < tiles) process(ka, pa, t * TILE);
    if (t + 1 < tiles) process(kb, pb, (t + 1) * TILE);
  }
  // Write the partially filled last line of every partition.
  for (uint32_t idx = tid; idx < F * L; idx += BLOCK) {
    const uint32_t i = idx / L;
    const uint32_t pos = flushed[i] + (idx % L);
    if (pos < fill[i] && pos >= ring_lo[i]) {
      const T t = ring[static_cast<size_t>(i) * RL + (pos & (RL - 1))];
      store_tuple<K>(part_dst<SPILL>(p, rel, cbase, i) + base[i] + pos, t.key, t.value);
    }
  }
}

// Launches the buffered scatter with the deepest ring (4, 2 or 1 lines per
// partition) whose LDS fits next to the kernel's static LDS; returns
// hipErrorNotSupported without launching when even one line does not fit.
template <typename K, int BLOCK, int V, bool SPILL>
hipError_t launch_swwc_t(ScatterParams p, uint32_t fanout, uint32_t chunks, uint32_t lds_total,
                         hipStream_t stream) {
  auto kern = scatter_swwc_kernel<K, BLOCK, V, SPILL>;
  const uint32_t budget = dynamic_lds_budget(reinterpret_cast<const void*>(kern), lds_total);
  uint32_t ring = 0;
  for (uint32_t r = 4; r >= 1; r /= 2) {
    if (swwc_lds_bytes(fanout, sizeof(Tuple<K>), r) <= budget) {
      ring = r;
      break;
    }
  }
  if (ring == 0) return hipErrorNotSupported;
  p.ring_lines = ring;
  const uint32_t lds = swwc_lds_bytes(fanout, sizeof(Tuple<K>), ring);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds));
  hipLaunchKernelGGL(kern, dim3(chunks), dim3(BLOCK), lds, stream, p);
  return hipGetLastError();
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
    constexpr int V = sizeof(K) == 4 ? 4 : 2;
    hipError_t e;
    if (spill)
      e = launch_swwc_t<K, 1024, V, true>(p, F, a.chunks, a.lds_bytes, stream);
    else if (a.block >= 1024)
      e = launch_swwc_t<K, 1024, V, false>(p, F, a.chunks, a.lds_bytes, stream);
    else if (a.block >= 512)
      e = launch_swwc_t<K, 512, V, false>(p, F, a.chunks, a.lds_bytes, stream);
    else
      e = launch_swwc_t<K, 256, V, false>(p, F, a.chunks, a.lds_bytes, stream);
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
  (void)tuple_bytes;
  const uint64_t bytes = static_cast<uint64_t>(fanout) * ring_lines * kAlignBytes  // rings
                         + static_cast<uint64_t>(fanout) * 4 * 5                  // per-partition words
                         + static_cast<uint64_t>(fanout) * ring_lines * 12;       // line queue
  return bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(bytes);
}

hipError_t launch_partition(const PartitionLaunch& a, hipStream_t stream) {
  if (a.type == KeyType::I32) return launch_partition_t<int32_t>(a, stream);
  return launch_partition_t<int64_t>(a, stream);
}

}  // namespace fij
