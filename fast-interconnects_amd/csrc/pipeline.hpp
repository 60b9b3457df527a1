// Device-visible state of the batched radix-join pipeline (pipeline.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "internal.hpp"

namespace fij {

// One relation's buffers.  rel1 is the pass-1 partitioned relation in the
// contiguous layout (each partition one range, 128 bytes of padding in front
// of it): chunk c writes partition p from local1[c][p] on, partition p starts
// at off1[p].  rel2 holds the pass-2 output: the concatenation of all pass-1
// partitions, each cut into tiles of tile2 tuples that are sorted by the
// pass-2 digit.
struct PipelineSide {
  const void* keys;           // input columns
  const void* pays;
  uint64_t n;
  uint64_t* off1;             // [F1] partition starts in rel1
  uint64_t* local1;           // [G1][F1] per-chunk write offsets
  uint32_t* counts1;          // [F1][G1] per-chunk histograms
  uint64_t* scan1;            // look-back scan state of the offsets scan
  uint64_t scan1_len;
  void* rel1;                 // Tuple[padded_len1]
  uint64_t padded_len1;       // n + F1 * pad
  uint32_t* plen;             // [F1]
  uint64_t* pstart;           // [F1 + 1] start of partition p in rel2
  uint32_t* tile_start;       // [F1 + 1] first pass-2 tile of partition p
  void* rel2;                 // Tuple[n]
  uint32_t* meta;             // [tiles + 1][F2] per tile and pass-2 digit: run start | run end << 16
};

// Join work counters, each on its own 128-byte line: one per XCD (work
// groups are dispatched to the 8 XCDs round-robin), then the ticket counter
// over the deferred-item flags, a word that is non-zero when any item was
// deferred, the number of block-table work units made of them (stored from
// the front of `units`), the ticket counter of the wave-table units, that of
// the block-table units, and the number of wave-table units (stored from the
// back of `units`); then the HBM-table counters: slots reserved in the HBM
// table arena (a 64-bit word), items given an HBM table, their build and
// probe work units, and the tickets over those.
constexpr uint32_t kJoinCounters = 8;
constexpr uint32_t kJoinCounterStride = 32;
constexpr uint32_t kOverflowTicket = kJoinCounters * kJoinCounterStride;
constexpr uint32_t kOverflowAny = (kJoinCounters + 1) * kJoinCounterStride;
constexpr uint32_t kUnitCount = (kJoinCounters + 2) * kJoinCounterStride;
constexpr uint32_t kUnitTicket = (kJoinCounters + 3) * kJoinCounterStride;
constexpr uint32_t kBlockTicket = (kJoinCounters + 4) * kJoinCounterStride;
constexpr uint32_t kWaveUnitCount = (kJoinCounters + 5) * kJoinCounterStride;
constexpr uint32_t kHbmSlots = (kJoinCounters + 6) * kJoinCounterStride;
constexpr uint32_t kBigCount = (kJoinCounters + 7) * kJoinCounterStride;
constexpr uint32_t kHbmBuildUnits = (kJoinCounters + 8) * kJoinCounterStride;
constexpr uint32_t kHbmProbeUnits = (kJoinCounters + 9) * kJoinCounterStride;
constexpr uint32_t kHbmBuildTicket = (kJoinCounters + 10) * kJoinCounterStride;
constexpr uint32_t kHbmProbeTicket = (kJoinCounters + 11) * kJoinCounterStride;
constexpr uint32_t kJoinCounterWords = (kJoinCounters + 12) * kJoinCounterStride;

struct PipelineDevice {
  PipelineSide build;
  PipelineSide probe;
  uint32_t F1, F2, G1;
  uint32_t bits1, bits2;
  uint32_t pad;
  uint32_t tile2;
  uint32_t sides;             // bit 0: build, bit 1: probe -- the sides the plan and pass-2 kernels process
  uint32_t* work_counter;     // [kJoinCounterWords]
  uint8_t* overflow;          // [F1 * F2] 1 = item deferred to the general join (zero between joins)
  uint4* units;               // [unit_cap] {item, first probe run, end probe run, 1 = wave table}
  uint32_t unit_cap;          // at least the items plus one per 65,536 probe tuples
  JoinAggregate* aggregates;  // [agg_len]
  uint32_t agg_len;
  // Many joins in one plan (batched joins): pass-1 partition p belongs to
  // join p >> job_shift, whose count and sums are added to
  // job_results[job_slot[join]] instead of the per-block aggregates.
  // job_results == nullptr: one join.
  JoinAggregate* job_results;
  const uint32_t* job_slot;
  uint32_t job_shift;
  // HBM hash tables of deferred items whose build side exceeds a block's
  // LDS table.  hbm_cap == 0: none; such items refill wave LDS tables.
  void* hbm_table;            // Tuple[hbm_cap], every word all ones between joins
  uint32_t hbm_cap;           // slots (power of two)
  uint32_t big_cap;           // entries of `big`
  uint4* big;                 // [big_cap] {item, first slot, log2 slots, 0}
  uint4* hbm_units;           // [hbm_unit_cap] {big index, first run, end run, 0}: build units from the front,
                              // probe units from the back
  uint32_t hbm_unit_cap;
  // single-pass plans: deferred items joined from 128 KiB block tables (one
  // block per CU) instead of 64 KiB ones (build partitions beyond the latter)
  uint32_t big_block_tables;
};

// One join side of a batched plan: its pass-1 region and where its
// partitions start in the batched plan's arrays (plan_batched_kernel).
struct BatchSide {
  uint64_t rel1_end;    // one past the side's pass-1 region in rel1 (tuples)
  uint64_t rel2_base;   // first tuple of the side's pass-2 output in rel2
  uint32_t part_base;   // first pass-1 partition (join * F1)
  uint32_t reserved;
};

// Allocates one side's buffers (capacity n tuples, F1 x F2 partitions, G1
// chunks; rel2 and meta only for two passes, meta with `tiles` + 1 rows) and
// records every allocation in `allocs`.  rel1 holds n + F1 * pad_tuples(t)
// tuples: the plan using it must set its pad to pad_tuples(t).
hipError_t alloc_pipeline_side(PipelineSide& s, KeyType t, uint64_t n, uint32_t F1, uint32_t F2, uint32_t G1,
                               bool two_pass, uint32_t tiles, std::vector<void*>& allocs);
// Pass 1 of one side: contiguous histogram + offsets scan, then the scatter.
hipError_t enqueue_side_prefix_sum(const PipelineDevice& d, PipelineSide& s, KeyType t, uint32_t block,
                                   hipStream_t stream);
hipError_t enqueue_side_scatter(const PipelineDevice& d, PipelineSide& s, KeyType t, uint32_t block,
                                uint32_t lds_total, hipStream_t stream);

uint32_t pipeline_tile2(KeyType t);
// Build tuples a deferred item may have to be joined from one block-shared
// LDS table (3/4 of its slots): the 64 KiB table, or the 128 KiB one of
// single-pass plans with big_block_tables.
uint32_t pipeline_block_table_tuples(KeyType t, bool big = false);

// Single-pass plans (fij_join_plan_bits, single-pass batched classes): 9
// bits, the widest pass-1 fanout whose line rings fit LDS.
constexpr uint32_t kOnePassBits = 9;
// Whether a build side's single-pass partitions hold on average at most 2/3
// of a (big) block table (half again of headroom for uneven ones).
inline bool block_tables_hold(KeyType t, uint64_t build_len, uint64_t partitions, bool big) {
  return build_len * 3 <= static_cast<uint64_t>(pipeline_block_table_tuples(t, big)) * partitions * 2;
}
// Single pass of kOnePassBits: with 128 KiB tables where 64 KiB ones do not
// hold the partitions (the big ones run one block per CU and are slower for
// build sides the small ones hold).
inline bool one_pass_fits(KeyType t, uint64_t build_len) {
  return block_tables_hold(t, build_len, 1ull << kOnePassBits, true);
}
// Pass-2 bits of a two-pass plan of 9 + b2 bits whose items hold about a
// third of a 64 KiB block table's build tuples (1,024 int64, 2,048 int32):
// for probe sides far longer than the build side, whose items go to the
// block join anyway, fewer pass-2 digits give longer runs.
inline uint32_t probe_heavy_bits2(KeyType t, uint64_t build_len) {
  const uint64_t per_item = pipeline_block_table_tuples(t, false) / 3;
  const uint64_t items = (build_len >> kOnePassBits) / per_item;
  uint32_t b2 = 1;
  while (b2 < 12 && (1ull << b2) < items) ++b2;
  return b2;
}
// big_block_tables of a single-pass plan of `partitions` partitions.
inline uint32_t big_block_tables_for(KeyType t, uint64_t build_len, uint64_t partitions) {
  return block_tables_hold(t, build_len, partitions, false) ? 0u : 1u;
}
size_t pipeline_sort2_lds(KeyType t, uint32_t F2);
hipError_t launch_pipeline_plan(const PipelineDevice& d, hipStream_t stream);
// Partition lengths, pass-2 starts and tile maps of every join side of a
// batched plan (sides[2 * join + side]; the tiles of all joins of a side
// form one range, side_tiles[2 * joins] is scratch); resets the join
// counters and the joins' results.  pstart / tile_start [F1] hold the
// totals.
hipError_t launch_pipeline_plan_batched(const PipelineDevice& d, const BatchSide* sides, uint32_t joins,
                                        uint32_t* side_tiles, uint64_t rel2_total_b, uint64_t rel2_total_p,
                                        hipStream_t stream);
// Pass 2 of both relations in one persistent launch of `grid` blocks.
hipError_t launch_pipeline_sort2(const PipelineDevice& d, KeyType t, uint32_t grid, hipStream_t stream);
// dst = the sum of the n per-block aggregates src (one block).
hipError_t launch_sum_aggregates(const JoinAggregate* src, uint32_t n, JoinAggregate* dst, hipStream_t stream);
// The register-resident join of every work item (deferred = false), then
// the join of the items it deferred (deferred = true).
hipError_t launch_pipeline_join(const PipelineDevice& d, KeyType t, bool large_tables, bool deferred, uint32_t grid,
                                hipStream_t stream);

}  // namespace fij
