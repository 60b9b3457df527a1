// Internal launch interface between the C-ABI layer (capi.cpp) and the kernel
// translation units.  Every launcher is asynchronous on `stream` and returns
// the hipError_t of its launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fij.h"

namespace fij {

// Which element type a launch works on (i32 keys + i32 payloads, or i64/i64).
enum class KeyType : uint32_t { I32 = 0, I64 = 1 };

inline uint32_t key_bytes(KeyType t) { return t == KeyType::I32 ? 4u : 8u; }
inline uint32_t tuple_bytes(KeyType t) { return 2u * key_bytes(t); }
inline uint32_t pad_tuples(KeyType t) { return 128u / tuple_bytes(t); }

// Status bits a kernel may raise in the library's device status word.
enum DeviceStatus : uint32_t {
  kStatusOk = 0,
  kStatusTransformLength = 1u << 0,   // transform length differs from caller's
  kStatusHashTableFull = 1u << 1,     // a build partition exceeded its table
  kStatusPartitionOverflow = 1u << 2, // offsets inconsistent with the input
};

// Decoupled look-back scan state: one ticket word plus one word per tile.
constexpr uint32_t kScanBlock = 1024;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;
constexpr uint64_t scan_state_words(uint64_t n) { return 1 + (n + kScanTile - 1) / kScanTile; }

// ---- prefix sums (hist.hip) ----------------------------------------------
// Chunked: one block per chunk; writes offsets[chunk][F] directly.
hipError_t launch_hist_chunked(KeyType t, const void* keys, uint64_t n, uint32_t chunks,
                               uint32_t radix_bits, uint32_t ignore_bits, uint64_t* offsets,
                               uint32_t block, hipStream_t stream);
// Contiguous, step 1: per-chunk counts in partition-major order counts[F][chunks];
// optionally copies keys/payloads (prefix_sum_and_copy_with_payload); clears
// `scan_state` for the following scan.
hipError_t launch_hist_counts(KeyType t, const void* keys, const void* pays, void* dst_keys,
                              void* dst_pays, uint64_t n, uint32_t chunks, uint32_t radix_bits,
                              uint32_t ignore_bits, uint32_t* counts, uint64_t* scan_state,
                              uint64_t scan_state_len, uint32_t block, hipStream_t stream);
// Contiguous transform, step 1: gathers partition `pid` of a (chunked or
// contiguous) partitioned relation into columnar dst arrays, and counts.
hipError_t launch_transform_counts(KeyType t, const void* src_rel, const uint64_t* src_offsets,
                                   uint32_t src_chunks, uint32_t src_radix_bits,
                                   uint64_t src_padded_len, uint32_t pid, uint64_t n_expected,
                                   void* dst_keys, void* dst_pays, uint32_t chunks,
                                   uint32_t radix_bits, uint32_t ignore_bits, uint32_t* counts,
                                   uint64_t* scan_state, uint64_t scan_state_len,
                                   uint32_t* status, uint32_t block, hipStream_t stream);
// Contiguous, step 2: decoupled look-back exclusive scan of counts[F][chunks],
// producing local_offsets[chunks][F] (+ (p+1)*pad) and offsets[F].
hipError_t launch_offsets_scan(const uint32_t* counts, uint32_t fanout, uint32_t chunks,
                               uint32_t pad, uint64_t* local_offsets, uint64_t* offsets,
                               uint64_t* scan_state, hipStream_t stream);

// Allocation-free contiguous prefix sum: at most kNoAllocCounters per-chunk
// counters (fanout x chunks).  noalloc_scratch returns the current device's
// copy of the library's static scratch array: counts[kNoAllocCounters] and
// the scan state behind them (scan_state_words(kNoAllocCounters) words).
constexpr uint64_t kNoAllocCounters = FIJ_NOALLOC_MAX_COUNTERS;
hipError_t noalloc_scratch(uint32_t** counts, uint64_t** scan_state);

// Histogram launch parameters (one relation, contiguous counts[F][chunks] or
// chunked offsets[chunks][F]).
struct HistParams {
  const void* keys;
  const void* pays;
  void* dst_keys;
  void* dst_pays;
  uint64_t n;
  uint64_t chunk_len;
  uint32_t radix_bits;
  uint32_t ignore_bits;
  uint32_t pad;
  uint32_t subhists;
  uint64_t* offsets;   // chunked: [G][F]
  uint32_t* counts;    // contiguous: [F][G]
  uint64_t* scan_state;
  uint64_t scan_state_len;
};

// One relation of a batched offsets launch: counts[F][chunks] in, write
// offsets local[chunks][F] (relative to the relation's region) and
// partition starts off1[F] (in the plan's pass-1 array) out.
struct BatchScanJob {
  const uint32_t* counts;
  uint64_t* local;
  uint64_t* off1;
  uint64_t region;
  uint32_t fanout, chunks, pad;
};
// Contiguous histogram parameters of one relation of a batched launch.
HistParams hist_params(KeyType t, const void* keys, uint64_t n, uint32_t chunks, uint32_t radix_bits,
                       uint32_t block, uint32_t* counts);
// Histograms of njobs relations in one launch: block b counts chunk
// b - prefix[j] of jobs[j] (prefix[njobs] = blocks, device memory).
hipError_t launch_hist_batched(KeyType t, const HistParams* jobs, const uint32_t* prefix, uint32_t njobs,
                               uint32_t blocks, uint32_t radix_bits, uint32_t block, hipStream_t stream);
// Offsets of njobs relations in one launch (one block each).
hipError_t launch_offsets_batched(const BatchScanJob* jobs, uint32_t njobs, hipStream_t stream);

// ---- partitioning (partition.hip) -----------------------------------------
struct PartitionLaunch {
  KeyType type;
  const void* keys;
  const void* pays;
  uint64_t n;
  uint32_t chunks;            // == grid
  uint32_t radix_bits;
  uint32_t ignore_bits;
  const uint64_t* chunk_offsets;  // [chunks][F] absolute tuple offsets
  void* relation;             // Tuple array
  uint32_t block;             // requested threads per block
  uint32_t lds_bytes;         // LDS the scatter may use per block
  bool buffered;              // software write-combining vs. direct scatter
  // Optional spill region: partitions spill_from.. are written to
  // `spill` instead of `relation`, spill[0] standing for tuple index
  // spill_base of the relation (so the relation's first spill_base tuples
  // live in `relation`, the rest in `spill`).  spill == nullptr: none.
  void* spill = nullptr;
  uint64_t spill_base = 0;
  uint32_t spill_from = 0;
};
hipError_t launch_partition(const PartitionLaunch& a, hipStream_t stream);
// Buffered-scatter parameters of one relation of a batched launch.
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
  uint32_t queue_len;    // line queue entries per step buffer
  uintptr_t spill_addr;  // address tuple index 0 would have in the spill region
  uint32_t spill_from;   // first partition written to the spill region
};

// Scatter descriptor of one relation of a batched launch: the line ring is
// the deepest that fits LDS for this block shape (512 or 1024 threads), at
// most ring_cap lines per partition when ring_cap is set.
hipError_t scatter_params_batched(KeyType t, const void* keys, const void* pays, uint64_t n, uint32_t chunks,
                                  uint32_t radix_bits, const uint64_t* chunk_offsets, void* relation,
                                  uint32_t lds_total, uint32_t block, uint32_t ring_cap, ScatterParams* out);
// Scatters of njobs relations of one fanout in one launch: block b scatters
// chunk b - prefix[j] of jobs[j] (prefix[njobs] = blocks, device memory);
// block and ring_cap as given to scatter_params_batched for every job.
hipError_t launch_partition_batched(KeyType t, const ScatterParams* jobs, const uint32_t* prefix, uint32_t njobs,
                                    uint32_t blocks, uint32_t radix_bits, uint32_t lds_total, uint32_t block,
                                    uint32_t ring_cap, hipStream_t stream);
// Pass 1 on the CPU (host memory in and out): the contiguous layout of the
// GPU's pass 1; offsets[F] receives the partition starts.
void cpu_partition_contiguous(KeyType t, const void* keys, const void* pays, uint64_t n, uint32_t bits,
                              uint32_t ignore, uint32_t threads, void* out, uint64_t* offsets);
// The GPU prefix sums' offsets on the host: chunked offsets[chunks][F], or
// contiguous offsets[F] and local[chunks][F].
void cpu_prefix_sum(KeyType t, const void* keys, uint64_t n, uint32_t chunks, uint64_t chunk_len, uint32_t bits,
                    uint32_t ignore, bool contiguous, uint32_t threads, uint64_t* offsets, uint64_t* local);
// LDS the buffered scatter needs for this fanout / tuple size / ring depth.
uint32_t swwc_lds_bytes(uint32_t fanout, uint32_t tuple_bytes, uint32_t ring_lines);

// ---- radix join (join.hip) --------------------------------------------------
struct JoinAggregate {
  uint64_t count;
  uint64_t sum_build;
  uint64_t sum_probe;
};
enum class JoinScheme : uint32_t { Perfect = 0, LinearProbing = 1, BucketChaining = 2 };
struct JoinLaunch {
  KeyType type;
  JoinScheme scheme;
  const void* build_rel;
  const uint64_t* build_offsets;
  uint64_t build_len;         // padded length of the build relation
  const void* probe_rel;
  const uint64_t* probe_offsets;
  uint64_t probe_len;         // padded length of the probe relation
  uint32_t radix_bits;
  uint32_t ignore_bits;       // table index = key >> ignore_bits (perfect)
  int64_t* result_set;        // [grid * block] per-thread SUM(probe.value), +=; may be null
  JoinAggregate* aggregates;  // [grid] per-block {count, sums}, +=; may be null
  uint32_t* task_assignments; // [grid + 1]
  uint32_t grid;
  uint32_t block;
  uint32_t lds_bytes;
  uint32_t* status;
};
hipError_t launch_assign_tasks(const uint64_t* probe_offsets, uint64_t probe_len, uint32_t pad,
                               uint32_t radix_bits, uint32_t grid, uint32_t* task_assignments,
                               hipStream_t stream);
hipError_t launch_join(const JoinLaunch& a, hipStream_t stream);

// ---- no-partitioning join (np_join.hip) -------------------------------------
hipError_t launch_np_table_init(KeyType t, void* table, uint64_t entries, bool null_values,
                                hipStream_t stream);
hipError_t launch_np_build(KeyType t, JoinScheme s, bool selective, void* table, uint64_t entries,
                           const void* keys, const void* pays, uint64_t n, uint32_t grid,
                           uint32_t block, hipStream_t stream);
hipError_t launch_np_probe(KeyType t, JoinScheme s, const void* table, uint64_t entries,
                           const void* keys, const void* pays, uint64_t n, uint64_t* result_set,
                           JoinAggregate* aggregates, uint32_t grid, uint32_t block,
                           hipStream_t stream);

// Probe of a resident linear-probing table of 2^lg entries; block b writes its
// aggregate to out[b] (grid blocks of 256 threads).
hipError_t launch_np_probe_resident(KeyType t, const void* table, uint32_t lg, const void* keys, const void* pays,
                                    uint64_t n, JoinAggregate* out, uint32_t grid, hipStream_t stream);

// ---- data generation (datagen.hip) -----------------------------------------
hipError_t launch_gen_perm_keys(KeyType t, void* out, uint64_t n, uint64_t seed, int64_t base,
                                hipStream_t stream);
hipError_t launch_gen_uniform(KeyType t, void* out, uint64_t n, uint64_t seed, uint64_t stream_id,
                              int64_t lo, uint64_t range, hipStream_t stream);
hipError_t launch_gen_iota(KeyType t, void* out, uint64_t n, int64_t base, hipStream_t stream);
hipError_t launch_gen_zipf(KeyType t, void* out, uint64_t n, uint64_t num_ranks, double exponent, uint64_t seed,
                           uint64_t stream_id, uint64_t perm_seed, int64_t base, hipStream_t stream);
hipError_t launch_fill_u64(uint64_t* out, uint64_t n, uint64_t value, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* out, uint64_t n, uint32_t value, hipStream_t stream);
hipError_t launch_copy(void* dst, const void* src, uint64_t words16, hipStream_t stream);
hipError_t launch_line_scatter(void* dst, const void* src, uint64_t words16, uint32_t partitions, uint32_t chunks,
                               hipStream_t stream);
hipError_t launch_run_gather(const void* src, uint64_t words16, uint32_t part_words, uint32_t f2, uint32_t run_words,
                             uint32_t shift, uint32_t* out, hipStream_t stream, uint32_t grid);

}  // namespace fij
