/*
 * fij.h -- C ABI of the MI355X (gfx950) radix-join library.
 *
 * This is the drop-in boundary for the reference's sql-ops GPU partitioner
 * and join operators (TU-Berlin-DIMA/fast-interconnects, sql-ops crate).
 * In the reference, the Rust operator structs launch CUDA kernels by name
 * from a fatbin (sql-ops/src/lib.rs:117-142).  Here every operator method is
 * one extern "C" function taking plain pointers, lengths and a hipStream_t
 * (passed as void*; NULL = default stream).  Functions are asynchronous on
 * that stream unless stated otherwise and return an fij_status; the message
 * of the last error on the calling thread is available from
 * fij_last_error().
 *
 * Layouts are the reference's: tuples are {key, value} rows of 8 (i32) or 16
 * (i64) bytes; every partition has 128 bytes of padding in front of it;
 * offsets are u64 tuple indices.  Device buffers are owned by the caller
 * except the partitioner's and join's internal state.
 */
#ifndef FIJ_H
#define FIJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sql-ops/src/error.rs:17-29 (ErrorKind). */
typedef enum fij_status {
  FIJ_OK = 0,
  FIJ_INVALID_ARGUMENT = 1,
  FIJ_INTEGER_OVERFLOW = 2,
  FIJ_DEVICE_ERROR = 3,   /* ErrorKind::CudaError */
  FIJ_RUNTIME_ERROR = 4,
  FIJ_UNSUPPORTED = 5,    /* combinations the reference leaves unimplemented!() */
  FIJ_OUT_OF_MEMORY = 6,
  FIJ_IO_ERROR = 7,       /* ErrorKind::IoError (data-store/src/error.rs) */
  FIJ_PARSE_ERROR = 8     /* ErrorKind::CsvError: a malformed relation file */
} fij_status;

/* Key/payload element type: Tuple<i32, i32> or Tuple<i64, i64>. */
typedef enum fij_type { FIJ_TYPE_I32 = 0, FIJ_TYPE_I64 = 1 } fij_type;

/* sql-ops/src/partition/gpu_radix_partition.rs:194-219 (GpuHistogramAlgorithm). */
typedef enum fij_histogram_algorithm {
  FIJ_HISTOGRAM_CHUNKED = 0,
  FIJ_HISTOGRAM_CONTIGUOUS = 1
} fij_histogram_algorithm;

/* sql-ops/src/partition/gpu_radix_partition.rs:241-293 (GpuRadixPartitionAlgorithm).
 * NC is the direct scatter; the four write-combining variants all select
 * this library's LDS line-buffered scatter, and where the fanout's line
 * rings exceed LDS (10 or more bits) its tile-sorting scatter, which sorts
 * each tile of the chunk by digit in LDS and writes runs; only where not even
 * a tile and its counters fit LDS does the direct scatter run. */
typedef enum fij_partition_algorithm {
  FIJ_PARTITION_NC = 0,
  FIJ_PARTITION_LASWWC = 1,
  FIJ_PARTITION_SSWWC_V2 = 2,
  FIJ_PARTITION_SSWWC_V2G = 3,
  FIJ_PARTITION_HSSWWC_V4 = 4
} fij_partition_algorithm;

/* sql-ops/src/partition.rs:55-60 (RadixPass). */
typedef enum fij_radix_pass { FIJ_PASS_FIRST = 0, FIJ_PASS_SECOND = 1, FIJ_PASS_THIRD = 2 } fij_radix_pass;

/* sql-ops/src/join/hashing_scheme.rs:19-43 (HashingScheme). */
typedef enum fij_hashing_scheme {
  FIJ_HASH_PERFECT = 0,
  FIJ_HASH_LINEAR_PROBING = 1,
  FIJ_HASH_BUCKET_CHAINING = 2
} fij_hashing_scheme;

/* sql-ops/src/partition.rs:62-120 (RadixBits): bits of each pass, a
 * negative value means the pass is not specified (None). */
typedef struct fij_radix_bits {
  int32_t pass_bits[3];
} fij_radix_bits;

#define FIJ_LEN_UNSET UINT64_MAX

/* sql-ops/src/partition/partitioned_relation.rs:63-70 (PartitionOffsets).
 * chunks = max_chunks for Chunked and 1 for Contiguous; local_offsets holds
 * max_chunks * fanout entries for Contiguous and is NULL for Chunked. */
typedef struct fij_partition_offsets {
  uint64_t *offsets;
  uint64_t *local_offsets;
  uint64_t data_len; /* FIJ_LEN_UNSET until a prefix sum ran */
  uint32_t chunks;
  uint32_t max_chunks;
  uint32_t radix_bits;
  uint32_t type; /* fij_type */
} fij_partition_offsets;

/* sql-ops/src/partition/partitioned_relation.rs:316-322 (PartitionedRelation). */
typedef struct fij_partitioned_relation {
  void *relation;        /* Tuple[relation_len] */
  uint64_t *offsets;     /* [chunks * fanout] */
  uint64_t relation_len; /* allocated tuples, padding included */
  uint64_t len;          /* tuples, padding excluded */
  uint32_t chunks;
  uint32_t radix_bits;
  uint32_t type;         /* fij_type */
  uint32_t reserved;
} fij_partitioned_relation;

/* Join aggregate extension: the reference reports only SUM(probe.value)
 * (sql-ops/src/join/cuda_radix_join.rs:17-30); this adds the match count and
 * SUM(build.value).  All sums are wrapping 64-bit. */
typedef struct fij_join_aggregate {
  uint64_t count;
  uint64_t sum_build;
  uint64_t sum_probe;
} fij_join_aggregate;

typedef struct fij_device_limits {
  uint32_t compute_units;
  uint32_t wave_size;
  uint32_t lds_bytes_per_block; /* usable LDS per block (opt-in maximum) */
  uint32_t max_threads_per_block;
  uint64_t global_mem_bytes;
  char arch_name[64];
} fij_device_limits;

/* ----------------------------------------------------------- library ---- */

/* Message of the last failed call on this thread ("" if none). */
const char *fij_last_error(void);
/* Library version string. */
const char *fij_version(void);
/* Device properties used for launch decisions (the reference reads
 * MultiprocessorCount / MaxSharedMemoryPerBlockOptin / WarpSize:
 * gpu_radix_partition.rs:570-571,620-621,981). Synchronous. */
fij_status fij_device_limits_get(int device, fij_device_limits *out);
/* Reads and optionally clears the current device's kernel status word
 * (bits: 1 transform length mismatch, 2 hash table overflow).  The
 * reference traps with device asserts instead (radix_join.cu:131-132,
 * 209-217).  Synchronous on `stream`. */
fij_status fij_device_status(void *stream, int clear, uint32_t *out);

/* ------------------------------------------------- runtime plumbing ---- */
/* HIP-backed stand-ins for the boundary types the reference's host code takes
 * from rustacuda and numa-gpu: DeviceBuffer / LockedBuffer and the Mem enum
 * (numa-gpu/src/runtime/memory.rs:36-95), Stream and Event.  Copies with
 * fij_memcpy complete before it returns; the fills are kernels (no memset
 * nodes), asynchronous on `stream`. */
fij_status fij_device_count(int *count);
fij_status fij_mem_alloc(uint64_t bytes, void **out);
fij_status fij_mem_free(void *ptr);
fij_status fij_mem_alloc_host(uint64_t bytes, void **out);
fij_status fij_mem_free_host(void *ptr);
fij_status fij_memcpy(void *dst, const void *src, uint64_t bytes, void *stream);
fij_status fij_memcpy_async(void *dst, const void *src, uint64_t bytes, void *stream);
fij_status fij_mem_fill_u32(void *ptr, uint32_t value, uint64_t count, void *stream);
fij_status fij_mem_fill_u64(void *ptr, uint64_t value, uint64_t count, void *stream);
/* A plain device copy kernel (16-byte loads and stores; buffers and size
 * 16-byte aligned), asynchronous on `stream`: the HBM bandwidth reference
 * measured beside the partition and join kernels (no counterpart in the
 * reference, whose measurements quote the nominal peak). */
fij_status fij_copy_kernel(void *dst, const void *src, uint64_t bytes, void *stream);
/* Calibration kernels, asynchronous on `stream` (no counterpart in the
 * reference): the access patterns of the pass-1 scatter and of the join's
 * gathers without their partitioning and table work, so that each kernel's
 * time can be set beside what its memory pattern alone costs on this GPU.
 * fij_calib_line_scatter copies `bytes` of 16-byte words from src to dst: block
 * c of `chunks` reads chunk c in order and writes whole 128-byte lines
 * round-robin to `partitions` regions, each holding the chunks' pieces in
 * chunk order (a chunk must be a multiple of 8192 words and of 8 x partitions
 * words).  fij_calib_run_gather reads src as partitions of part_words words,
 * each f2 items of part_words / f2 <= 1024 words gathered by one wave as runs
 * of run_words words (run r in tile r of the partition, tiles of f2 runs),
 * every run but a tile's last shifted by shift <= run_words words; the xor of
 * the loaded words goes into *checksum (device memory). */
fij_status fij_calib_line_scatter(void *dst, const void *src, uint64_t bytes, uint32_t partitions, uint32_t chunks,
                                  void *stream);
fij_status fij_calib_run_gather(const void *src, uint64_t bytes, uint32_t part_words, uint32_t f2,
                                uint32_t run_words, uint32_t shift, uint32_t *checksum, uint32_t grid, void *stream);
fij_status fij_stream_create(void **out);
fij_status fij_stream_destroy(void *stream);
fij_status fij_stream_synchronize(void *stream);
fij_status fij_device_synchronize(void);
fij_status fij_event_create(void **out);
fij_status fij_event_destroy(void *event);
fij_status fij_event_record(void *event, void *stream);
/* Milliseconds between two recorded events; waits for `stop`. */
fij_status fij_event_elapsed_ms(void *start, void *stop, float *ms);

/* --------------------------------------------- layout helpers (host) ---- */

/* partitioned_relation.rs:26-29: padding tuples per partition. */
uint32_t fij_padding_len(fij_type type);
/* partition_input_chunk.rs:24-35: canonical chunk length; INTEGER_OVERFLOW
 * when a chunk reaches u32::MAX tuples. */
fij_status fij_input_chunk_size(fij_type type, uint64_t data_len, uint32_t num_chunks, uint64_t *chunk_len);
/* PartitionOffsets::new sizes (partitioned_relation.rs:74-103). */
fij_status fij_partition_offsets_layout(fij_histogram_algorithm algo, uint32_t max_chunks, uint32_t radix_bits,
                                        uint32_t *chunks, uint64_t *offsets_len, uint64_t *local_offsets_len);
/* PartitionedRelation::new sizes (partitioned_relation.rs:335-362). */
fij_status fij_partitioned_relation_layout(fij_type type, uint64_t len, fij_histogram_algorithm algo,
                                           uint32_t radix_bits, uint32_t max_chunks, uint32_t *chunks,
                                           uint64_t *relation_len, uint64_t *offsets_len);
/* PartitionedRelation::padded_len / capacity / resize
 * (partitioned_relation.rs:365-420). */
uint64_t fij_partitioned_relation_padded_len(const fij_partitioned_relation *rel);
uint64_t fij_partitioned_relation_capacity(const fij_partitioned_relation *rel);
fij_status fij_partitioned_relation_resize(fij_partitioned_relation *rel, uint64_t new_len);
/* partition_len of PartitionedRelation / PartitionOffsets
 * (partitioned_relation.rs:429-459, :162-195).  `host_offsets` must be a
 * host-readable copy of the offsets. */
fij_status fij_partitioned_relation_partition_len(const fij_partitioned_relation *rel,
                                                  const uint64_t *host_offsets, uint32_t partition_id,
                                                  uint64_t *len);
fij_status fij_partition_offsets_partition_len(const fij_partition_offsets *offs, const uint64_t *host_offsets,
                                               uint32_t partition_id, uint64_t *len);

/* ------------------------------------------------------ partitioner ---- */

typedef struct fij_partitioner fij_partitioner;

/* GpuRadixPartitioner::new (gpu_radix_partition.rs:329-383).  block is
 * clamped to 1024 threads (the reference's default warp_size * 32 is 2048
 * with 64-wide waves). */
fij_status fij_partitioner_create(fij_histogram_algorithm histogram_algorithm,
                                  fij_partition_algorithm partition_algorithm, fij_radix_bits radix_bits,
                                  uint32_t grid_size, uint32_t block_size, uint64_t dmem_buffer_bytes,
                                  fij_partitioner **out);
fij_status fij_partitioner_destroy(fij_partitioner *p);

/* GpuRadixPartitioner::prefix_sum (gpu_radix_partition.rs:398-406, impl
 * :561-696; kernels radix_partition.cu:40-199).  Sets offsets->data_len. */
fij_status fij_prefix_sum(fij_partitioner *p, fij_radix_pass pass, fij_type type, const void *partition_attr,
                          uint64_t len, fij_partition_offsets *offsets, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fij_prefix_sum with the same arguments, checks and
 * results, but it calls no hipMalloc/hipFree.  The Contiguous histogram's
 * per-chunk counters and look-back state live in a static device array of
 * the library (FIJ_NOALLOC_MAX_COUNTERS counters, about 4 MiB, one copy per
 * GPU, held until the process ends) instead of the partitioner's buffers;
 * fanout x grid above FIJ_NOALLOC_MAX_COUNTERS is rejected with
 * FIJ_INVALID_ARGUMENT.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race.  The Chunked
 * histogram needs no scratch and runs as in fij_prefix_sum. */
#define FIJ_NOALLOC_MAX_COUNTERS (1u << 20)
fij_status fij_prefix_sum_noalloc(fij_partitioner *p, fij_radix_pass pass, fij_type type,
                                  const void *partition_attr, uint64_t len, fij_partition_offsets *offsets,
                                  void *stream);

/* GpuRadixPartitioner::prefix_sum_and_copy_with_payload
 * (gpu_radix_partition.rs:431-451, impl :698-824; radix_partition.cu:207-315). */
fij_status fij_prefix_sum_and_copy_with_payload(fij_partitioner *p, fij_radix_pass pass, fij_type type,
                                                const void *src_partition_attr, const void *src_payload_attr,
                                                uint64_t len, void *dst_partition_attr,
                                                void *dst_payload_attr, uint64_t dst_len,
                                                fij_partition_offsets *offsets, void *stream);

/* GpuRadixPartitioner::prefix_sum_and_transform (gpu_radix_partition.rs:480-500,
 * impl :826-957; radix_partition.cu:323-467).  dst_len is the length of the
 * dst slices, i.e. the partition's length. */
fij_status fij_prefix_sum_and_transform(fij_partitioner *p, fij_radix_pass pass, fij_type type,
                                        uint32_t partition_id, const fij_partitioned_relation *src_relation,
                                        void *dst_partition_attr, void *dst_payload_attr, uint64_t dst_len,
                                        fij_partition_offsets *offsets, void *stream);

/* GpuRadixPartitioner::preallocate_partition_state (gpu_radix_partition.rs:511-526). */
fij_status fij_preallocate_partition_state(fij_partitioner *p, fij_radix_pass pass, fij_type type);

/* GpuRadixPartitioner::partition (gpu_radix_partition.rs:535-554, impl
 * :1022-1239; kernels radix_partition.cu:472-1477).  Afterwards the
 * relation owns the offsets: relation->offsets and offsets->offsets are
 * swapped (:1117-1121). */
fij_status fij_partition(fij_partitioner *p, fij_radix_pass pass, fij_type type, const void *partition_attr,
                         const void *payload_attr, uint64_t len, fij_partition_offsets *offsets,
                         fij_partitioned_relation *relation, void *stream);

/* --------------------------------------------------------- radix join ---- */

typedef struct fij_radix_join fij_radix_join;

/* CudaRadixJoin::new (sql-ops/src/join/cuda_radix_join.rs:116-130). */
fij_status fij_radix_join_create(fij_radix_pass pass, fij_radix_bits radix_bits, fij_hashing_scheme scheme,
                                 uint32_t grid_size, uint32_t block_size, fij_radix_join **out);
fij_status fij_radix_join_destroy(fij_radix_join *j);

/* CudaRadixJoin::join (cuda_radix_join.rs:133-164, impl :172-276; kernels
 * radix_join.cu:58-311).  result_set: grid*block per-thread SUM(probe.value)
 * partials, added to (caller zeroes), may be NULL; aggregates: grid
 * per-block {count, sum_build, sum_probe}, added to, may be NULL;
 * task_assignments: grid + 1 entries of device scratch. */
fij_status fij_radix_join_join(fij_radix_join *j, fij_type type, const fij_partitioned_relation *build_rel,
                               const fij_partitioned_relation *probe_rel, int64_t *result_set,
                               uint64_t result_set_len, uint32_t *task_assignments,
                               uint64_t task_assignments_len, fij_join_aggregate *aggregates, void *stream);

/* gpu_radix_join_assign_tasks alone (radix_join.cu:58-98). */
fij_status fij_join_assign_tasks(const fij_partitioned_relation *probe_rel, uint32_t radix_bits,
                                 uint32_t grid_size, uint32_t *task_assignments, void *stream);

/* ---------------------------------------------- no-partitioning join ---- */

/* HashTable::new_on_gpu (sql-ops/src/join/no_partitioning_join.rs:588-634):
 * sets every 32-bit word of the table to the null key, as the reference's
 * cuMemsetD32 does. */
fij_status fij_np_table_init(fij_type type, void *table, uint64_t entries, void *stream);
/* Sets every tuple of a relation buffer to {null key, 0}, the padding value
 * the reference's partition tests allocate with
 * (sql-ops/tests/test_gpu_radix_partition.rs:88-97). */
fij_status fij_relation_fill_null(fij_type type, void *relation, uint64_t tuples, void *stream);
/* CudaHashJoin::build (no_partitioning_join.rs:255-268, impl :313-375; kernels
 * no_partitioning_join.cu:115-132, :251-319).  is_selective skips null
 * build keys (perfect hashing only). */
fij_status fij_np_build(fij_type type, fij_hashing_scheme scheme, int is_selective, void *table, uint64_t entries,
                        const void *keys, const void *payloads, uint64_t len, uint32_t grid_size,
                        uint32_t block_size, void *stream);
/* CudaHashJoin::probe_sum (no_partitioning_join.rs:270-284, impl :377-440; kernels
 * no_partitioning_join.cu:225-249, :321-379).  result_set: grid*block u64
 * partials of SUM(probe.value), added to; aggregates: grid entries, added
 * to; either may be NULL. */
fij_status fij_np_probe_sum(fij_type type, fij_hashing_scheme scheme, const void *table, uint64_t entries,
                            const void *keys, const void *payloads, uint64_t len, uint64_t *result_set,
                            uint64_t result_set_len, fij_join_aggregate *aggregates, uint32_t grid_size,
                            uint32_t block_size, void *stream);

/* ----------------------------------------- radix-join execution method ---- */
/* The radix-join binary's GpuRadixJoinTwoPass execution method
 * (radix-join/src/execution_methods/gpu_radix_join.rs:90-565), which the
 * reference builds from the operators above with a host loop of six launches
 * per pass-1 partition.  A plan allocates every intermediate buffer of one
 * join shape once; executing it runs pass 1 (chunked histogram + buffered
 * scatter), pass 2 (per-partition LDS counting sort of tiles) and the LDS
 * hash join for all partitions with a fixed, small number of launches, and
 * returns {count, SUM(build.value), SUM(probe.value)}. */

typedef struct fij_join_plan fij_join_plan;

typedef struct fij_join_plan_config {
  uint32_t type;         /* fij_type */
  uint32_t radix_bits_1; /* pass-1 bits (RadixBits first pass) */
  uint32_t radix_bits_2; /* pass-2 bits; 0 = single pass */
  uint32_t pass1_chunks; /* pass-1 chunks (= blocks); 0 = one per compute unit (fewer for short relations) */
  uint32_t pass1_block;  /* threads per pass-1 block; 0 = 1024 */
  uint32_t join_blocks;  /* resident blocks of the join; 0 = two per compute unit */
  uint64_t build_len;    /* |R| */
  uint64_t probe_len;    /* |S| */
} fij_join_plan_config;

/* Phase times as in the reference's RadixJoinPoint
 * (radix-join/src/measurement/data_point.rs:29-74): prefix sum (pass-1
 * histograms), partition (pass-1 scatter), join (pass 2 + build/probe);
 * join_ms splits into pass2_ms (plan + pass-2 sort), build_probe_ms (the
 * register-resident build/probe launch) and deferred_ms (the launch that
 * joins the partitions too large for it). */
typedef struct fij_join_timing {
  float prefix_sum_ms;
  float partition_ms;
  float join_ms;
  float total_ms;
  float pass2_ms;
  float build_probe_ms;
  float deferred_ms;
} fij_join_timing;

/* Radix bits for one join run alone (a plan, or a served request): one
 * pass of 9 bits when the build side's pass-1 partitions hold on average at
 * most 2/3 of a 128 KiB block-shared LDS table (|R| <= 2^21 at int64, 2^22
 * at int32; plans whose partitions also fit 64 KiB tables use those), so
 * that the probe side is partitioned once and its partitions are streamed
 * against them.  Otherwise two passes: for a probe side of 16 |R| or more,
 * 9 + b2 bits with items of about a third of a 64 KiB block table's build
 * tuples (1,024 int64, 2,048 int32), else the bits of fij_join_autotune. 
 * No counterpart in the reference, whose caller chooses the bits. */
fij_status fij_join_plan_bits(fij_type type, uint64_t build_len, uint64_t probe_len, uint32_t *radix_bits_1,
                              uint32_t *radix_bits_2);
fij_status fij_join_plan_create(const fij_join_plan_config *config, fij_join_plan **out);
fij_status fij_join_plan_destroy(fij_join_plan *plan);
/* Device bytes the plan holds (partitions, offsets, tile metadata). */
fij_status fij_join_plan_device_bytes(const fij_join_plan *plan, uint64_t *bytes);
/* Enqueues one join on `stream`; the result is read with fij_join_plan_result. */
fij_status fij_join_plan_enqueue(fij_join_plan *plan, const void *build_keys, const void *build_payloads,
                                 const void *probe_keys, const void *probe_payloads, void *stream);
/* Enqueues one join over columns of build_len and probe_len tuples (at most
 * the plan's lengths) and writes its {count, sums} to *device_result (device
 * memory); no events, no host copies and no allocations, so it can be
 * recorded into a graph.  Joins enqueued on one stream run one after
 * another; a plan serves one stream at a time. */
fij_status fij_join_plan_enqueue_sized(fij_join_plan *plan, const void *build_keys, const void *build_payloads,
                                       uint64_t build_len, const void *probe_keys, const void *probe_payloads,
                                       uint64_t probe_len, fij_join_aggregate *device_result, void *stream);
/* Waits for the enqueued join; fills result and (optionally) phase times. */
fij_status fij_join_plan_result(fij_join_plan *plan, void *stream, fij_join_aggregate *result,
                                fij_join_timing *timing);
/* enqueue + result. */
fij_status fij_join_plan_execute(fij_join_plan *plan, const void *build_keys, const void *build_payloads,
                                 const void *probe_keys, const void *probe_payloads, fij_join_aggregate *result,
                                 fij_join_timing *timing, void *stream);

/* ------------------------------------------------------ batched joins ---- */
/* Many independent joins with one submission (no counterpart in the
 * reference, which joins one pair per run of the radix-join binary).  Joins
 * whose build side fits one block's LDS table are joined together by one
 * persistent launch without partitioning; every other join is partitioned
 * in one or two passes with the radix bits of its size class
 * (fij_join_autotune: one pass of 9 bits for build sides from 2^18 tuples up
 * to what 64 KiB block tables hold after it: 2^20 int64, 2^21 int32),
 * all joins of a class together: one launch per pass (histograms, offsets,
 * scatter, pass 2, join) covers every join of the class.  The submission is
 * recorded into one HIP graph when use_graph is set.  The size classes are
 * independent and are spread over `lanes` streams (forked from the
 * submission stream and joined back into it), so classes that do not fill
 * the GPU overlap.  Results are per join, identical to joining each pair
 * alone.  Columns need only be aligned to their element size. */

typedef struct fij_join_desc {
  const void *build_keys;
  const void *build_payloads;
  uint64_t build_len;
  const void *probe_keys;
  const void *probe_payloads;
  uint64_t probe_len;
} fij_join_desc;

typedef struct fij_join_batch_config {
  uint32_t type;      /* fij_type */
  uint32_t max_joins; /* joins per execute call */
  uint32_t use_graph; /* record the submission into a HIP graph */
  uint32_t lanes;     /* streams the independent size classes are spread over (0 or 1: one; at most 4) */
} fij_join_batch_config;

typedef struct fij_join_batch_stats {
  float gpu_ms;      /* device time of the submission */
  float prepare_ms;  /* host: classification, plan growth, recording */
  float total_ms;    /* host: whole execute call */
  uint32_t small_joins;
  uint32_t partitioned_joins;
  uint32_t size_classes;
  uint32_t graph;
  float grow_ms;     /* host, part of prepare_ms: growing size-class buffers (device allocation) */
  float record_ms;   /* host, part of prepare_ms: recording and instantiating the graph */
} fij_join_batch_stats;

typedef struct fij_join_batch fij_join_batch;

/* The size-class choice: small = 1 when one block joins the pair from LDS;
 * otherwise the plan's radix bits for these lengths.  Looked up in the
 * tuning table (fij_join_tuning_lookup); lengths of a class the table does
 * not hold use the closed-form rule. */
fij_status fij_join_autotune(fij_type type, uint64_t build_len, uint64_t probe_len, uint32_t *small,
                             uint32_t *radix_bits_1, uint32_t *radix_bits_2);

/* Size-class tuning (no counterpart in the reference, whose radix bits and
 * buffer sizes are fixed per run by scripts/benchmark_gpu_radix_join.py:55-67
 * and sql-ops/build.rs:65-67).  A class is (ceil(log2 |R|), ceil(log2 |S|));
 * its entry holds the configuration measured fastest for batched joins of
 * that class on this GPU: the small-join path or the two-pass plan with its
 * radix bits, the pass-1 line buffer depth and the pass-1 block shape.  The
 * library carries a built-in table measured on MI355X
 * (fast-interconnects_amd/autotune.py writes it). */
typedef struct fij_join_tuning {
  uint32_t build_log2;   /* class: ceil(log2 build_len) */
  uint32_t probe_log2;   /* class: ceil(log2 probe_len) */
  uint32_t small;        /* 1: one block joins the pair from an LDS table */
  uint32_t radix_bits_1; /* two-pass plan: pass-1 and pass-2 radix bits */
  uint32_t radix_bits_2;
  uint32_t ring_lines;   /* pass-1 line buffer: 128-byte lines per partition, 0 = deepest that fits LDS */
  uint32_t block;        /* pass-1 threads per block: 512 or 1024 */
  uint32_t from_table;   /* lookup output: 1 = a table entry, 0 = the closed-form rule */
} fij_join_tuning;

/* The configuration batched joins of these lengths run with. */
fij_status fij_join_tuning_lookup(fij_type type, uint64_t build_len, uint64_t probe_len, fij_join_tuning *out);
/* Replaces the active table of `type` with entries[0..n) (n = 0: the rule
 * only); a later call with use_builtin = 1 restores the built-in table.
 * Entries are validated (bits, block, ring); batches executed afterwards use
 * the new table. */
fij_status fij_join_tuning_set(fij_type type, const fij_join_tuning *entries, uint32_t n, int use_builtin);
fij_status fij_join_batch_create(const fij_join_batch_config *config, fij_join_batch **out);
fij_status fij_join_batch_destroy(fij_join_batch *batch);
/* Joins joins[0..n) (device columns) and writes results[0..n) (host).  A
 * size class's buffers are kept between calls and grown (by a quarter more
 * than asked) when a call needs more.  Synchronous. */
fij_status fij_join_batch_execute(fij_join_batch *batch, uint32_t n, const fij_join_desc *joins,
                                  fij_join_aggregate *results, fij_join_batch_stats *stats, void *stream);
fij_status fij_join_batch_device_bytes(const fij_join_batch *batch, uint64_t *bytes);

/* ------------------------------------------------------- served joins ---- */
/* A request loop over resident build sides (no counterpart in the reference,
 * whose radix-join binary partitions both relations of every join).  Build
 * sides are partitioned once and kept in device memory; each request brings
 * a probe side, waits in arrival order for a free slot that fits it (a
 * preallocated probe workspace with its own stream), partitions its probe
 * side and joins it with the resident partitions.  Admitted requests overlap
 * on the device; results equal joining each pair alone.  Memory: slots and
 * resident builds are charged against hbm_budget (0 = unlimited). */

#define FIJ_SERVER_MAX_SLOT_CLASSES 4

typedef struct fij_join_server_config {
  uint32_t type;                                         /* fij_type */
  uint32_t slot_classes;                                 /* <= FIJ_SERVER_MAX_SLOT_CLASSES */
  uint32_t slots[FIJ_SERVER_MAX_SLOT_CLASSES];           /* slots per class */
  uint64_t slot_probe_len[FIJ_SERVER_MAX_SLOT_CLASSES];  /* largest probe side of a class's slot */
  uint64_t max_probe_len;                                /* largest probe side expected (radix bits of builds) */
  uint64_t hbm_budget;                                   /* bytes; 0 = unlimited */
  uint64_t table_build_max;                              /* build sides of at most this many tuples are kept as
                                                            one linear-probing table (2x their tuples, a power of
                                                            two) and probed in one launch; 0 = every build side
                                                            is partitioned */
} fij_join_server_config;

typedef struct fij_join_request_times {
  double queued_ms;  /* arrival to admission */
  double latency_ms; /* arrival to result */
} fij_join_request_times;

typedef struct fij_join_server fij_join_server;

fij_status fij_join_server_create(const fij_join_server_config *config, fij_join_server **out);
/* Waits for every admitted request, then releases everything. */
fij_status fij_join_server_destroy(fij_join_server *server);
/* Partitions a build side (device columns, read only during the call) and
 * keeps it resident; synchronous. */
fij_status fij_join_server_add_build(fij_join_server *server, const void *keys, const void *payloads, uint64_t len,
                                     uint32_t *build_id);
/* Queues a request (device probe columns, valid until its result is taken);
 * returns at once.  Thread-safe. */
fij_status fij_join_server_submit(fij_join_server *server, uint32_t build_id, const void *probe_keys,
                                  const void *probe_payloads, uint64_t probe_len, uint64_t *request_id);
/* Waits for a request's result and its queueing time and latency. */
fij_status fij_join_server_wait(fij_join_server *server, uint64_t request_id, fij_join_aggregate *result,
                                fij_join_request_times *times);
fij_status fij_join_server_memory(const fij_join_server *server, uint64_t *used_bytes, uint32_t *builds);

/* ------------------------------------------------ host (CPU) partitioner ---- */
/* The GPU prefix sum's offsets computed on host threads, from host keys into
 * host offset buffers (CpuRadixPartitioner::prefix_sum,
 * sql-ops/src/partition/cpu_radix_partition.rs:46-62 "CPU and GPU histograms
 * ... can be mixed and matched"; radix_partition.cpp:228-267).  Chunked and
 * Contiguous give exactly fij_prefix_sum's offsets for max_chunks chunks, so
 * a caller can copy them into device offsets and run fij_partition, reading
 * the keys over the interconnect once instead of twice. */
fij_status fij_cpu_prefix_sum(fij_histogram_algorithm algorithm, fij_type type, const void *keys, uint64_t len,
                              uint32_t radix_bits, uint32_t ignore_bits, uint32_t threads,
                              fij_partition_offsets *offsets);
/* Pass 1 on host threads into host memory, in the contiguous layout
 * (relation: len + 2^radix_bits * pad tuples; offsets[2^radix_bits] the
 * partition starts), as the GPU's contiguous pass 1 lays it out
 * (cpu_radix_partition.rs:364-640, radix_partition.cpp:448-557 write-combining
 * scatter).  Order within a partition: by thread, then input order. */
fij_status fij_cpu_radix_partition(fij_type type, const void *keys, const void *payloads, uint64_t len,
                                   uint32_t radix_bits, uint32_t ignore_bits, uint32_t threads, void *relation,
                                   uint64_t *offsets);

/* ---------------------------------------------------- out-of-core join ---- */
/* Relations larger than device memory (radix-join/src/execution_methods/
 * gpu_triton_join.rs:99-638, the Triton join; with pass1_cpu = 1
 * cpu_partitioned_radix_join.rs:118-769).  Pass 1 runs on the GPU, reading
 * the relations in place (device or pinned host memory), or on host threads
 * (any host memory); its partitions are cached in device memory up to
 * cache_bytes (GPU pass 1: split between the relations in proportion to
 * their lengths; CPU pass 1: the build relation's, copied while the host
 * threads partition the probe relation slice by slice, every slice joined
 * on the GPU while the next is partitioned) and spilled to pinned host
 * memory beyond it.  Pass 2 and
 * the join then run over groups of consecutive pass-1 partitions (at most
 * staging_tuples tuples of either relation, or one partition when a single
 * one is longer) on two streams, spilled groups copied into device memory
 * first.  The result equals the in-memory join's.  Synchronous; allocates
 * and frees its buffers (not part of the timing). */
typedef struct fij_ooc_config {
  uint32_t type;           /* fij_type */
  uint32_t radix_bits_1;   /* 1..12 */
  uint32_t radix_bits_2;   /* 1..12 */
  uint32_t pass1_cpu;      /* 0: pass 1 on the GPU, 1: on host threads */
  uint32_t cpu_threads;    /* host threads of a CPU pass 1 (0 = all) */
  uint32_t probe_slices;   /* CPU pass 1: the probe relation is partitioned in this many slices of its
                              input, each joined on the GPU while the next is partitioned (0 = one per
                              staging_tuples, at most 8) */
  uint64_t cache_bytes;    /* device memory for pass-1 partitions */
  uint64_t staging_tuples; /* group size (0 = 2^24) */
} fij_ooc_config;

/* RadixJoinPoint (radix-join/src/measurement/data_point.rs:29-74) */
typedef struct fij_ooc_timing {
  float prefix_sum_ms;
  float partition_ms;
  float join_ms;       /* pass 2 + join over all groups, copies included */
  float total_ms;
  uint64_t cached_build_tuples;
  uint64_t cached_probe_tuples;
  uint32_t groups;
  uint32_t reserved;
} fij_ooc_timing;

fij_status fij_ooc_join(const fij_ooc_config *config, const void *build_keys, const void *build_payloads,
                        uint64_t build_len, const void *probe_keys, const void *probe_payloads, uint64_t probe_len,
                        fij_join_aggregate *result, fij_ooc_timing *timing);

/* ----------------------------------------------------- relation files ---- */
/* The reference's relation files (data-store/src/join_data.rs:173-280 reads,
 * datagen/src/bin/generate.rs:343-365 writes): a header record, then one
 * `key<delim>value` record per tuple; the reader uses delimiter ' '
 * (delim = 0 selects it), skips the header and blank lines, accepts CRLF and
 * gzip-compressed files, and parses each field as an integer of the key
 * type (optional sign, decimal digits, in range) -- anything else is
 * FIJ_PARSE_ERROR with the line number.  Columns are host memory. */

/* Number of records (tuples) in a relation file. */
fij_status fij_relation_file_len(const char *path, char delim, uint64_t *len);
/* Reads up to capacity tuples into keys / payloads; *len = tuples read. */
fij_status fij_relation_file_read(const char *path, char delim, fij_type type, void *keys, void *payloads,
                                  uint64_t capacity, uint64_t *len);
/* Writes a relation file ("key<delim>value" header); gzip when the path ends in "gz". */
fij_status fij_relation_file_write(const char *path, char delim, fij_type type, const void *keys,
                                   const void *payloads, uint64_t len);

/* ---------------------------------------------------- data generation ---- */
/* Seeded replacements for the reference's unseeded generators
 * (datagen/src/relation.rs:76-290); each element is a pure function of
 * (seed, stream, index). */

/* out[i] = base + perm(i), perm a seeded bijection on [0, n) (primary keys). */
fij_status fij_gen_permutation(fij_type type, void *out, uint64_t n, uint64_t seed, int64_t base, void *stream);
/* out[i] = lo + floor(u_i * range), u_i uniform (foreign keys, attributes). */
fij_status fij_gen_uniform(fij_type type, void *out, uint64_t n, uint64_t seed, uint64_t stream_id, int64_t lo,
                           uint64_t range, void *stream);
/* out[i] = base + i (payloads). */
fij_status fij_gen_iota(fij_type type, void *out, uint64_t n, int64_t base, void *stream);
/* Skewed foreign keys (the datagen Zipf relation, datagen/src/relation.rs:218-290):
 * out[i] = base + perm(rank_i - 1), rank_i ~ Zipf(exponent) over 1..num_ranks
 * by rejection-inversion and perm the permutation fij_gen_permutation uses for
 * num_ranks keys with perm_seed -- so the keys are those of that relation and
 * the most frequent one sits at a seeded position. */
fij_status fij_gen_zipf(fij_type type, void *out, uint64_t n, uint64_t num_ranks, double exponent, uint64_t seed,
                        uint64_t stream_id, uint64_t perm_seed, int64_t base, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FIJ_H */
