/*
 * oracle.h -- CPU restatement of the reference radix-join path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (fast-interconnects_amd/) never links, loads
 * or calls it, and must fail loudly when its own HIP library is missing.
 *
 * Every function cites the reference file:line it restates.  Pins (what checks
 * each function against something other than itself) are listed per function;
 * see tests/test_oracle_*.py.
 *
 * Conventions shared with the reference (sql-ops/include/gpu_common.h:50-123,
 * sql-ops/src/partition/partitioned_relation.rs:26-29, sql-ops/build.rs:44-72):
 *   - tuples are row structs {key, value}, 8 B (i32) or 16 B (i64);
 *   - the null key is all ones (-1);
 *   - PADDING_BYTES = ALIGN_BYTES = 128, so pad = 128 / sizeof(tuple).
 */
#ifndef FIJ_ORACLE_H
#define FIJ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ALIGN_BYTES 128u
#define ORC_PADDING_BYTES 128u

typedef struct { int32_t key; int32_t value; } orc_tuple32;
typedef struct { int64_t key; int64_t value; } orc_tuple64;

/* Join aggregate: match count and wrapping 64-bit sums of build (R) and
 * probe (S) payloads over all matching pairs. */
typedef struct {
  uint64_t count;
  uint64_t sum_build;
  uint64_t sum_probe;
} orc_join_result;

/* ---------------- generators (seeded, counter-based) ------------------- */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_rand_u64(uint64_t seed, uint64_t stream, uint64_t i);
uint64_t orc_feistel_perm(uint64_t i, uint64_t n, uint64_t seed);
void orc_gen_perm_keys_i32(int32_t *out, uint64_t n, uint64_t seed, int64_t base);
void orc_gen_perm_keys_i64(int64_t *out, uint64_t n, uint64_t seed, int64_t base);
void orc_gen_uniform_i32(int32_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                         int64_t lo, uint64_t range);
void orc_gen_uniform_i64(int64_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                         int64_t lo, uint64_t range);
void orc_gen_iota_i32(int32_t *out, uint64_t n, int64_t base);
void orc_gen_iota_i64(int64_t *out, uint64_t n, int64_t base);
int orc_gen_zipf_ranks(uint64_t *out, uint64_t n, uint64_t num_ranks, double exponent,
                       uint64_t seed, uint64_t stream);
/* base + perm(rank - 1) for Zipf ranks over num_ranks keys (see oracle.c). */
int orc_gen_zipf_keys_i32(int32_t *out, uint64_t n, uint64_t num_ranks, double s, uint64_t seed, uint64_t stream,
                          uint64_t perm_seed, int64_t base);
int orc_gen_zipf_keys_i64(int64_t *out, uint64_t n, uint64_t num_ranks, double s, uint64_t seed, uint64_t stream,
                          uint64_t perm_seed, int64_t base);

/* ---------------- partitioning ------------------------------------------ */
uint32_t orc_key_to_partition_i32(int32_t key, uint32_t radix_bits, uint32_t ignore_bits);
uint32_t orc_key_to_partition_i64(int64_t key, uint32_t radix_bits, uint32_t ignore_bits);
/* returns 0 on success, -1 on the reference's IntegerOverflow condition */
int orc_input_chunk_size(uint64_t data_len, uint32_t num_chunks, uint32_t key_bytes,
                         uint64_t *chunk_len);

void orc_chunked_prefix_sum_i32(const int32_t *keys, uint64_t n, uint32_t chunks,
                                uint32_t radix_bits, uint32_t ignore_bits,
                                uint64_t *offsets /* [chunks * F] */);
void orc_chunked_prefix_sum_i64(const int64_t *keys, uint64_t n, uint32_t chunks,
                                uint32_t radix_bits, uint32_t ignore_bits,
                                uint64_t *offsets);
void orc_contiguous_prefix_sum_i32(const int32_t *keys, uint64_t n, uint32_t chunks,
                                   uint32_t radix_bits, uint32_t ignore_bits,
                                   uint64_t *offsets /* [F] */,
                                   uint64_t *local_offsets /* [chunks * F] */);
void orc_contiguous_prefix_sum_i64(const int64_t *keys, uint64_t n, uint32_t chunks,
                                   uint32_t radix_bits, uint32_t ignore_bits,
                                   uint64_t *offsets, uint64_t *local_offsets);
void orc_partition_i32(const int32_t *keys, const int32_t *pays, uint64_t n,
                       uint32_t chunks, uint32_t radix_bits, uint32_t ignore_bits,
                       const uint64_t *chunk_offsets /* [chunks * F] */,
                       orc_tuple32 *relation);
void orc_partition_i64(const int64_t *keys, const int64_t *pays, uint64_t n,
                       uint32_t chunks, uint32_t radix_bits, uint32_t ignore_bits,
                       const uint64_t *chunk_offsets, orc_tuple64 *relation);
uint64_t orc_partition_len(const uint64_t *offsets, uint32_t chunks, uint32_t radix_bits,
                           uint64_t padded_len, uint32_t pad, uint32_t partition_id);
uint64_t orc_transform_i32(const orc_tuple32 *src, const uint64_t *src_offsets,
                           uint32_t src_chunks, uint32_t src_radix_bits,
                           uint64_t src_padded_len, uint32_t partition_id,
                           int32_t *dst_key, int32_t *dst_pay);
uint64_t orc_transform_i64(const orc_tuple64 *src, const uint64_t *src_offsets,
                           uint32_t src_chunks, uint32_t src_radix_bits,
                           uint64_t src_padded_len, uint32_t partition_id,
                           int64_t *dst_key, int64_t *dst_pay);

/* ---------------- join ---------------------------------------------------- */
void orc_assign_tasks(const uint64_t *probe_offsets, uint32_t probe_len,
                      uint32_t probe_pad, uint32_t radix_bits, uint32_t grid,
                      uint32_t *task_assignments /* [grid + 1] */);
int32_t orc_mult_shift_hash_i32(int32_t value, uint32_t log2_buckets);
int64_t orc_mult_shift_hash_i64(int64_t value, uint32_t log2_buckets);
uint32_t orc_fmix32(uint32_t h);
uint64_t orc_fmix64(uint64_t k);
int32_t orc_murmur3_hash_i32(int32_t value, uint32_t log2_buckets);
int64_t orc_murmur3_hash_i64(int64_t value, uint32_t log2_buckets);

/* Per-partition bucket-chaining join over two contiguous partitioned
 * relations (one chunk each); returns 0, or -1 if a build partition exceeds
 * max_build (the reference's LDS-capacity assert). */
int orc_partitioned_join_i32(const orc_tuple32 *build, const uint64_t *build_offsets,
                             uint64_t build_padded_len, const orc_tuple32 *probe,
                             const uint64_t *probe_offsets, uint64_t probe_padded_len,
                             uint32_t radix_bits, uint32_t ignore_bits,
                             uint64_t max_build, orc_join_result *res);
int orc_partitioned_join_i64(const orc_tuple64 *build, const uint64_t *build_offsets,
                             uint64_t build_padded_len, const orc_tuple64 *probe,
                             const uint64_t *probe_offsets, uint64_t probe_padded_len,
                             uint32_t radix_bits, uint32_t ignore_bits,
                             uint64_t max_build, orc_join_result *res);
/* Perfect-hash variant: table index = key >> ignore_bits (dense unique keys). */
int orc_partitioned_join_perfect_i32(const orc_tuple32 *build, const uint64_t *build_offsets,
                                     uint64_t build_padded_len, const orc_tuple32 *probe,
                                     const uint64_t *probe_offsets, uint64_t probe_padded_len,
                                     uint32_t radix_bits, uint32_t ignore_bits,
                                     uint64_t ht_entries, orc_join_result *res);
int orc_partitioned_join_perfect_i64(const orc_tuple64 *build, const uint64_t *build_offsets,
                                     uint64_t build_padded_len, const orc_tuple64 *probe,
                                     const uint64_t *probe_offsets, uint64_t probe_padded_len,
                                     uint32_t radix_bits, uint32_t ignore_bits,
                                     uint64_t ht_entries, orc_join_result *res);

/* No-partitioning linear-probing join (global table of next_pow2 entries). */
void orc_np_build_lp_i32(orc_tuple32 *table, uint64_t entries, const int32_t *keys,
                         const int32_t *pays, uint64_t n);
void orc_np_build_lp_i64(orc_tuple64 *table, uint64_t entries, const int64_t *keys,
                         const int64_t *pays, uint64_t n);
void orc_np_probe_lp_i32(const orc_tuple32 *table, uint64_t entries, const int32_t *keys,
                         const int32_t *pays, uint64_t n, orc_join_result *res);
void orc_np_probe_lp_i64(const orc_tuple64 *table, uint64_t entries, const int64_t *keys,
                         const int64_t *pays, uint64_t n, orc_join_result *res);

/* End-to-end radix join (1 or 2 passes, bits2 == 0 means one pass) with
 * T threads; R = build, S = probe.  Returns 0 on success. */
int orc_radix_join_i32(const int32_t *r_key, const int32_t *r_pay, uint64_t r_n,
                       const int32_t *s_key, const int32_t *s_pay, uint64_t s_n,
                       uint32_t bits1, uint32_t bits2, uint32_t threads,
                       orc_join_result *res);
int orc_radix_join_i64(const int64_t *r_key, const int64_t *r_pay, uint64_t r_n,
                       const int64_t *s_key, const int64_t *s_pay, uint64_t s_n,
                       uint32_t bits1, uint32_t bits2, uint32_t threads,
                       orc_join_result *res);

#ifdef __cplusplus
}
#endif
#endif /* FIJ_ORACLE_H */
