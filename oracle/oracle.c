/*
 * oracle.c -- CPU restatement of the reference radix-join path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for tests/, smoke()
 * and bench.py's cpu_baseline leg; never part of the product path.
 *
 * Reference files restated (paths relative to the reference repository):
 *   sql-ops/include/gpu_radix_partition.h:121-127   partition function
 *   sql-ops/src/partition/partition_input_chunk.rs:24-35   chunk length
 *   sql-ops/cudautils/radix_partition.cu:40-96      chunked prefix sum
 *   sql-ops/cpputils/radix_partition.cpp:228-267    (same, CPU form)
 *   sql-ops/cudautils/radix_partition.cu:104-199    contiguous prefix sum
 *   sql-ops/cpputils/radix_partition.cpp:409-446    NC partition
 *   sql-ops/src/partition/partitioned_relation.rs:429-459   partition_len
 *   sql-ops/cudautils/radix_partition.cu:343-410    pass-2 transform order
 *   sql-ops/cudautils/radix_join.cu:58-98           task assignment
 *   sql-ops/cudautils/radix_join.cu:103-194,199-311 LDS join semantics
 *   sql-ops/include/gpu_common.h:143-212            hash functions
 *   sql-ops/cpputils/no_partitioning_join.cpp:27-163 NP linear probing join
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================= */
/* Generators: counter-based so that any element can be produced alone (the
 * GPU library generates the same integer sequences in parallel).  The
 * reference's own generators are unseeded (datagen/src/relation.rs:94-173),
 * so these are this build's definitions, pinned by distribution tests. */

uint64_t orc_mix64(uint64_t z) {
  /* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

uint64_t orc_rand_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t k = orc_mix64(seed ^ orc_mix64(stream + 0x9e3779b97f4a7c15ULL));
  return orc_mix64(k + (i + 1) * 0x9e3779b97f4a7c15ULL);
}

/* Bijection on [0, n): a 4-round balanced Feistel network on 2h bits
 * (4^h >= n) with cycle walking. */
uint64_t orc_feistel_perm(uint64_t i, uint64_t n, uint64_t seed) {
  if (n <= 1) return 0;
  unsigned b = 64u - (unsigned)__builtin_clzll(n - 1);
  unsigned h = (b + 1u) / 2u;
  uint64_t mask = (h >= 64u) ? ~0ULL : ((1ULL << h) - 1ULL);
  uint64_t x = i;
  do {
    uint64_t l = x >> h, r = x & mask;
    for (unsigned round = 0; round < 4; ++round) {
      uint64_t f = orc_mix64(r ^ orc_mix64(seed + (round + 1) * 0x9e3779b97f4a7c15ULL)) & mask;
      uint64_t nl = r;
      r = l ^ f;
      l = nl;
    }
    x = (l << h) | r;
  } while (x >= n);
  return x;
}

void orc_gen_perm_keys_i32(int32_t *out, uint64_t n, uint64_t seed, int64_t base) {
  for (uint64_t i = 0; i < n; ++i) out[i] = (int32_t)(base + (int64_t)orc_feistel_perm(i, n, seed));
}
void orc_gen_perm_keys_i64(int64_t *out, uint64_t n, uint64_t seed, int64_t base) {
  for (uint64_t i = 0; i < n; ++i) out[i] = base + (int64_t)orc_feistel_perm(i, n, seed);
}

static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
}

void orc_gen_uniform_i32(int32_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                         int64_t lo, uint64_t range) {
  for (uint64_t i = 0; i < n; ++i)
    out[i] = (int32_t)(lo + (int64_t)mulhi64(orc_rand_u64(seed, stream, i), range));
}
void orc_gen_uniform_i64(int64_t *out, uint64_t n, uint64_t seed, uint64_t stream,
                         int64_t lo, uint64_t range) {
  for (uint64_t i = 0; i < n; ++i)
    out[i] = lo + (int64_t)mulhi64(orc_rand_u64(seed, stream, i), range);
}
void orc_gen_iota_i32(int32_t *out, uint64_t n, int64_t base) {
  for (uint64_t i = 0; i < n; ++i) out[i] = (int32_t)(base + (int64_t)i);
}
void orc_gen_iota_i64(int64_t *out, uint64_t n, int64_t base) {
  for (uint64_t i = 0; i < n; ++i) out[i] = base + (int64_t)i;
}

/* Zipf(s) over ranks 1..N by rejection-inversion, after W. Hormann and
 * G. Derflinger, "Rejection-inversion to generate variates from monotone
 * discrete distributions", ACM TOMACS 6(3), 1996.  H is the antiderivative
 * of h(x) = x^-s written with log1p/expm1 forms for stability near s = 1. */
static double zh1(double x) { return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x)); }
static double zh2(double x) { return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x)); }
static double zipf_H(double x, double s) { double lx = log(x); return zh2((1.0 - s) * lx) * lx; }
static double zipf_Hinv(double y, double s) {
  double t = y * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zh1(t) * y);
}
static double zipf_h(double x, double s) { return exp(-s * log(x)); }

int orc_gen_zipf_ranks(uint64_t *out, uint64_t n, uint64_t num_ranks, double s,
                       uint64_t seed, uint64_t stream) {
  if (num_ranks < 1 || s < 0.0) return -1;
  const double N = (double)num_ranks;
  const double hx1 = zipf_H(1.5, s) - 1.0;
  const double hN = zipf_H(N + 0.5, s);
  const double sc = 2.0 - zipf_Hinv(zipf_H(2.5, s) - zipf_h(2.0, s), s);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = 1;
    for (uint64_t a = 0; a < 64; ++a) {
      double u01 = (double)(orc_rand_u64(seed, stream, i * 64 + a) >> 11) * 0x1.0p-53;
      double u = hN + u01 * (hx1 - hN);
      double x = zipf_Hinv(u, s);
      double kd = floor(x + 0.5);
      if (kd < 1.0) kd = 1.0;
      if (kd > N) kd = N;
      k = (uint64_t)kd;
      if (kd - x <= sc || u >= zipf_H(kd + 0.5, s) - zipf_h(kd, s)) break;
    }
    out[i] = k;
  }
  return 0;
}

/* Zipf keys over the keys of a permutation relation: rank r (1..N, from
 * orc_gen_zipf_ranks) becomes base + perm(r - 1), perm the permutation
 * gen_perm_keys uses for N keys with perm_seed -- so rank 1, the hottest key,
 * sits at a seeded position among R's keys. */
int orc_gen_zipf_keys_i32(int32_t *out, uint64_t n, uint64_t num_ranks, double s, uint64_t seed, uint64_t stream,
                          uint64_t perm_seed, int64_t base) {
  uint64_t *r = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  if (!r) return -1;
  int rc = orc_gen_zipf_ranks(r, n, num_ranks, s, seed, stream);
  for (uint64_t i = 0; rc == 0 && i < n; ++i) out[i] = (int32_t)(base + (int64_t)orc_feistel_perm(r[i] - 1, num_ranks, perm_seed));
  free(r);
  return rc;
}
int orc_gen_zipf_keys_i64(int64_t *out, uint64_t n, uint64_t num_ranks, double s, uint64_t seed, uint64_t stream,
                          uint64_t perm_seed, int64_t base) {
  uint64_t *r = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  if (!r) return -1;
  int rc = orc_gen_zipf_ranks(r, n, num_ranks, s, seed, stream);
  for (uint64_t i = 0; rc == 0 && i < n; ++i) out[i] = base + (int64_t)orc_feistel_perm(r[i] - 1, num_ranks, perm_seed);
  free(r);
  return rc;
}

/* ======================================================================= */
/* Partitioning                                                             */

/* gpu_radix_partition.h:121-127: the key is widened to 64 bits (i32 keys are
 * sign-extended), masked with (F-1) << ignore_bits and shifted down. */
uint32_t orc_key_to_partition_i32(int32_t key, uint32_t radix_bits, uint32_t ignore_bits) {
  uint64_t mask = ((uint64_t)((1u << radix_bits) - 1u)) << ignore_bits;
  return (uint32_t)(((uint64_t)(int64_t)key & mask) >> ignore_bits);
}
uint32_t orc_key_to_partition_i64(int64_t key, uint32_t radix_bits, uint32_t ignore_bits) {
  uint64_t mask = ((uint64_t)((1u << radix_bits) - 1u)) << ignore_bits;
  return (uint32_t)(((uint64_t)key & mask) >> ignore_bits);
}

/* partition_input_chunk.rs:24-35. */
int orc_input_chunk_size(uint64_t data_len, uint32_t num_chunks, uint32_t key_bytes,
                         uint64_t *chunk_len) {
  uint64_t align_mask = ~((uint64_t)(ORC_ALIGN_BYTES / key_bytes) - 1ULL);
  uint64_t c = ((data_len + num_chunks - 1) / num_chunks) & align_mask;
  *chunk_len = c;
  return (c >= 0xFFFFFFFFULL) ? -1 : 0;
}

static inline uint64_t chunk_begin(uint64_t n, uint32_t chunks, uint32_t c, uint64_t clen) {
  (void)n; (void)chunks;
  return clen * c;
}
static inline uint64_t chunk_length(uint64_t n, uint32_t chunks, uint32_t c, uint64_t clen) {
  return (c + 1 == chunks) ? n - clen * c : clen;
}

#define ORC_DEFINE_PARTITION(SUF, K, T)                                                    \
  void orc_chunked_prefix_sum_##SUF(const K *keys, uint64_t n, uint32_t chunks,            \
                                    uint32_t radix_bits, uint32_t ignore_bits,             \
                                    uint64_t *offsets) {                                   \
    const uint32_t F = 1u << radix_bits;                                                   \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                    \
    uint64_t clen;                                                                         \
    orc_input_chunk_size(n, chunks, sizeof(K), &clen);                                     \
    uint64_t *hist = (uint64_t *)malloc(sizeof(uint64_t) * F);                             \
    for (uint32_t c = 0; c < chunks; ++c) {                                                \
      memset(hist, 0, sizeof(uint64_t) * F);                                               \
      const K *ck = keys + chunk_begin(n, chunks, c, clen);                                \
      uint64_t len = chunk_length(n, chunks, c, clen);                                     \
      for (uint64_t i = 0; i < len; ++i)                                                   \
        hist[orc_key_to_partition_##SUF(ck[i], radix_bits, ignore_bits)] += 1;             \
      /* radix_partition.cpp:252-266: running offset starts after the first pad. */       \
      uint64_t off = (clen + (uint64_t)pad * F) * c + pad;                                 \
      for (uint32_t p = 0; p < F; ++p) {                                                   \
        offsets[(uint64_t)c * F + p] = off;                                                \
        off += hist[p] + pad;                                                              \
      }                                                                                    \
    }                                                                                      \
    free(hist);                                                                            \
  }                                                                                        \
                                                                                           \
  void orc_contiguous_prefix_sum_##SUF(const K *keys, uint64_t n, uint32_t chunks,         \
                                       uint32_t radix_bits, uint32_t ignore_bits,          \
                                       uint64_t *offsets, uint64_t *local_offsets) {       \
    const uint32_t F = 1u << radix_bits;                                                   \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                    \
    uint64_t clen;                                                                         \
    orc_input_chunk_size(n, chunks, sizeof(K), &clen);                                     \
    uint64_t *hist = (uint64_t *)calloc((size_t)chunks * F, sizeof(uint64_t));             \
    for (uint32_t c = 0; c < chunks; ++c) {                                                \
      const K *ck = keys + chunk_begin(n, chunks, c, clen);                                \
      uint64_t len = chunk_length(n, chunks, c, clen);                                     \
      for (uint64_t i = 0; i < len; ++i)                                                   \
        hist[(uint64_t)c * F + orc_key_to_partition_##SUF(ck[i], radix_bits, ignore_bits)] \
            += 1;                                                                          \
    }                                                                                      \
    /* radix_partition.cu:156-198: exclusive scan in partition-major order             */ \
    /* [F][chunks], plus (p+1)*pad, transposed back to [chunks][F]. */                    \
    uint64_t run = 0;                                                                      \
    for (uint32_t p = 0; p < F; ++p) {                                                     \
      for (uint32_t c = 0; c < chunks; ++c) {                                              \
        local_offsets[(uint64_t)c * F + p] = run + (uint64_t)(p + 1) * pad;                \
        run += hist[(uint64_t)c * F + p];                                                  \
      }                                                                                    \
    }                                                                                      \
    for (uint32_t p = 0; p < F; ++p) offsets[p] = local_offsets[p];                        \
    free(hist);                                                                            \
  }                                                                                        \
                                                                                           \
  /* radix_partition.cpp:409-446 (NC); chunk c uses offset row c. */                      \
  void orc_partition_##SUF(const K *keys, const K *pays, uint64_t n, uint32_t chunks,      \
                           uint32_t radix_bits, uint32_t ignore_bits,                      \
                           const uint64_t *chunk_offsets, T *relation) {                   \
    const uint32_t F = 1u << radix_bits;                                                   \
    uint64_t clen;                                                                         \
    orc_input_chunk_size(n, chunks, sizeof(K), &clen);                                     \
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * F);                              \
    for (uint32_t c = 0; c < chunks; ++c) {                                                \
      memcpy(cur, chunk_offsets + (uint64_t)c * F, sizeof(uint64_t) * F);                  \
      uint64_t b = chunk_begin(n, chunks, c, clen);                                        \
      uint64_t len = chunk_length(n, chunks, c, clen);                                     \
      for (uint64_t i = b; i < b + len; ++i) {                                             \
        uint32_t p = orc_key_to_partition_##SUF(keys[i], radix_bits, ignore_bits);         \
        T t;                                                                               \
        t.key = keys[i];                                                                   \
        t.value = pays[i];                                                                 \
        relation[cur[p]++] = t;                                                            \
      }                                                                                    \
    }                                                                                      \
    free(cur);                                                                             \
  }                                                                                        \
                                                                                           \
  /* radix_partition.cu:343-410: concatenate the chunk pieces of one partition  */        \
  /* in chunk order into columnar arrays; the very last piece ends at the      */         \
  /* padded relation length. */                                                           \
  uint64_t orc_transform_##SUF(const T *src, const uint64_t *src_offsets,                  \
                               uint32_t src_chunks, uint32_t src_radix_bits,               \
                               uint64_t src_padded_len, uint32_t pid, K *dst_key,          \
                               K *dst_pay) {                                               \
    const uint32_t F = 1u << src_radix_bits;                                               \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                    \
    uint64_t out = 0;                                                                      \
    for (uint32_t c = 0; c < src_chunks; ++c) {                                            \
      uint64_t begin = src_offsets[(uint64_t)c * F + pid];                                 \
      uint64_t end = (c + 1 == src_chunks && pid + 1 == F)                                 \
                         ? src_padded_len                                                  \
                         : src_offsets[(uint64_t)c * F + pid + 1] - pad;                   \
      for (uint64_t i = begin; i < end; ++i) {                                             \
        if (dst_key) dst_key[out] = src[i].key;                                            \
        if (dst_pay) dst_pay[out] = src[i].value;                                          \
        ++out;                                                                             \
      }                                                                                    \
    }                                                                                      \
    return out;                                                                            \
  }

ORC_DEFINE_PARTITION(i32, int32_t, orc_tuple32)
ORC_DEFINE_PARTITION(i64, int64_t, orc_tuple64)

/* partitioned_relation.rs:429-459. */
uint64_t orc_partition_len(const uint64_t *offsets, uint32_t chunks, uint32_t radix_bits,
                           uint64_t padded_len, uint32_t pad, uint32_t pid) {
  const uint32_t F = 1u << radix_bits;
  const uint64_t olen = (uint64_t)chunks * F;
  uint64_t len = 0;
  for (uint32_t c = 0; c < chunks; ++c) {
    uint64_t ofi = (uint64_t)c * F + pid;
    uint64_t begin = offsets[ofi];
    uint64_t end = (ofi + 1 < olen) ? offsets[ofi + 1] - pad : padded_len;
    len += end - begin;
  }
  return len;
}

/* ======================================================================= */
/* Join                                                                     */

/* radix_join.cu:58-98 (greedy, u32 arithmetic as in the reference). */
void orc_assign_tasks(const uint64_t *probe_offsets, uint32_t probe_len, uint32_t pad,
                      uint32_t radix_bits, uint32_t grid, uint32_t *ta) {
  const uint32_t F = 1u << radix_bits;
  const uint32_t probe_tuples = probe_len - pad * F;
  const uint32_t avg = (probe_tuples + grid - 1u) / grid;
  ta[0] = 0;
  uint32_t task_id = 1, task_size = 0;
  for (uint32_t p = 0; p < F && task_id < grid; ++p) {
    uint32_t upper = (p + 1 < F) ? (uint32_t)(probe_offsets[p + 1] - pad) : probe_len;
    uint32_t size = (uint32_t)(upper - probe_offsets[p]);
    task_size += size;
    if (task_size >= avg) {
      ta[task_id] = p + 1;
      task_size = 0;
      task_id += 1;
    }
  }
  for (uint32_t t = task_id; t <= grid; ++t) ta[t] = F;
}

/* gpu_common.h:143-212. */
int32_t orc_mult_shift_hash_i32(int32_t value, uint32_t log2_buckets) {
  uint32_t product = (uint32_t)value * 2654435769u;
  return (int32_t)(product >> (32u - log2_buckets));
}
int64_t orc_mult_shift_hash_i64(int64_t value, uint32_t log2_buckets) {
  uint64_t product = (uint64_t)value * 11400714819323198485ULL;
  return (int64_t)(product >> (64u - log2_buckets));
}
/* MurmurHash3's finalizers fmix32 / fmix64 (Appleby's public-domain
 * MurmurHash3.cpp), which gpu_common.h:181-212 applies to the key and then
 * masks to log2_buckets bits.  Exported whole so that tests can run the full
 * MurmurHash3_x86_32 / x64_128 on top of them against published vectors. */
uint32_t orc_fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}
uint64_t orc_fmix64(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
  return k;
}
int32_t orc_murmur3_hash_i32(int32_t value, uint32_t log2_buckets) {
  return (int32_t)(orc_fmix32((uint32_t)value) & ((1u << log2_buckets) - 1u));
}
int64_t orc_murmur3_hash_i64(int64_t value, uint32_t log2_buckets) {
  return (int64_t)(orc_fmix64((uint64_t)value) & ((1ULL << log2_buckets) - 1ULL));
}

static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

/* Per-partition equi-join of build[0..nb) and probe[0..np): every pair with
 * equal keys contributes (count 1, build payload, probe payload); this is the
 * result of radix_join.cu:199-311 (bucket chaining visits every matching build
 * entry).  Scratch: a linear-probing index over the build tuples. */
#define ORC_DEFINE_PART_JOIN(SUF, K, T)                                                  \
  typedef struct { uint32_t *slots; uint64_t cap; } part_scratch_##SUF;                  \
  static void part_join_##SUF(const T *build, uint64_t nb, const T *probe, uint64_t np,  \
                              part_scratch_##SUF *scr, orc_join_result *res) {           \
    if (nb == 0 || np == 0) return;                                                      \
    uint64_t cap = next_pow2(nb * 2);                                                    \
    if (cap > scr->cap) {                                                                \
      free(scr->slots);                                                                  \
      scr->slots = (uint32_t *)malloc(sizeof(uint32_t) * cap);                           \
      scr->cap = cap;                                                                    \
    }                                                                                    \
    uint32_t *slots = scr->slots;                                                        \
    const uint32_t lg = (uint32_t)__builtin_ctzll(cap);                                  \
    for (uint64_t i = 0; i < cap; ++i) slots[i] = 0xFFFFFFFFu;                           \
    for (uint64_t i = 0; i < nb; ++i) {                                                  \
      if (build[i].key == (K)-1) continue; /* null keys never join */                    \
      uint64_t h = (uint64_t)orc_mult_shift_hash_i64((int64_t)build[i].key, lg);         \
      while (slots[h] != 0xFFFFFFFFu) h = (h + 1) & (cap - 1);                           \
      slots[h] = (uint32_t)i;                                                            \
    }                                                                                    \
    uint64_t cnt = 0, sr = 0, ss = 0;                                                    \
    for (uint64_t j = 0; j < np; ++j) {                                                  \
      K key = probe[j].key;                                                              \
      if (key == (K)-1) continue;                                                        \
      uint64_t h = (uint64_t)orc_mult_shift_hash_i64((int64_t)key, lg);                  \
      while (slots[h] != 0xFFFFFFFFu) {                                                  \
        const T *b = &build[slots[h]];                                                   \
        if (b->key == key) {                                                             \
          cnt += 1;                                                                      \
          sr += (uint64_t)(int64_t)b->value;                                             \
          ss += (uint64_t)(int64_t)probe[j].value;                                       \
        }                                                                                \
        h = (h + 1) & (cap - 1);                                                         \
      }                                                                                  \
    }                                                                                    \
    res->count += cnt;                                                                   \
    res->sum_build += sr;                                                                \
    res->sum_probe += ss;                                                                \
  }                                                                                      \
                                                                                         \
  static inline uint64_t part_end_##SUF(const uint64_t *off, uint32_t F, uint32_t p,     \
                                        uint64_t padded_len, uint32_t pad) {             \
    return (p + 1 < F) ? off[p + 1] - pad : padded_len;                                  \
  }                                                                                      \
                                                                                         \
  int orc_partitioned_join_##SUF(const T *build, const uint64_t *boff, uint64_t bplen,   \
                                 const T *probe, const uint64_t *poff, uint64_t pplen,   \
                                 uint32_t radix_bits, uint32_t ignore_bits,              \
                                 uint64_t max_build, orc_join_result *res) {             \
    (void)ignore_bits;                                                                   \
    const uint32_t F = 1u << radix_bits;                                                 \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                  \
    part_scratch_##SUF scr = {0, 0};                                                     \
    memset(res, 0, sizeof(*res));                                                        \
    int rc = 0;                                                                          \
    for (uint32_t p = 0; p < F; ++p) {                                                   \
      uint64_t b0 = boff[p], b1 = part_end_##SUF(boff, F, p, bplen, pad);                \
      uint64_t p0 = poff[p], p1 = part_end_##SUF(poff, F, p, pplen, pad);                \
      if (max_build && b1 - b0 > max_build) rc = -1;                                     \
      part_join_##SUF(build + b0, b1 - b0, probe + p0, p1 - p0, &scr, res);              \
    }                                                                                    \
    free(scr.slots);                                                                     \
    return rc;                                                                           \
  }                                                                                      \
                                                                                         \
  /* radix_join.cu:103-194: ht[key >> ignore] = (key, value); a probe tuple        */   \
  /* matches when the slot holds its key.  Later build tuples overwrite earlier.   */   \
  int orc_partitioned_join_perfect_##SUF(const T *build, const uint64_t *boff,           \
                                         uint64_t bplen, const T *probe,                 \
                                         const uint64_t *poff, uint64_t pplen,           \
                                         uint32_t radix_bits, uint32_t ignore_bits,      \
                                         uint64_t ht_entries, orc_join_result *res) {    \
    const uint32_t F = 1u << radix_bits;                                                 \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                  \
    T *ht = (T *)malloc(sizeof(T) * ht_entries);                                         \
    memset(res, 0, sizeof(*res));                                                        \
    int rc = 0;                                                                          \
    for (uint32_t p = 0; p < F; ++p) {                                                   \
      for (uint64_t i = 0; i < ht_entries; ++i) {                                        \
        ht[i].key = (K)-1;                                                               \
        ht[i].value = 0;                                                                 \
      }                                                                                  \
      uint64_t b0 = boff[p], b1 = part_end_##SUF(boff, F, p, bplen, pad);                \
      uint64_t p0 = poff[p], p1 = part_end_##SUF(poff, F, p, pplen, pad);                \
      if (b1 - b0 > ht_entries) rc = -1;                                                 \
      for (uint64_t i = b0; i < b1; ++i) {                                               \
        if (build[i].key == (K)-1) continue;                                             \
        uint64_t idx = (uint32_t)((uint64_t)(int64_t)build[i].key >> ignore_bits);       \
        if (idx < ht_entries) ht[idx] = build[i];                                        \
        else rc = -1;                                                                    \
      }                                                                                  \
      for (uint64_t j = p0; j < p1; ++j) {                                               \
        uint64_t idx = (uint32_t)((uint64_t)(int64_t)probe[j].key >> ignore_bits);       \
        if (probe[j].key != (K)-1 && idx < ht_entries && ht[idx].key == probe[j].key) {  \
          res->count += 1;                                                               \
          res->sum_build += (uint64_t)(int64_t)ht[idx].value;                            \
          res->sum_probe += (uint64_t)(int64_t)probe[j].value;                           \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
    free(ht);                                                                            \
    return rc;                                                                           \
  }                                                                                      \
                                                                                         \
  /* no_partitioning_join.cpp:27-163: mult-shift hash into a power-of-two table, */     \
  /* linear probing; the probe continues past a match so duplicates all count. */       \
  void orc_np_build_lp_##SUF(T *table, uint64_t entries, const K *keys, const K *pays,   \
                             uint64_t n) {                                               \
    const uint32_t lg = 63u - (uint32_t)__builtin_clzll(entries);                        \
    const uint64_t mask = (1ULL << lg) - 1ULL;                                           \
    for (uint64_t i = 0; i < n; ++i) {                                                   \
      if (keys[i] == (K)-1) continue; /* null keys never join */                         \
      uint64_t idx = (uint64_t)orc_mult_shift_hash_##SUF(keys[i], lg);                   \
      for (uint64_t a = 0; a <= mask; ++a, idx = (idx + 1) & mask) {                     \
        if (table[idx].key == (K)-1) {                                                   \
          table[idx].key = keys[i];                                                      \
          table[idx].value = pays[i];                                                    \
          break;                                                                         \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
  }                                                                                      \
  void orc_np_probe_lp_##SUF(const T *table, uint64_t entries, const K *keys,            \
                             const K *pays, uint64_t n, orc_join_result *res) {          \
    const uint32_t lg = 63u - (uint32_t)__builtin_clzll(entries);                        \
    const uint64_t mask = (1ULL << lg) - 1ULL;                                           \
    memset(res, 0, sizeof(*res));                                                        \
    for (uint64_t i = 0; i < n; ++i) {                                                   \
      if (keys[i] == (K)-1) continue; /* the reference's probe never ends on it */       \
      uint64_t idx = (uint64_t)orc_mult_shift_hash_##SUF(keys[i], lg);                   \
      for (uint64_t a = 0; a <= mask; ++a, idx = (idx + 1) & mask) {                     \
        if (table[idx].key == keys[i]) {                                                 \
          res->count += 1;                                                               \
          res->sum_build += (uint64_t)(int64_t)table[idx].value;                         \
          res->sum_probe += (uint64_t)(int64_t)pays[i];                                  \
        } else if (table[idx].key == (K)-1) {                                            \
          break;                                                                         \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
  }

ORC_DEFINE_PART_JOIN(i32, int32_t, orc_tuple32)
ORC_DEFINE_PART_JOIN(i64, int64_t, orc_tuple64)

/* ======================================================================= */
/* End-to-end multi-threaded radix join (the CPU baseline).  Pass 1 follows */
/* the reference's chunk-per-thread CPU partitioner (radix_partition.cpp)    */
/* extended with ignore_bits; pass 2 and the join run per pass-1 partition, */
/* taken dynamically by threads.                                             */

#define ORC_DEFINE_RADIX_JOIN(SUF, K, T)                                                 \
  typedef struct {                                                                       \
    const K *rk, *rp, *sk, *sp;                                                          \
    uint64_t rn, sn, rclen, sclen;                                                       \
    uint32_t bits1, bits2, threads;                                                      \
    uint64_t *roff, *soff; /* [threads][F1] */                                           \
    T *r1, *s1;                                                                          \
    uint64_t r1len, s1len;                                                               \
    atomic_uint next_part;                                                               \
    orc_join_result *partial; /* [threads] */                                            \
    pthread_barrier_t bar;                                                               \
  } rj_ctx_##SUF;                                                                        \
  typedef struct { rj_ctx_##SUF *ctx; uint32_t tid; } rj_arg_##SUF;                      \
                                                                                         \
  static void rj_hist_##SUF(const K *k, uint64_t len, uint32_t bits, uint32_t ign,       \
                            uint64_t *h) {                                               \
    const uint32_t F = 1u << bits;                                                       \
    memset(h, 0, sizeof(uint64_t) * F);                                                  \
    for (uint64_t i = 0; i < len; ++i) h[orc_key_to_partition_##SUF(k[i], bits, ign)]++; \
  }                                                                                      \
  /* gather pass-1 partition p of a chunked relation into columnar arrays */            \
  static uint64_t rj_gather_##SUF(const T *rel, const uint64_t *off, uint32_t chunks,    \
                                  uint32_t bits, uint64_t plen, uint32_t p, K *dk,       \
                                  K *dp) {                                               \
    return orc_transform_##SUF(rel, off, chunks, bits, plen, p, dk, dp);                 \
  }                                                                                      \
                                                                                         \
  static void *rj_worker_##SUF(void *varg) {                                             \
    rj_arg_##SUF *a = (rj_arg_##SUF *)varg;                                              \
    rj_ctx_##SUF *c = a->ctx;                                                            \
    const uint32_t t = a->tid, T_ = c->threads;                                          \
    const uint32_t F1 = 1u << c->bits1;                                                  \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                  \
    for (int rel = 0; rel < 2; ++rel) {                                                  \
      const K *k = rel ? c->sk : c->rk;                                                  \
      const K *py = rel ? c->sp : c->rp;                                                 \
      uint64_t n = rel ? c->sn : c->rn, clen = rel ? c->sclen : c->rclen;                \
      uint64_t *off = (rel ? c->soff : c->roff) + (uint64_t)t * F1;                      \
      T *out = rel ? c->s1 : c->r1;                                                      \
      uint64_t b = clen * t, len = (t + 1 == T_) ? n - clen * t : clen;                  \
      uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * F1);                           \
      rj_hist_##SUF(k + b, len, c->bits1, 0, h);                                         \
      uint64_t o = (clen + (uint64_t)pad * F1) * t + pad;                                \
      for (uint32_t p = 0; p < F1; ++p) {                                                \
        off[p] = o;                                                                      \
        o += h[p] + pad;                                                                 \
      }                                                                                  \
      memcpy(h, off, sizeof(uint64_t) * F1);                                             \
      for (uint64_t i = b; i < b + len; ++i) {                                           \
        uint32_t p = orc_key_to_partition_##SUF(k[i], c->bits1, 0);                      \
        T tu;                                                                            \
        tu.key = k[i];                                                                   \
        tu.value = py[i];                                                                \
        out[h[p]++] = tu;                                                                \
      }                                                                                  \
      free(h);                                                                           \
    }                                                                                    \
    pthread_barrier_wait(&c->bar);                                                       \
    part_scratch_##SUF scr = {0, 0};                                                     \
    orc_join_result res = {0, 0, 0};                                                     \
    K *bk = 0, *bp = 0, *pk = 0, *pp = 0;                                                \
    T *b2 = 0, *p2 = 0;                                                                  \
    uint64_t bcap = 0, pcap = 0;                                                         \
    const uint32_t F2 = 1u << c->bits2;                                                  \
    uint64_t *h2 = (uint64_t *)malloc(sizeof(uint64_t) * F2 * 4);                        \
    for (;;) {                                                                           \
      uint32_t p = atomic_fetch_add(&c->next_part, 1u);                                  \
      if (p >= F1) break;                                                                \
      uint64_t nb = orc_partition_len(c->roff, T_, c->bits1, c->r1len, pad, p);          \
      uint64_t np_ = orc_partition_len(c->soff, T_, c->bits1, c->s1len, pad, p);         \
      if (nb > bcap) {                                                                   \
        bcap = nb * 2;                                                                   \
        bk = (K *)realloc(bk, sizeof(K) * bcap);                                         \
        bp = (K *)realloc(bp, sizeof(K) * bcap);                                         \
        b2 = (T *)realloc(b2, sizeof(T) * (bcap + (uint64_t)F2 * pad));                  \
      }                                                                                  \
      if (np_ > pcap) {                                                                  \
        pcap = np_ * 2;                                                                  \
        pk = (K *)realloc(pk, sizeof(K) * pcap);                                         \
        pp = (K *)realloc(pp, sizeof(K) * pcap);                                         \
        p2 = (T *)realloc(p2, sizeof(T) * (pcap + (uint64_t)F2 * pad));                  \
      }                                                                                  \
      rj_gather_##SUF(c->r1, c->roff, T_, c->bits1, c->r1len, p, bk, bp);                \
      rj_gather_##SUF(c->s1, c->soff, T_, c->bits1, c->s1len, p, pk, pp);                \
      if (c->bits2 == 0) {                                                               \
        /* single pass: join the gathered partition directly */                         \
        T *bt = b2, *pt = p2;                                                            \
        for (uint64_t i = 0; i < nb; ++i) { bt[i].key = bk[i]; bt[i].value = bp[i]; }    \
        for (uint64_t i = 0; i < np_; ++i) { pt[i].key = pk[i]; pt[i].value = pp[i]; }   \
        part_join_##SUF(bt, nb, pt, np_, &scr, &res);                                    \
        continue;                                                                        \
      }                                                                                  \
      uint64_t *boff2 = h2, *poff2 = h2 + F2, *cur = h2 + 2 * F2;                        \
      orc_contiguous_prefix_sum_##SUF(bk, nb, 1, c->bits2, c->bits1, boff2, cur);        \
      orc_partition_##SUF(bk, bp, nb, 1, c->bits2, c->bits1, cur, b2);                   \
      orc_contiguous_prefix_sum_##SUF(pk, np_, 1, c->bits2, c->bits1, poff2, cur);       \
      orc_partition_##SUF(pk, pp, np_, 1, c->bits2, c->bits1, cur, p2);                  \
      for (uint32_t q = 0; q < F2; ++q) {                                                \
        uint64_t b0 = boff2[q], b1 = (q + 1 < F2) ? boff2[q + 1] - pad                   \
                                                  : nb + (uint64_t)F2 * pad;             \
        uint64_t p0 = poff2[q], p1 = (q + 1 < F2) ? poff2[q + 1] - pad                   \
                                                  : np_ + (uint64_t)F2 * pad;            \
        part_join_##SUF(b2 + b0, b1 - b0, p2 + p0, p1 - p0, &scr, &res);                 \
      }                                                                                  \
    }                                                                                    \
    c->partial[t] = res;                                                                 \
    free(h2); free(bk); free(bp); free(pk); free(pp); free(b2); free(p2);                \
    free(scr.slots);                                                                     \
    return 0;                                                                            \
  }                                                                                      \
                                                                                         \
  int orc_radix_join_##SUF(const K *r_key, const K *r_pay, uint64_t r_n, const K *s_key, \
                           const K *s_pay, uint64_t s_n, uint32_t bits1, uint32_t bits2, \
                           uint32_t threads, orc_join_result *out) {                     \
    if (threads == 0) threads = 1;                                                       \
    rj_ctx_##SUF c;                                                                      \
    memset(&c, 0, sizeof(c));                                                            \
    c.rk = r_key; c.rp = r_pay; c.sk = s_key; c.sp = s_pay;                              \
    c.rn = r_n; c.sn = s_n; c.bits1 = bits1; c.bits2 = bits2; c.threads = threads;       \
    if (orc_input_chunk_size(r_n, threads, sizeof(K), &c.rclen)) return -1;              \
    if (orc_input_chunk_size(s_n, threads, sizeof(K), &c.sclen)) return -1;              \
    const uint32_t F1 = 1u << bits1;                                                     \
    const uint32_t pad = ORC_PADDING_BYTES / sizeof(T);                                  \
    c.r1len = r_n + (uint64_t)F1 * threads * pad;                                        \
    c.s1len = s_n + (uint64_t)F1 * threads * pad;                                        \
    c.roff = (uint64_t *)malloc(sizeof(uint64_t) * F1 * threads);                        \
    c.soff = (uint64_t *)malloc(sizeof(uint64_t) * F1 * threads);                        \
    c.r1 = (T *)malloc(sizeof(T) * c.r1len);                                             \
    c.s1 = (T *)malloc(sizeof(T) * c.s1len);                                             \
    c.partial = (orc_join_result *)calloc(threads, sizeof(orc_join_result));             \
    if (!c.roff || !c.soff || !c.r1 || !c.s1 || !c.partial) return -2;                   \
    atomic_init(&c.next_part, 0u);                                                       \
    pthread_barrier_init(&c.bar, 0, threads);                                            \
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);                    \
    rj_arg_##SUF *args = (rj_arg_##SUF *)malloc(sizeof(rj_arg_##SUF) * threads);         \
    for (uint32_t t = 0; t < threads; ++t) {                                             \
      args[t].ctx = &c;                                                                  \
      args[t].tid = t;                                                                   \
      pthread_create(&th[t], 0, rj_worker_##SUF, &args[t]);                              \
    }                                                                                    \
    memset(out, 0, sizeof(*out));                                                        \
    for (uint32_t t = 0; t < threads; ++t) {                                             \
      pthread_join(th[t], 0);                                                            \
      out->count += c.partial[t].count;                                                  \
      out->sum_build += c.partial[t].sum_build;                                          \
      out->sum_probe += c.partial[t].sum_probe;                                          \
    }                                                                                    \
    pthread_barrier_destroy(&c.bar);                                                     \
    free(th); free(args); free(c.roff); free(c.soff); free(c.r1); free(c.s1);            \
    free(c.partial);                                                                     \
    return 0;                                                                            \
  }

ORC_DEFINE_RADIX_JOIN(i32, int32_t, orc_tuple32)
ORC_DEFINE_RADIX_JOIN(i64, int64_t, orc_tuple64)
