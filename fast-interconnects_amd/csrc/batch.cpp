// Batched execution of many independent joins (fij_join_batch_*).
//
// The reference joins one pair of relations per run of the radix-join binary
// (radix-join/src/execution_methods/gpu_radix_join.rs:90-565), allocating and
// launching per join.  A batch joins n pairs with one submission:
//   * joins whose build side fits one block's LDS table ("small") go to one
//     persistent launch that joins them without partitioning (batch.hip);
//   * every other join runs the two-pass radix pipeline with the radix bits
//     of its size class (fij_join_autotune); all joins of a class go
//     through each pass together, one launch per pass: join j's pass-1
//     partitions are partitions j * F1 .. j * F1 + F1 - 1 of one batched
//     plan (pipeline.hpp), the histogram, offsets and scatter launches map
//     their blocks to (join, side, chunk), pass 2 and the join kernels run
//     over all partitions and add each join's count and sums to its result;
//     the class's buffers are pooled and grown when a submission needs more;
//   * the whole submission (the small-join launch and every class's
//     launches) is recorded into one HIP graph and launched once; results
//     are written to device slots and copied back with one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/fij.h"
#include "batch.hpp"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

using namespace fij;

namespace {

fij_status batch_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(s, buf);
}

#define BATCH_HIP(expr)                                                                                      \
  do {                                                                                                       \
    hipError_t e_ = (expr);                                                                                  \
    if (e_ != hipSuccess)                                                                                    \
      return batch_fail(e_ == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, "%s: %s", #expr, \
                        hipGetErrorString(e_));                                                              \
  } while (0)

uint32_t ceil_log2(uint64_t x) { return x <= 1 ? 0 : 64u - static_cast<uint32_t>(__builtin_clzll(x - 1)); }

// Buffers of one size class's batched plan and the capacities they were
// allocated for.
struct ClassArena {
  std::vector<void*> allocs;
  PipelineDevice d{};
  HistParams* hist = nullptr;      // [2 * joins]
  BatchScanJob* scan = nullptr;    // [2 * joins]
  ScatterParams* scat = nullptr;   // [2 * joins]
  BatchSide* sides = nullptr;      // [2 * joins]
  uint32_t* prefix = nullptr;      // [2 * joins + 1] first chunk of each join side
  uint32_t* job_slot = nullptr;    // [joins]
  uint32_t* side_tiles = nullptr;  // [2 * joins] plan scratch: tiles of each join side
  uint32_t* counts = nullptr;      // per-chunk histograms of every join side
  uint64_t* local = nullptr;       // per-chunk write offsets of every join side
  // capacities
  uint64_t joins = 0, rel1[2] = {0, 0}, rel2[2] = {0, 0}, tiles[2] = {0, 0}, chunk_words = 0, units = 0;
  // this submission: joins, chunks of all join sides, pass-2 tiles per side
  uint32_t sub_joins = 0, sub_blocks = 0, sub_tiles[2] = {0, 0};
};

// Streams a submission's independent size classes are spread over.
constexpr uint32_t kMaxBatchLanes = 4;

// A size class's configuration as batches run it: radix bits, pass-1 line
// ring cap (0 = deepest) and pass-1 block shape.
struct ClassKey {
  uint32_t b1, b2, ring, block;
  bool operator<(const ClassKey& o) const {
    return std::tie(b1, b2, ring, block) < std::tie(o.b1, o.b2, o.ring, o.block);
  }
};

// One join side's place in a class arena.
struct SideLayout {
  uint64_t n;
  uint32_t chunks;
  uint64_t counts_at, local_at, rel1_at, rel2_at;
};

}  // namespace

struct fij_join_batch {
  fij_join_batch_config cfg;
  KeyType kt;
  uint32_t cus;
  uint32_t lds_total;
  SmallJob* d_jobs = nullptr;
  uint32_t* d_ticket = nullptr;
  JoinAggregate* d_results = nullptr;
  SmallJob* h_jobs = nullptr;
  JoinAggregate* h_results = nullptr;
  hipStream_t own = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::map<ClassKey, ClassArena> arenas;
  // parallel lanes: lane 0 is the submission stream, lanes 1.. their own
  // streams, forked from it and joined back by events
  uint32_t lanes = 1;
  hipStream_t lane[kMaxBatchLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxBatchLanes] = {};
};

namespace {

void free_arena(ClassArena& a) {
  for (void* p : a.allocs) (void)hipFree(p);
  a = ClassArena{};
}

template <typename T>
hipError_t arena_alloc(ClassArena& a, T** out, uint64_t count) {
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, std::max<uint64_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  a.allocs.push_back(p);
  *out = static_cast<T*>(p);
  return hipSuccess;
}

// Chunks of a batched join side: about 32768 tuples each, at most one per
// compute unit, and never a chunk length that leaves the last chunk empty.
uint32_t side_chunks(fij_type type, uint64_t n, uint32_t cus) {
  uint32_t g = static_cast<uint32_t>(std::min<uint64_t>(cus, std::max<uint64_t>(1, (n + 32767) / 32768)));
  auto fits = [&](uint32_t chunks) {
    uint64_t c = 0;
    return fij_input_chunk_size(type, n, chunks, &c) == FIJ_OK && c * (chunks - 1) <= n;
  };
  while (g > 1 && !fits(g)) --g;
  return g;
}

// Grows the class arena to hold `joins` joins with these totals; every
// buffer is allocated anew when any capacity is short (the old contents are
// not needed between submissions).
fij_status grow_arena(ClassArena& a, KeyType kt, uint32_t b1, uint32_t b2, uint64_t joins, const uint64_t (&rel1)[2],
                      const uint64_t (&rel2)[2], const uint64_t (&tiles)[2], uint64_t chunk_words, uint64_t units,
                      hipStream_t st) {
  if (a.joins >= joins && a.rel1[0] >= rel1[0] && a.rel1[1] >= rel1[1] && a.rel2[0] >= rel2[0] &&
      a.rel2[1] >= rel2[1] && a.tiles[0] >= tiles[0] && a.tiles[1] >= tiles[1] && a.chunk_words >= chunk_words &&
      a.units >= units)
    return FIJ_OK;
  BATCH_HIP(hipStreamSynchronize(st));
  ClassArena n{};
  // room for growth: a quarter more than this submission needs
  auto more = [](uint64_t x) { return x + x / 4; };
  n.joins = more(std::max(a.joins, joins));
  for (int s = 0; s < 2; ++s) {
    n.rel1[s] = more(std::max(a.rel1[s], rel1[s]));
    n.rel2[s] = more(std::max(a.rel2[s], rel2[s]));
    n.tiles[s] = more(std::max(a.tiles[s], tiles[s]));
  }
  n.chunk_words = more(std::max(a.chunk_words, chunk_words));
  n.units = more(std::max(a.units, units));
  free_arena(a);
  const uint32_t F1 = 1u << b1, F2 = 1u << b2;
  const uint64_t parts = n.joins * F1;
  const uint64_t tb = tuple_bytes(kt);
  PipelineDevice& d = n.d;
  hipError_t e = hipSuccess;
  for (int s = 0; s < 2 && e == hipSuccess; ++s) {
    PipelineSide& ps = s ? d.probe : d.build;
    uint8_t *r1 = nullptr, *r2 = nullptr;
    if ((e = arena_alloc(n, &ps.off1, parts)) != hipSuccess || (e = arena_alloc(n, &ps.plen, parts)) != hipSuccess ||
        (e = arena_alloc(n, &ps.pstart, parts + 1)) != hipSuccess ||
        (e = arena_alloc(n, &ps.tile_start, parts + 1)) != hipSuccess ||
        // one spare tuple after rel1 (pass 2 reads 8-byte tuples in aligned pairs)
        (e = arena_alloc(n, &r1, (n.rel1[s] + 1) * tb)) != hipSuccess ||
        (e = arena_alloc(n, &r2, n.rel2[s] * tb)) != hipSuccess ||
        (e = arena_alloc(n, &ps.meta, (n.tiles[s] + 1) * F2)) != hipSuccess)
      break;
    ps.rel1 = r1;
    ps.rel2 = r2;
  }
  if (e == hipSuccess) e = arena_alloc(n, &n.counts, n.chunk_words);
  if (e == hipSuccess) e = arena_alloc(n, &n.local, n.chunk_words);
  if (e == hipSuccess) e = arena_alloc(n, &d.work_counter, kJoinCounterWords);
  if (e == hipSuccess) e = arena_alloc(n, &d.overflow, parts * F2);
  if (e == hipSuccess) e = arena_alloc(n, &d.units, n.units);
  d.unit_cap = static_cast<uint32_t>(n.units);
  if (e == hipSuccess) e = arena_alloc(n, &d.aggregates, 1);
  if (e == hipSuccess) e = arena_alloc(n, &n.hist, 2 * n.joins);
  if (e == hipSuccess) e = arena_alloc(n, &n.scan, 2 * n.joins);
  if (e == hipSuccess) e = arena_alloc(n, &n.scat, 2 * n.joins);
  if (e == hipSuccess) e = arena_alloc(n, &n.sides, 2 * n.joins);
  if (e == hipSuccess) e = arena_alloc(n, &n.prefix, 2 * n.joins + 1);
  if (e == hipSuccess) e = arena_alloc(n, &n.job_slot, n.joins);
  if (e == hipSuccess) e = arena_alloc(n, &n.side_tiles, 2 * n.joins);
  // the deferred-item flags start at zero; every join leaves them at zero
  if (e == hipSuccess) e = hipMemset(d.overflow, 0, parts * F2);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    free_arena(n);
    return batch_fail(e == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, "batched plan buffers: %s",
                      hipGetErrorString(e));
  }
  a = n;
  return FIJ_OK;
}

// Records one size class's passes over its joins (join indices `ids`).
fij_status enqueue_class(fij_join_batch* b, ClassArena& a, const ClassKey& ck, const std::vector<uint32_t>& ids,
                         const fij_join_desc* joins, hipStream_t st, uint32_t lds_total, std::vector<uint8_t>& host,
                         double* grow_ms) {
  const uint32_t b1 = ck.b1, b2 = ck.b2;
  const KeyType kt = b->kt;
  const fij_type type = static_cast<fij_type>(b->cfg.type);
  const uint32_t F1 = 1u << b1, F2 = 1u << b2;
  const uint32_t pad = pad_tuples(kt);
  const uint32_t L = 128 / tuple_bytes(kt);
  const uint32_t tile2 = pipeline_tile2(kt);
  const uint32_t J = static_cast<uint32_t>(ids.size());
  std::vector<SideLayout> lay(2 * J);
  uint64_t rel1[2] = {0, 0}, rel2[2] = {0, 0}, tiles[2] = {0, 0}, words = 0, probe_total = 0;
  uint32_t blocks = 0;
  for (uint32_t j = 0; j < J; ++j) {
    const fij_join_desc& jd = joins[ids[j]];
    for (int s = 0; s < 2; ++s) {
      SideLayout& l = lay[2 * j + s];
      l.n = s ? jd.probe_len : jd.build_len;
      l.chunks = side_chunks(type, l.n, b->cus);
      l.counts_at = words;
      l.local_at = words;
      words += static_cast<uint64_t>(F1) * l.chunks;
      l.rel1_at = rel1[s];
      rel1[s] += (l.n + static_cast<uint64_t>(F1) * pad + L - 1) / L * L;  // regions start on 128-byte lines
      l.rel2_at = rel2[s];
      if (b2) {  // single-pass classes join from rel1
        rel2[s] += l.n;
        tiles[s] += (l.n + tile2 - 1) / tile2 + F1;  // at most
      }
      blocks += l.chunks;
    }
    probe_total += jd.probe_len;
  }
  if (static_cast<uint64_t>(J) * F1 * F2 >= 0xFFFFFFFFull || tiles[0] >= 0xFFFFFFFFull || tiles[1] >= 0xFFFFFFFFull)
    return batch_fail(FIJ_INTEGER_OVERFLOW, "size class %u+%u: too many joins for one submission", b1, b2);
  const uint64_t units = static_cast<uint64_t>(J) * F1 * F2 + probe_total / 65536 + J + 1;
  const auto t_grow = std::chrono::steady_clock::now();
  fij_status s = grow_arena(a, kt, b1, b2, J, rel1, rel2, tiles, words, units, st);
  *grow_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_grow).count();
  if (s != FIJ_OK) return s;
  // descriptors
  const size_t n_hist = 2ull * J * sizeof(HistParams), n_scan = 2ull * J * sizeof(BatchScanJob),
               n_scat = 2ull * J * sizeof(ScatterParams), n_side = 2ull * J * sizeof(BatchSide),
               n_pre = (2ull * J + 1) * 4, n_slot = static_cast<size_t>(J) * 4;
  host.assign(n_hist + n_scan + n_scat + n_side + n_pre + n_slot, 0);
  HistParams* hh = reinterpret_cast<HistParams*>(host.data());
  BatchScanJob* hs = reinterpret_cast<BatchScanJob*>(host.data() + n_hist);
  ScatterParams* hc = reinterpret_cast<ScatterParams*>(host.data() + n_hist + n_scan);
  BatchSide* hb = reinterpret_cast<BatchSide*>(host.data() + n_hist + n_scan + n_scat);
  uint32_t* hp = reinterpret_cast<uint32_t*>(host.data() + n_hist + n_scan + n_scat + n_side);
  uint32_t* hj = reinterpret_cast<uint32_t*>(host.data() + n_hist + n_scan + n_scat + n_side + n_pre);
  const uint64_t tb = tuple_bytes(kt);
  uint32_t at = 0;
  for (uint32_t j = 0; j < J; ++j) {
    const fij_join_desc& jd = joins[ids[j]];
    hj[j] = ids[j];
    for (int side = 0; side < 2; ++side) {
      const uint32_t js = 2 * j + side;
      const SideLayout& l = lay[js];
      PipelineSide& ps = side ? a.d.probe : a.d.build;
      const void* keys = side ? jd.probe_keys : jd.build_keys;
      const void* pays = side ? jd.probe_payloads : jd.build_payloads;
      hh[js] = hist_params(kt, keys, l.n, l.chunks, b1, ck.block, a.counts + l.counts_at);
      hs[js] = BatchScanJob{a.counts + l.counts_at, a.local + l.local_at, ps.off1 + static_cast<uint64_t>(j) * F1,
                            l.rel1_at, F1, l.chunks, pad};
      void* region = static_cast<uint8_t*>(ps.rel1) + l.rel1_at * tb;
      const hipError_t e = scatter_params_batched(kt, keys, pays, l.n, l.chunks, b1, a.local + l.local_at, region,
                                                  lds_total, ck.block, ck.ring, &hc[js]);
      if (e != hipSuccess) return batch_fail(FIJ_UNSUPPORTED, "size class %u+%u: pass-1 buffers do not fit LDS", b1, b2);
      hb[js] = BatchSide{l.rel1_at + l.n + static_cast<uint64_t>(F1) * pad, l.rel2_at, j * F1, 0};
      hp[js] = at;
      at += l.chunks;
    }
  }
  hp[2 * J] = at;
  BATCH_HIP(hipMemcpyAsync(a.hist, hh, n_hist, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipMemcpyAsync(a.scan, hs, n_scan, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipMemcpyAsync(a.scat, hc, n_scat, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipMemcpyAsync(a.sides, hb, n_side, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipMemcpyAsync(a.prefix, hp, n_pre, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipMemcpyAsync(a.job_slot, hj, n_slot, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipStreamSynchronize(st));
  // the plan's device view
  PipelineDevice& d = a.d;
  d.F1 = J * F1;
  d.F2 = F2;
  d.G1 = 1;
  d.bits1 = b1;
  d.bits2 = b2;
  d.pad = pad;
  d.tile2 = tile2;
  d.sides = 3;
  d.agg_len = 0;
  d.job_results = b->d_results;
  d.job_slot = a.job_slot;
  d.job_shift = b1;
  d.build.n = rel2[0];
  d.probe.n = rel2[1];
  d.big_block_tables = 0;
  if (b2 == 0)
    for (uint32_t j = 0; j < J; ++j) d.big_block_tables |= big_block_tables_for(kt, lay[2 * j].n, F1);
  a.sub_joins = J;
  a.sub_blocks = blocks;
  a.sub_tiles[0] = static_cast<uint32_t>(tiles[0]);
  a.sub_tiles[1] = static_cast<uint32_t>(tiles[1]);
  return FIJ_OK;
}

// The class's launches (after enqueue_class prepared its descriptors): one
// per pass over all its joins.
hipError_t launch_class(fij_join_batch* b, ClassArena& a, const ClassKey& ck, hipStream_t st, uint32_t lds_total) {
  const KeyType kt = b->kt;
  const uint32_t b1 = ck.b1;
  const uint32_t js = 2 * a.sub_joins;
  hipError_t e;
  if ((e = launch_hist_batched(kt, a.hist, a.prefix, js, a.sub_blocks, b1, ck.block, st)) != hipSuccess) return e;
  if ((e = launch_offsets_batched(a.scan, js, st)) != hipSuccess) return e;
  if ((e = launch_partition_batched(kt, a.scat, a.prefix, js, a.sub_blocks, b1, lds_total, ck.block, ck.ring, st)) !=
      hipSuccess)
    return e;
  if ((e = launch_pipeline_plan_batched(a.d, a.sides, a.sub_joins, a.side_tiles, a.d.build.n, a.d.probe.n, st)) !=
      hipSuccess)
    return e;
  const bool one_pass = ck.b2 == 0;
  if (!one_pass) {
    const uint64_t tiles = static_cast<uint64_t>(a.sub_tiles[0]) + a.sub_tiles[1];
    const uint32_t sort_grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(b->cus, tiles)));
    if ((e = launch_pipeline_sort2(a.d, kt, sort_grid, st)) != hipSuccess) return e;
  }
  // the register join: four blocks per CU, as many as the lighter join shapes
  // of batched classes (fewer registers, smaller tables) keep resident; the
  // others start as earlier blocks finish and find the items taken.  A
  // single-pass class's items are whole pass-1 partitions (large tables);
  // most go on to the block-table join.
  if ((e = launch_pipeline_join(a.d, kt, one_pass, false, 4 * b->cus, st)) != hipSuccess) return e;
  return launch_pipeline_join(a.d, kt, one_pass, true, 2 * b->cus, st);
}

// The closed-form size-class rule (the built-in table's fallback).
fij_join_tuning class_rule(KeyType kt, uint64_t build_len, uint64_t probe_len) {
  fij_join_tuning t{};
  t.build_log2 = ceil_log2(build_len);
  t.probe_log2 = ceil_log2(probe_len);
  // one block joins a small pair: its build side fits one LDS table fill and
  // its probe side is short enough for one block to stream
  t.small = build_len <= small_join_table_tuples(kt) && probe_len <= (1u << 22) ? 1u : 0u;
  // Size classes of the radix plan: at most 512 build and 512 probe tuples
  // per (pass-1, pass-2) partition pair on average, so that pairs fit the
  // join's registers (at most 640 per side), split evenly over the passes
  // (pass 1 at most 9 bits: its LDS line buffers), and enough pass-1
  // partitions that a pair's probe side has at most 64 pass-2 tiles of 8192
  // tuples (int32 tiles hold 16,382: at most 33 of them).
  uint32_t b = std::max<uint32_t>(
      2, std::max(ceil_log2((build_len + 511) / 512), ceil_log2((probe_len + 511) / 512)));
  uint32_t b1 = std::min<uint32_t>(9, (b + 1) / 2);
  const uint32_t need = ceil_log2(probe_len) > 19 ? ceil_log2(probe_len) - 19 : 0;
  b1 = std::min<uint32_t>(9, std::max(b1, need));
  b = std::max(b, b1 + 1);
  t.radix_bits_1 = b1;
  t.radix_bits_2 = std::min<uint32_t>(12, b - b1);
  t.ring_lines = 0;
  t.block = 1024;
  t.from_table = 0;
  return t;
}

// Built-in tables: the configurations measured fastest per size class on
// MI355X (tuning_table.inc, written by fast-interconnects_amd/autotune.py).
// Each list ends with an entry of block 0.
#include "tuning_table.inc"

std::vector<fij_join_tuning> builtin_table(KeyType kt) {
  std::vector<fij_join_tuning> out;
  for (const fij_join_tuning* e = kt == KeyType::I32 ? kTunedI32 : kTunedI64; e->block; ++e) {
    out.push_back(*e);
    out.back().from_table = 1;
  }
  return out;
}

std::mutex g_tuning_mutex;
std::vector<fij_join_tuning> g_tuning[2];
bool g_tuning_loaded[2] = {false, false};
bool g_tuning_user[2] = {false, false};  // the table came from fij_join_tuning_set entries

// The configuration of a pair of these lengths: an entry of a table set by
// fij_join_tuning_set; else a single-pass plan where it applies; else the
// built-in table's entry of its class (the small path only where the pair
// qualifies for it); else the rule.
fij_join_tuning tuned_class(KeyType kt, uint64_t build_len, uint64_t probe_len) {
  const fij_join_tuning rule = class_rule(kt, build_len, probe_len);
  const int ti = kt == KeyType::I32 ? 0 : 1;
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  if (!g_tuning_loaded[ti]) {
    g_tuning[ti] = builtin_table(kt);
    g_tuning_loaded[ti] = true;
    g_tuning_user[ti] = false;
  }
  const fij_join_tuning* entry = nullptr;
  for (const fij_join_tuning& e : g_tuning[ti])
    if (e.build_log2 == rule.build_log2 && e.probe_log2 == rule.probe_log2) {
      entry = &e;
      break;
    }
  // Build sides of 2^18 tuples up to what 64 KiB block tables hold after one
  // pass of 9 bits: a single-pass plan, measured faster than the two-pass
  // classes there (1.2-1.45x for single int32 joins at 2^19-2^21, up to 2.9x
  // for long probe sides at int64; batched int32 joins of log-uniform sizes
  // 6 % faster); below 2^18 the two are level.  Going on to the 128 KiB
  // tables that single joins take up to twice that size (fij_join_plan_bits)
  // left the batched mix as fast (0.238 s either way on one box), so batched
  // classes stop here.
  if (!(entry && g_tuning_user[ti]) && !rule.small && build_len >= (1u << 18) &&
      block_tables_hold(kt, build_len, 1ull << kOnePassBits, false)) {
    fij_join_tuning t = rule;
    t.radix_bits_1 = kOnePassBits;
    t.radix_bits_2 = 0;
    t.ring_lines = 0;
    t.block = 1024;
    return t;
  }
  if (entry) {
    fij_join_tuning t = *entry;
    t.small = entry->small && rule.small;
    t.from_table = 1;
    return t;
  }
  return rule;
}

}  // namespace

extern "C" {

fij_status fij_join_autotune(fij_type type, uint64_t build_len, uint64_t probe_len, uint32_t* small,
                             uint32_t* radix_bits_1, uint32_t* radix_bits_2) {
  if (!small || !radix_bits_1 || !radix_bits_2) return batch_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  fij_join_tuning t{};
  const fij_status s = fij_join_tuning_lookup(type, build_len, probe_len, &t);
  if (s != FIJ_OK) return s;
  *small = t.small;
  *radix_bits_1 = t.radix_bits_1;
  *radix_bits_2 = t.radix_bits_2;
  return FIJ_OK;
}

fij_status fij_join_tuning_lookup(fij_type type, uint64_t build_len, uint64_t probe_len, fij_join_tuning* out) {
  if ((type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) || !out) return batch_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  *out = tuned_class(type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64, build_len, probe_len);
  return FIJ_OK;
}

fij_status fij_join_tuning_set(fij_type type, const fij_join_tuning* entries, uint32_t n, int use_builtin) {
  if ((type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) || (n && !entries))
    return batch_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  std::vector<fij_join_tuning> table;
  if (use_builtin) {
    table = builtin_table(type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64);
  } else {
    for (uint32_t i = 0; i < n; ++i) {
      const fij_join_tuning& e = entries[i];
      const bool pow2_ring = e.ring_lines == 0 || (e.ring_lines <= 256 && (e.ring_lines & (e.ring_lines - 1)) == 0);
      if (e.radix_bits_1 < 1 || e.radix_bits_1 > 9 || e.radix_bits_2 < 1 || e.radix_bits_2 > 12 || !pow2_ring ||
          (e.block != 512 && e.block != 1024) || e.small > 1 || e.build_log2 > 40 || e.probe_log2 > 40)
        return batch_fail(FIJ_INVALID_ARGUMENT,
                          "tuning entry %u: radix bits %u+%u (1..9 + 1..12), ring %u (0 or a power of two <= 256), "
                          "block %u (512 or 1024)",
                          i, e.radix_bits_1, e.radix_bits_2, e.ring_lines, e.block);
      table.push_back(e);
      table.back().from_table = 1;
    }
  }
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_tuning[type == FIJ_TYPE_I32 ? 0 : 1] = std::move(table);
  g_tuning_loaded[type == FIJ_TYPE_I32 ? 0 : 1] = true;
  g_tuning_user[type == FIJ_TYPE_I32 ? 0 : 1] = !use_builtin;
  return FIJ_OK;
}

fij_status fij_join_batch_create(const fij_join_batch_config* cfg, fij_join_batch** out) {
  if (!cfg || !out || (cfg->type != FIJ_TYPE_I32 && cfg->type != FIJ_TYPE_I64) || cfg->max_joins == 0)
    return batch_fail(FIJ_INVALID_ARGUMENT, "invalid batch configuration");
  int dev = 0;
  BATCH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  BATCH_HIP(hipGetDeviceProperties(&prop, dev));
  fij_join_batch* b = new fij_join_batch();
  b->cfg = *cfg;
  b->kt = cfg->type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  b->cus = static_cast<uint32_t>(prop.multiProcessorCount);
  int optin = 0;
  b->lds_total = static_cast<uint32_t>(prop.sharedMemPerBlock);
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
    b->lds_total = std::max<uint32_t>(b->lds_total, static_cast<uint32_t>(optin));
  const uint64_t n = cfg->max_joins;
  fij_status st = FIJ_OK;
  if (hipMalloc(&b->d_jobs, n * sizeof(SmallJob)) != hipSuccess ||
      hipMalloc(&b->d_ticket, 256) != hipSuccess ||
      hipMalloc(&b->d_results, n * sizeof(JoinAggregate)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_jobs), n * sizeof(SmallJob), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_results), n * sizeof(JoinAggregate), hipHostMallocDefault) !=
          hipSuccess ||
      hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess)
    st = batch_fail(FIJ_OUT_OF_MEMORY, "cannot allocate the batch's descriptors and results");
  b->lanes = std::min<uint32_t>(kMaxBatchLanes, std::max<uint32_t>(1, cfg->lanes));
  if (st == FIJ_OK && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    st = batch_fail(FIJ_DEVICE_ERROR, "cannot create the batch's events");
  for (uint32_t l = 1; l < b->lanes && st == FIJ_OK; ++l)
    if (hipStreamCreateWithFlags(&b->lane[l], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming) != hipSuccess)
      st = batch_fail(FIJ_DEVICE_ERROR, "cannot create the batch's lane streams");
  if (st != FIJ_OK) {
    fij_join_batch_destroy(b);
    return st;
  }
  *out = b;
  return FIJ_OK;
}

fij_status fij_join_batch_destroy(fij_join_batch* b) {
  if (!b) return FIJ_OK;
  if (b->own) (void)hipStreamSynchronize(b->own);
  for (auto& kv : b->arenas) free_arena(kv.second);
  (void)hipFree(b->d_jobs);
  (void)hipFree(b->d_ticket);
  (void)hipFree(b->d_results);
  if (b->h_jobs) (void)hipHostFree(b->h_jobs);
  if (b->h_results) (void)hipHostFree(b->h_results);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (uint32_t l = 1; l < kMaxBatchLanes; ++l) {
    if (b->lane[l]) {
      (void)hipStreamSynchronize(b->lane[l]);
      (void)hipStreamDestroy(b->lane[l]);
    }
    if (b->join[l]) (void)hipEventDestroy(b->join[l]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->own) (void)hipStreamDestroy(b->own);
  delete b;
  return FIJ_OK;
}

fij_status fij_join_batch_execute(fij_join_batch* b, uint32_t n, const fij_join_desc* joins,
                                  fij_join_aggregate* results, fij_join_batch_stats* stats, void* stream) {
  if (!b || (n && (!joins || !results))) return batch_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (n > b->cfg.max_joins) return batch_fail(FIJ_INVALID_ARGUMENT, "more joins than the batch was created for");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : b->own;
  const size_t kb = key_bytes(b->kt);
  // classify
  uint32_t n_small = 0, n_large = 0;
  std::map<ClassKey, std::vector<uint32_t>> classes;  // configuration -> join indices
  for (uint32_t i = 0; i < n; ++i) {
    const fij_join_desc& j = joins[i];
    for (const void* c : {j.build_keys, j.build_payloads, j.probe_keys, j.probe_payloads})
      if (c && reinterpret_cast<uintptr_t>(c) % kb)
        return batch_fail(FIJ_INVALID_ARGUMENT, "join %u: columns must be aligned to their element size", i);
    if ((j.build_len && (!j.build_keys || !j.build_payloads)) || (j.probe_len && (!j.probe_keys || !j.probe_payloads)))
      return batch_fail(FIJ_INVALID_ARGUMENT, "join %u: NULL input column", i);
    const fij_join_tuning tc = tuned_class(b->kt, j.build_len, j.probe_len);
    if (tc.small || j.probe_len == 0) {
      SmallJob& sj = b->h_jobs[n_small++];
      sj.build_keys = j.build_keys;
      sj.build_pays = j.build_payloads;
      sj.build_len = j.build_len;
      sj.probe_keys = j.probe_keys;
      sj.probe_pays = j.probe_payloads;
      sj.probe_len = j.probe_len;
      sj.result = b->d_results + i;
    } else {
      classes[ClassKey{tc.radix_bits_1, tc.radix_bits_2, tc.ring_lines, tc.block}].push_back(i);
      ++n_large;
    }
  }
  // every class's descriptors and (grown) buffers
  std::vector<uint8_t> host;
  double grow_ms = 0;
  for (auto& kv : classes) {
    fij_status s = enqueue_class(b, b->arenas[kv.first], kv.first, kv.second, joins, st, b->lds_total, host, &grow_ms);
    if (s != FIJ_OK) return s;
  }
  if (n_small) {
    BATCH_HIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, n_small * sizeof(SmallJob), hipMemcpyHostToDevice, st));
    BATCH_HIP(hipStreamSynchronize(st));
  }
  // record the submission (or launch it directly)
  const auto t_rec = std::chrono::steady_clock::now();
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const bool use_graph = b->cfg.use_graph != 0;
  if (use_graph) BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fij_status s = FIJ_OK;
  if (!use_graph) BATCH_HIP(hipEventRecord(b->ev[0], st));
  if (n_small) {
    hipError_t e = launch_small_joins(b->kt, b->d_jobs, n_small, b->d_ticket, b->cus, st);
    if (e != hipSuccess) s = batch_fail(FIJ_DEVICE_ERROR, "small joins: %s", hipGetErrorString(e));
  }
  // Size classes are independent (each has its own buffers and result
  // slots), so they are spread over the lanes: largest class first, each to
  // the lane with the fewest tuples so far.  Lane 0 is the submission stream
  // (which also runs the small joins); the other lanes wait for it to reach
  // this point and it waits for all of them at the end.
  std::vector<std::pair<uint64_t, const ClassKey*>> order;
  for (auto& kv : classes) {
    uint64_t tuples = 0;
    for (uint32_t i : kv.second) tuples += joins[i].build_len + joins[i].probe_len;
    order.emplace_back(tuples, &kv.first);
  }
  std::sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  const uint32_t lanes = std::min<uint32_t>(b->lanes, std::max<uint32_t>(1, static_cast<uint32_t>(order.size())));
  uint64_t lane_load[kMaxBatchLanes] = {};
  if (lanes > 1 && s == FIJ_OK) {
    hipError_t e = hipEventRecord(b->fork, st);
    for (uint32_t l = 1; l < lanes && e == hipSuccess; ++l) e = hipStreamWaitEvent(b->lane[l], b->fork, 0);
    if (e != hipSuccess) s = batch_fail(FIJ_DEVICE_ERROR, "forking the lanes: %s", hipGetErrorString(e));
  }
  for (const auto& [tuples, key] : order) {
    if (s != FIJ_OK) break;
    const uint32_t l = static_cast<uint32_t>(std::min_element(lane_load, lane_load + lanes) - lane_load);
    lane_load[l] += tuples;
    hipError_t e = launch_class(b, b->arenas[*key], *key, l ? b->lane[l] : st, b->lds_total);
    if (e != hipSuccess)
      s = batch_fail(FIJ_DEVICE_ERROR, "size class %u+%u: %s", key->b1, key->b2, hipGetErrorString(e));
  }
  // every forked lane joins back, also after an error (a capture must end
  // with every stream it forked joined)
  for (uint32_t l = 1; l < lanes; ++l) {
    hipError_t e = hipEventRecord(b->join[l], b->lane[l]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, b->join[l], 0);
    if (e != hipSuccess && s == FIJ_OK) s = batch_fail(FIJ_DEVICE_ERROR, "joining the lanes: %s", hipGetErrorString(e));
  }
  if (use_graph) {
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (s != FIJ_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return s;
    }
    if (e != hipSuccess) return batch_fail(FIJ_DEVICE_ERROR, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return batch_fail(FIJ_DEVICE_ERROR, "graph instantiation: %s", hipGetErrorString(e));
    }
  } else if (s != FIJ_OK) {
    return s;
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (use_graph) {
    BATCH_HIP(hipEventRecord(b->ev[0], st));
    BATCH_HIP(hipGraphLaunch(exec, st));
  }
  BATCH_HIP(hipEventRecord(b->ev[1], st));
  if (n) BATCH_HIP(hipMemcpyAsync(b->h_results, b->d_results, n * sizeof(JoinAggregate), hipMemcpyDeviceToHost, st));
  BATCH_HIP(hipStreamSynchronize(st));
  const auto t2 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < n; ++i) {
    results[i].count = b->h_results[i].count;
    results[i].sum_build = b->h_results[i].sum_build;
    results[i].sum_probe = b->h_results[i].sum_probe;
  }
  if (stats) {
    float ms = 0;
    BATCH_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    stats->gpu_ms = ms;
    stats->prepare_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    stats->total_ms = std::chrono::duration<float, std::milli>(t2 - t0).count();
    stats->small_joins = n_small;
    stats->partitioned_joins = n_large;
    stats->size_classes = static_cast<uint32_t>(classes.size());
    stats->graph = use_graph ? 1u : 0u;
    stats->grow_ms = static_cast<float>(grow_ms);
    stats->record_ms = std::chrono::duration<float, std::milli>(t1 - t_rec).count();
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return FIJ_OK;
}

fij_status fij_join_batch_device_bytes(const fij_join_batch* b, uint64_t* bytes) {
  if (!b || !bytes) return batch_fail(FIJ_INVALID_ARGUMENT, "NULL argument");
  uint64_t total = static_cast<uint64_t>(b->cfg.max_joins) * (sizeof(SmallJob) + sizeof(JoinAggregate)) + 256;
  const uint64_t tb = tuple_bytes(b->kt);
  for (const auto& kv : b->arenas) {
    const ClassArena& a = kv.second;
    const uint64_t parts = a.joins << kv.first.b1, F2 = 1ull << kv.first.b2;
    for (int s = 0; s < 2; ++s)
      total += (a.rel1[s] + a.rel2[s]) * tb + parts * (8 + 4 + 8 + 4) + (a.tiles[s] + 1) * F2 * 4;
    total += a.chunk_words * (4 + 8) + parts * F2 + a.units * 16 +
             a.joins * 2 * (sizeof(HistParams) + sizeof(BatchScanJob) + sizeof(ScatterParams) + sizeof(BatchSide) + 4);
  }
  *bytes = total;
  return FIJ_OK;
}

}  // extern "C"
