// Host side of the batched radix-join execution method (fij_join_plan_*).
//
// A plan owns every intermediate buffer of one join shape (allocated once,
// as the reference allocates its partitions and stream states before the
// timed phases: radix-join/src/execution_methods/gpu_radix_join.rs:153-205,
// :307-396), so that executing it launches kernels only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/fij.h"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

using namespace fij;

struct fij_join_plan {
  fij_join_plan_config cfg;
  KeyType kt;
  PipelineDevice d;
  uint32_t max_tiles[2];
  uint32_t sort2_grid;
  uint32_t join_grid;
  uint32_t g1_cap;
  uint32_t block1;
  uint32_t lds_total;
  bool large_tables;
  std::vector<void*> allocs;
  hipEvent_t ev[6];
  JoinAggregate* host_agg;
};

namespace {

fij_status plan_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(s, buf);
}

fij_status hip_fail(hipError_t e, const char* what) {
  return plan_fail(e == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, "%s: %s", what,
                   hipGetErrorString(e));
}

#define PLAN_HIP(expr)                                \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

template <typename T>
fij_status alloc(fij_join_plan* p, T** out, uint64_t count) {
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, std::max<uint64_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  p->allocs.push_back(ptr);
  *out = static_cast<T*>(ptr);
  return FIJ_OK;
}

void release(fij_join_plan* p) {
  for (void* a : p->allocs) (void)hipFree(a);
  p->allocs.clear();
  for (hipEvent_t& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  if (p->host_agg) (void)hipHostFree(p->host_agg);
}

fij_status alloc_side(fij_join_plan* p, PipelineSide& s, uint64_t n, uint32_t tiles) {
  const hipError_t e = alloc_pipeline_side(s, p->kt, n, p->d.F1, p->d.F2, p->d.G1, p->d.bits2 != 0, tiles, p->allocs);
  return e == hipSuccess ? FIJ_OK : hip_fail(e, "hipMalloc");
}

}  // namespace

extern "C" {

fij_status fij_join_plan_bits(fij_type type, uint64_t build_len, uint64_t probe_len, uint32_t* radix_bits_1,
                              uint32_t* radix_bits_2) {
  if (!radix_bits_1 || !radix_bits_2) return plan_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) return plan_fail(FIJ_INVALID_ARGUMENT, "unknown type");
  const KeyType kt = type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  if (one_pass_fits(kt, build_len)) {
    *radix_bits_1 = kOnePassBits;
    *radix_bits_2 = 0;
    return FIJ_OK;
  }
  // a probe side 16 times the build side or longer: items of about a third
  // of a block table, joined from block tables over long runs
  if (probe_len >= 16 * build_len) {
    *radix_bits_1 = kOnePassBits;
    *radix_bits_2 = probe_heavy_bits2(kt, build_len);
    return FIJ_OK;
  }
  uint32_t small = 0;
  return fij_join_autotune(type, build_len, probe_len, &small, radix_bits_1, radix_bits_2);
}

fij_status fij_join_plan_create(const fij_join_plan_config* cfg, fij_join_plan** out) {
  if (!cfg || !out) return plan_fail(FIJ_INVALID_ARGUMENT, "NULL argument");
  if (cfg->type != FIJ_TYPE_I32 && cfg->type != FIJ_TYPE_I64) return plan_fail(FIJ_INVALID_ARGUMENT, "unknown type");
  if (cfg->radix_bits_1 < 1 || cfg->radix_bits_1 > 12)
    return plan_fail(FIJ_UNSUPPORTED, "pass-1 radix bits must be in [1, 12]");
  if (cfg->radix_bits_2 > 12) return plan_fail(FIJ_UNSUPPORTED, "pass-2 radix bits must be in [0, 12]");
  const KeyType kt = cfg->type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  if (cfg->radix_bits_1 + cfg->radix_bits_2 > 8 * key_bytes(kt))
    return plan_fail(FIJ_INVALID_ARGUMENT, "more radix bits than key bits");
  int dev = 0;
  PLAN_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PLAN_HIP(hipGetDeviceProperties(&prop, dev));
  int optin = 0;
  uint32_t lds_total = static_cast<uint32_t>(prop.sharedMemPerBlock);
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
    lds_total = std::max<uint32_t>(lds_total, static_cast<uint32_t>(optin));
  const uint32_t cus = static_cast<uint32_t>(prop.multiProcessorCount);

  fij_join_plan* p = new fij_join_plan();
  std::memset(p->ev, 0, sizeof(p->ev));
  p->host_agg = nullptr;
  p->cfg = *cfg;
  p->kt = kt;
  p->lds_total = lds_total;
  std::memset(&p->d, 0, sizeof(p->d));
  p->d.bits1 = cfg->radix_bits_1;
  p->d.bits2 = cfg->radix_bits_2;
  p->d.F1 = 1u << cfg->radix_bits_1;
  p->d.F2 = 1u << cfg->radix_bits_2;
  p->d.G1 = cfg->pass1_chunks ? cfg->pass1_chunks : cus;
  if (!cfg->pass1_chunks) {
    // automatic: one chunk per compute unit, fewer for short relations
    auto fits = [&](uint64_t n, uint32_t chunks) {
      uint64_t c = 0;
      return fij_input_chunk_size(static_cast<fij_type>(cfg->type), n, chunks, &c) == FIJ_OK && c * (chunks - 1) <= n;
    };
    while (p->d.G1 > 1 && !(fits(cfg->build_len, p->d.G1) && fits(cfg->probe_len, p->d.G1))) --p->d.G1;
  }
  p->g1_cap = p->d.G1;
  p->d.pad = pad_tuples(kt);
  p->d.tile2 = pipeline_tile2(kt);
  p->d.sides = 3;
  p->block1 = cfg->pass1_block ? std::min<uint32_t>(cfg->pass1_block, 1024) : 1024;
  p->large_tables = cfg->radix_bits_2 == 0;
  p->join_grid = cfg->join_blocks ? cfg->join_blocks : cus * 2;
  p->d.agg_len = p->join_grid;
  fij_status st = FIJ_OK;
  const uint64_t lens[2] = {cfg->build_len, cfg->probe_len};
  for (int rel = 0; rel < 2 && st == FIJ_OK; ++rel) {
    uint64_t chunk = 0;
    st = fij_input_chunk_size(static_cast<fij_type>(cfg->type), lens[rel], p->d.G1, &chunk);
    if (st == FIJ_OK && chunk * (p->d.G1 - 1) > lens[rel])
      st = plan_fail(FIJ_INVALID_ARGUMENT, "relation of %llu tuples is too short for %u chunks",
                     static_cast<unsigned long long>(lens[rel]), p->d.G1);
    // the scatter addresses a chunk's writes relative to its first partition
    // with 32-bit offsets; in the contiguous layout they span the relation
    if (st == FIJ_OK && lens[rel] + static_cast<uint64_t>(p->d.F1) * p->d.pad * 2 >= 0xFFFFFFFFull)
      st = plan_fail(FIJ_INTEGER_OVERFLOW, "relation of %llu tuples exceeds the 32-bit pass-1 offsets",
                     static_cast<unsigned long long>(lens[rel]));
    p->max_tiles[rel] = p->d.bits2 ? static_cast<uint32_t>((lens[rel] + p->d.tile2 - 1) / p->d.tile2 + p->d.F1) : 0;
  }
  if (st == FIJ_OK && p->d.bits2 && pipeline_sort2_lds(kt, p->d.F2) > lds_total)
    st = plan_fail(FIJ_UNSUPPORTED, "pass-2 tile does not fit LDS for this fanout / chunk count");
  // pass 2: one persistent block per compute unit (its LDS tile fills the CU)
  p->sort2_grid = p->d.bits2 ? std::min<uint32_t>(cus, p->max_tiles[0] + p->max_tiles[1]) : 0;
  if (st == FIJ_OK) st = alloc_side(p, p->d.build, cfg->build_len, p->max_tiles[0]);
  if (st == FIJ_OK) st = alloc_side(p, p->d.probe, cfg->probe_len, p->max_tiles[1]);
  if (st == FIJ_OK) st = alloc(p, &p->d.work_counter, kJoinCounterWords);
  const uint64_t items = static_cast<uint64_t>(p->d.F1) * (p->d.bits2 ? p->d.F2 : 1);
  if (st == FIJ_OK) st = alloc(p, &p->d.overflow, items);
  // work units of deferred items: at most one per item plus one per 65536
  // probe tuples
  p->d.unit_cap = static_cast<uint32_t>(items + cfg->probe_len / 65536 + 1);
  if (st == FIJ_OK) st = alloc(p, &p->d.units, p->d.unit_cap);
  // the deferred-item flags start at zero (complete before any stream uses
  // them); every join leaves them at zero
  if (st == FIJ_OK && (hipMemset(p->d.overflow, 0, items) != hipSuccess || hipDeviceSynchronize() != hipSuccess))
    st = plan_fail(FIJ_DEVICE_ERROR, "clearing the deferred-item flags failed");
  // HBM hash tables of deferred items with large build sides: an arena of
  // 2^k slots, twice the build relation (at most 2^26 slots), every word all
  // ones; one table per item at most, each at least 256 slots
  {
    uint64_t slots = 1024;
    while (slots < 2 * cfg->build_len && slots < (1ull << 26)) slots <<= 1;
    p->d.hbm_cap = static_cast<uint32_t>(slots);
    p->d.big_cap = static_cast<uint32_t>(std::min<uint64_t>(items, slots / 256));
    p->d.hbm_unit_cap =
        static_cast<uint32_t>(2ull * p->d.big_cap + (cfg->build_len + cfg->probe_len) / 65536 + 2);
    void* arena = nullptr;
    if (st == FIJ_OK && hipMalloc(&arena, slots * tuple_bytes(kt)) != hipSuccess)
      st = plan_fail(FIJ_OUT_OF_MEMORY, "cannot allocate the HBM hash-table arena");
    if (arena) p->allocs.push_back(arena);
    p->d.hbm_table = arena;
    if (st == FIJ_OK) st = alloc(p, &p->d.big, p->d.big_cap);
    if (st == FIJ_OK) st = alloc(p, &p->d.hbm_units, p->d.hbm_unit_cap);
    if (st == FIJ_OK && (hipMemset(arena, 0xFF, slots * tuple_bytes(kt)) != hipSuccess ||
                         hipDeviceSynchronize() != hipSuccess))
      st = plan_fail(FIJ_DEVICE_ERROR, "initialising the HBM hash-table arena failed");
  }
  if (st == FIJ_OK) st = alloc(p, &p->d.aggregates, p->join_grid);
  for (int i = 0; i < 6 && st == FIJ_OK; ++i)
    if (hipEventCreate(&p->ev[i]) != hipSuccess) st = plan_fail(FIJ_DEVICE_ERROR, "hipEventCreate failed");
  if (st == FIJ_OK && hipHostMalloc(reinterpret_cast<void**>(&p->host_agg), sizeof(JoinAggregate) * p->join_grid,
                                    hipHostMallocDefault) != hipSuccess)
    st = plan_fail(FIJ_OUT_OF_MEMORY, "cannot allocate pinned aggregates");
  if (st != FIJ_OK) {
    release(p);
    delete p;
    return st;
  }
  *out = p;
  return FIJ_OK;
}

fij_status fij_join_plan_destroy(fij_join_plan* plan) {
  if (!plan) return FIJ_OK;
  release(plan);
  delete plan;
  return FIJ_OK;
}

}  // extern "C"

namespace {

// Chunks of pass 1 for relations of n_b and n_p tuples: the plan's chunk
// count, fewer when a relation is too short for it (every chunk but the last
// has the canonical length, the last one the rest).
uint32_t exec_chunks(const fij_join_plan* p, uint64_t n_b, uint64_t n_p) {
  uint32_t g = p->g1_cap;
  auto fits = [&](uint64_t n, uint32_t chunks) {
    uint64_t c = 0;
    return fij_input_chunk_size(static_cast<fij_type>(p->cfg.type), n, chunks, &c) == FIJ_OK && c * (chunks - 1) <= n;
  };
  while (g > 1 && !(fits(n_b, g) && fits(n_p, g))) g = g > 64 ? g - 32 : g - 1;
  return g;
}

// Enqueues one join of the plan's shape over columns of n_b and n_p tuples
// (at most the plan's lengths).  With events, the phases are timed and the
// per-block aggregates are copied to the host; otherwise (capturable: no
// events, no copies) they are summed into *device_result.
fij_status enqueue_join(fij_join_plan* p, const void* build_keys, const void* build_pays, uint64_t n_b,
                        const void* probe_keys, const void* probe_pays, uint64_t n_p, bool events,
                        fij_join_aggregate* device_result, hipStream_t st) {
  if (n_b > p->cfg.build_len || n_p > p->cfg.probe_len)
    return plan_fail(FIJ_INVALID_ARGUMENT, "join of %llu x %llu tuples exceeds the plan's %llu x %llu",
                     static_cast<unsigned long long>(n_b), static_cast<unsigned long long>(n_p),
                     static_cast<unsigned long long>(p->cfg.build_len),
                     static_cast<unsigned long long>(p->cfg.probe_len));
  if ((n_b && (!build_keys || !build_pays)) || (n_p && (!probe_keys || !probe_pays)))
    return plan_fail(FIJ_INVALID_ARGUMENT, "NULL input column");
  const size_t kb = key_bytes(p->kt);
  for (const void* c : {build_keys, build_pays, probe_keys, probe_pays})
    if (c && reinterpret_cast<uintptr_t>(c) % kb)
      return plan_fail(FIJ_INVALID_ARGUMENT, "input columns must be aligned to their element size");
  PipelineDevice& d = p->d;
  d.build.keys = build_keys;
  d.build.pays = build_pays;
  d.probe.keys = probe_keys;
  d.probe.pays = probe_pays;
  d.build.n = n_b;
  d.probe.n = n_p;
  d.build.padded_len1 = n_b + static_cast<uint64_t>(d.F1) * d.pad;
  d.probe.padded_len1 = n_p + static_cast<uint64_t>(d.F1) * d.pad;
  d.G1 = exec_chunks(p, n_b, n_p);
  d.big_block_tables = d.bits2 ? 0u : big_block_tables_for(p->kt, n_b, d.F1);
  if (events) PLAN_HIP(hipEventRecord(p->ev[0], st));
  for (PipelineSide* s : {&d.build, &d.probe}) PLAN_HIP(enqueue_side_prefix_sum(d, *s, p->kt, p->block1, st));
  if (events) PLAN_HIP(hipEventRecord(p->ev[1], st));
  for (PipelineSide* s : {&d.build, &d.probe})
    PLAN_HIP(enqueue_side_scatter(d, *s, p->kt, p->block1, p->lds_total, st));
  if (events) PLAN_HIP(hipEventRecord(p->ev[2], st));
  PLAN_HIP(launch_pipeline_plan(d, st));
  if (d.bits2) PLAN_HIP(launch_pipeline_sort2(d, p->kt, p->sort2_grid, st));
  if (events) PLAN_HIP(hipEventRecord(p->ev[3], st));
  PLAN_HIP(launch_pipeline_join(d, p->kt, p->large_tables, false, p->join_grid, st));
  if (events) PLAN_HIP(hipEventRecord(p->ev[4], st));
  PLAN_HIP(launch_pipeline_join(d, p->kt, p->large_tables, true, p->join_grid, st));
  if (events) {
    PLAN_HIP(hipEventRecord(p->ev[5], st));
    PLAN_HIP(hipMemcpyAsync(p->host_agg, d.aggregates, sizeof(JoinAggregate) * p->join_grid, hipMemcpyDeviceToHost,
                            st));
  }
  if (device_result)
    PLAN_HIP(launch_sum_aggregates(d.aggregates, p->join_grid, reinterpret_cast<JoinAggregate*>(device_result), st));
  return FIJ_OK;
}

}  // namespace

extern "C" {

fij_status fij_join_plan_enqueue(fij_join_plan* p, const void* build_keys, const void* build_pays,
                                 const void* probe_keys, const void* probe_pays, void* stream) {
  if (!p) return plan_fail(FIJ_INVALID_ARGUMENT, "plan is NULL");
  return enqueue_join(p, build_keys, build_pays, p->cfg.build_len, probe_keys, probe_pays, p->cfg.probe_len, true,
                      nullptr, static_cast<hipStream_t>(stream));
}

fij_status fij_join_plan_enqueue_sized(fij_join_plan* p, const void* build_keys, const void* build_pays,
                                       uint64_t build_len, const void* probe_keys, const void* probe_pays,
                                       uint64_t probe_len, fij_join_aggregate* device_result, void* stream) {
  if (!p || !device_result) return plan_fail(FIJ_INVALID_ARGUMENT, "plan or result is NULL");
  return enqueue_join(p, build_keys, build_pays, build_len, probe_keys, probe_pays, probe_len, false, device_result,
                      static_cast<hipStream_t>(stream));
}

fij_status fij_join_plan_result(fij_join_plan* p, void* stream, fij_join_aggregate* result, fij_join_timing* timing) {
  if (!p) return plan_fail(FIJ_INVALID_ARGUMENT, "plan is NULL");
  PLAN_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  if (result) {
    uint64_t c = 0, b = 0, s = 0;
    for (uint32_t i = 0; i < p->join_grid; ++i) {
      c += p->host_agg[i].count;
      b += p->host_agg[i].sum_build;
      s += p->host_agg[i].sum_probe;
    }
    result->count = c;
    result->sum_build = b;
    result->sum_probe = s;
  }
  if (timing) {
    float ms[6] = {0, 0, 0, 0, 0, 0};
    const int from[6] = {0, 1, 2, 0, 2, 3}, to[6] = {1, 2, 5, 5, 3, 4};
    for (int i = 0; i < 6; ++i) PLAN_HIP(hipEventElapsedTime(&ms[i], p->ev[from[i]], p->ev[to[i]]));
    float deferred = 0;
    PLAN_HIP(hipEventElapsedTime(&deferred, p->ev[4], p->ev[5]));
    timing->prefix_sum_ms = ms[0];
    timing->partition_ms = ms[1];
    timing->join_ms = ms[2];
    timing->total_ms = ms[3];
    timing->pass2_ms = ms[4];
    timing->build_probe_ms = ms[5];
    timing->deferred_ms = deferred;
  }
  return FIJ_OK;
}

fij_status fij_join_plan_execute(fij_join_plan* p, const void* build_keys, const void* build_pays,
                                 const void* probe_keys, const void* probe_pays, fij_join_aggregate* result,
                                 fij_join_timing* timing, void* stream) {
  fij_status s = fij_join_plan_enqueue(p, build_keys, build_pays, probe_keys, probe_pays, stream);
  if (s != FIJ_OK) return s;
  return fij_join_plan_result(p, stream, result, timing);
}

fij_status fij_join_plan_device_bytes(const fij_join_plan* p, uint64_t* bytes) {
  if (!p || !bytes) return plan_fail(FIJ_INVALID_ARGUMENT, "NULL argument");
  const uint64_t tb = tuple_bytes(p->kt);
  uint64_t total = 0;
  for (const PipelineSide* s : {&p->d.build, &p->d.probe}) {
    total += s->padded_len1 * tb + static_cast<uint64_t>(p->d.F1) * (p->d.G1 + 1) * 8;
    total += static_cast<uint64_t>(p->d.F1) * p->d.G1 * 4 + s->scan1_len * 8 + p->d.F1 * 4 + (p->d.F1 + 1) * 12;
    if (p->d.bits2) total += s->n * tb;
  }
  if (p->d.bits2) total += (static_cast<uint64_t>(p->max_tiles[0]) + p->max_tiles[1] + 2) * p->d.F2 * 4;
  total += static_cast<uint64_t>(p->d.hbm_cap) * tb + (static_cast<uint64_t>(p->d.big_cap) + p->d.hbm_unit_cap) * 16;
  *bytes = total;
  return FIJ_OK;
}

}  // extern "C"
