// Served joins (fij_join_server_*): a request loop over resident build sides.
//
// The reference has no serving mode; each run of its radix-join binary
// partitions both relations of one join
// (radix-join/src/execution_methods/gpu_radix_join.rs:90-565).  Here:
//   * build sides are registered once: partitioned (both passes) and kept in
//     device memory, charged against the server's HBM budget;
//   * a request names a build side and brings a probe side (device columns);
//     requests wait in arrival order and are admitted as soon as a slot that
//     fits them is free -- a slot is a preallocated probe-side workspace
//     with its own stream, so admitted requests overlap on the device
//     (continuous batching: a request arriving while others run starts
//     whenever a slot frees up, not when the whole batch ends);
//   * a request partitions only its probe side, with the build side's radix
//     bits, and joins it with the resident partitions; its {count, sums} are
//     identical to joining the pair alone.
// A scheduler thread admits requests and polls completions; callers submit
// from any thread and wait for results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/fij.h"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

using namespace fij;

namespace {

using Clock = std::chrono::steady_clock;

fij_status server_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(s, buf);
}

fij_status hip_status(hipError_t e, const char* what) {
  return server_fail(e == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, "%s: %s", what,
                     hipGetErrorString(e));
}

#define SRV_HIP(expr)                                \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_status(e_, #expr); \
  } while (0)

constexpr uint32_t kMaxF1 = 512;
constexpr uint32_t kMaxF2 = 4096;

uint64_t allocated_bytes(const std::vector<void*>& allocs) {
  uint64_t total = 0;
  for (void* p : allocs) {
    size_t sz = 0;
    if (hipMemPtrGetInfo(p, &sz) == hipSuccess) total += sz;
  }
  return total;
}

void free_all(std::vector<void*>& allocs) {
  for (void* p : allocs) (void)hipFree(p);
  allocs.clear();
}

struct ResidentBuild {
  uint32_t bits1 = 0, bits2 = 0;
  uint64_t n = 0;
  PipelineSide side{};
  // a build side of at most table_build_max tuples: one linear-probing table
  // of 2^table_lg entries instead of partitions (null: partitioned)
  void* table = nullptr;
  uint32_t table_lg = 0;
  std::vector<void*> allocs;
  uint64_t bytes = 0;
};

struct Slot {
  uint64_t probe_cap = 0;
  PipelineSide probe{};
  uint32_t* work_counter = nullptr;
  uint8_t* overflow = nullptr;
  uint4* units = nullptr;
  uint32_t unit_cap = 0;
  JoinAggregate* aggregates = nullptr;
  JoinAggregate* d_result = nullptr;
  JoinAggregate* h_result = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  int64_t request = -1;
};

struct Request {
  uint64_t id;
  uint32_t build;
  const void* keys;
  const void* pays;
  uint64_t len;
  Clock::time_point arrival, start, end;
  fij_join_aggregate result{};
  fij_status status = FIJ_OK;
  bool done = false;
};

}  // namespace

struct fij_join_server {
  fij_join_server_config cfg;
  KeyType kt;
  int device = 0;
  uint32_t cus = 0, lds_total = 0, join_grid = 0;
  uint64_t budget_used = 0;
  std::vector<std::unique_ptr<ResidentBuild>> builds;
  std::vector<Slot> slots;
  // registration scratch: counters and aggregates the plan kernel resets
  uint32_t* scratch_counters = nullptr;
  JoinAggregate* scratch_agg = nullptr;
  hipStream_t setup = nullptr;
  mutable std::mutex mu;
  std::condition_variable cv;
  std::deque<uint64_t> queue;
  std::map<uint64_t, Request> requests;
  uint64_t next_id = 0;
  bool stop = false;
  fij_status fault = FIJ_OK;
  std::thread scheduler;
};

namespace {

PipelineDevice base_device(const fij_join_server* s, const ResidentBuild& b) {
  PipelineDevice d{};
  d.build = b.side;
  d.bits1 = b.bits1;
  d.bits2 = b.bits2;
  d.F1 = 1u << b.bits1;
  d.F2 = 1u << b.bits2;
  d.G1 = s->cus;
  d.pad = pad_tuples(s->kt);
  d.tile2 = pipeline_tile2(s->kt);
  d.agg_len = s->join_grid;
  return d;
}

uint32_t chunks_for(KeyType kt, uint64_t n, uint32_t cap) {
  uint32_t g = cap;
  auto fits = [&](uint32_t chunks) {
    uint64_t c = 0;
    return fij_input_chunk_size(kt == KeyType::I32 ? FIJ_TYPE_I32 : FIJ_TYPE_I64, n, chunks, &c) == FIJ_OK &&
           c * (chunks - 1) <= n;
  };
  while (g > 1 && !fits(g)) g = g > 64 ? g - 32 : g - 1;
  return g;
}

// Launches request r on slot sl (scheduler thread, server lock held).
fij_status launch_request(fij_join_server* s, Slot& sl, const Request& r) {
  const ResidentBuild& b = *s->builds[r.build];
  if (b.table) {
    // a resident table: one probe launch, no partitioning of the probe side
    hipStream_t st = sl.stream;
    SRV_HIP(launch_np_probe_resident(s->kt, b.table, b.table_lg, r.keys, r.pays, r.len, sl.aggregates, s->join_grid,
                                     st));
    SRV_HIP(launch_sum_aggregates(sl.aggregates, s->join_grid, sl.d_result, st));
    SRV_HIP(hipMemcpyAsync(sl.h_result, sl.d_result, sizeof(JoinAggregate), hipMemcpyDeviceToHost, st));
    SRV_HIP(hipEventRecord(sl.done, st));
    return FIJ_OK;
  }
  PipelineDevice d = base_device(s, b);
  d.probe = sl.probe;
  d.probe.keys = r.keys;
  d.probe.pays = r.pays;
  d.probe.n = r.len;
  d.probe.padded_len1 = r.len + static_cast<uint64_t>(d.F1) * d.pad;
  d.probe.scan1_len = scan_state_words(static_cast<uint64_t>(d.F1) * s->cus);
  d.G1 = chunks_for(s->kt, r.len, s->cus);
  d.big_block_tables = d.bits2 ? 0u : big_block_tables_for(s->kt, b.n, d.F1);
  d.sides = 2;
  d.work_counter = sl.work_counter;
  d.overflow = sl.overflow;
  d.units = sl.units;
  d.unit_cap = sl.unit_cap;
  d.aggregates = sl.aggregates;
  hipStream_t st = sl.stream;
  SRV_HIP(enqueue_side_prefix_sum(d, d.probe, s->kt, 1024, st));
  SRV_HIP(enqueue_side_scatter(d, d.probe, s->kt, 1024, s->lds_total, st));
  SRV_HIP(launch_pipeline_plan(d, st));
  if (d.bits2) {
    const uint32_t tiles = static_cast<uint32_t>(r.len / d.tile2 + d.F1);
    SRV_HIP(launch_pipeline_sort2(d, s->kt, std::max(1u, std::min(s->cus, tiles)), st));
  }
  const bool large = d.bits2 == 0;
  SRV_HIP(launch_pipeline_join(d, s->kt, large, false, s->join_grid, st));
  SRV_HIP(launch_pipeline_join(d, s->kt, large, true, s->join_grid, st));
  SRV_HIP(launch_sum_aggregates(sl.aggregates, s->join_grid, sl.d_result, st));
  SRV_HIP(hipMemcpyAsync(sl.h_result, sl.d_result, sizeof(JoinAggregate), hipMemcpyDeviceToHost, st));
  SRV_HIP(hipEventRecord(sl.done, st));
  return FIJ_OK;
}

void scheduler_loop(fij_join_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progress = false;
    // completions
    for (Slot& sl : s->slots) {
      if (sl.request < 0) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      Request& r = s->requests[static_cast<uint64_t>(sl.request)];
      if (q != hipSuccess) {
        r.status = hip_status(q, "request");
        s->fault = r.status;
      } else {
        r.result.count = sl.h_result->count;
        r.result.sum_build = sl.h_result->sum_build;
        r.result.sum_probe = sl.h_result->sum_probe;
      }
      r.end = Clock::now();
      r.done = true;
      sl.request = -1;
      progress = true;
    }
    // admissions: in arrival order, each request to the smallest free slot
    // that fits it; a request that fits no free slot waits without holding
    // back later requests that fit another free slot
    for (auto it = s->queue.begin(); it != s->queue.end();) {
      Request& r = s->requests[*it];
      Slot* best = nullptr;
      for (Slot& sl : s->slots)
        if (sl.request < 0 && sl.probe_cap >= r.len && (!best || sl.probe_cap < best->probe_cap)) best = &sl;
      if (!best) {
        ++it;
        continue;
      }
      r.start = Clock::now();
      const fij_status st = launch_request(s, *best, r);
      if (st != FIJ_OK) {
        r.status = st;
        r.end = Clock::now();
        r.done = true;
      } else {
        best->request = static_cast<int64_t>(r.id);
      }
      it = s->queue.erase(it);
      progress = true;
    }
    if (progress) s->cv.notify_all();
    bool busy = false;
    for (const Slot& sl : s->slots) busy |= sl.request >= 0;
    if (s->stop && !busy && s->queue.empty()) return;
    if (busy) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    } else if (!progress) {
      s->cv.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

fij_status make_slot(fij_join_server* s, Slot& sl, uint64_t probe_cap) {
  sl.probe_cap = probe_cap;
  const uint32_t tiles = static_cast<uint32_t>(probe_cap / pipeline_tile2(s->kt) + kMaxF1);
  SRV_HIP(alloc_pipeline_side(sl.probe, s->kt, probe_cap, kMaxF1, kMaxF2, s->cus, true, tiles, sl.allocs));
  auto alloc = [&](void** p, uint64_t bytes) -> hipError_t {
    const hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e == hipSuccess) sl.allocs.push_back(*p);
    return e;
  };
  const uint64_t items = static_cast<uint64_t>(kMaxF1) * kMaxF2;
  SRV_HIP(alloc(reinterpret_cast<void**>(&sl.work_counter), kJoinCounterWords * 4));
  SRV_HIP(alloc(reinterpret_cast<void**>(&sl.overflow), items));
  SRV_HIP(hipMemset(sl.overflow, 0, items));
  sl.unit_cap = static_cast<uint32_t>(items + probe_cap / 65536 + 1);
  SRV_HIP(alloc(reinterpret_cast<void**>(&sl.units), sl.unit_cap * sizeof(uint4)));
  SRV_HIP(alloc(reinterpret_cast<void**>(&sl.aggregates), s->join_grid * sizeof(JoinAggregate)));
  SRV_HIP(alloc(reinterpret_cast<void**>(&sl.d_result), sizeof(JoinAggregate)));
  SRV_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_result), sizeof(JoinAggregate), hipHostMallocDefault));
  SRV_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
  SRV_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  sl.bytes = allocated_bytes(sl.allocs);
  return FIJ_OK;
}

void destroy_slot(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  free_all(sl.allocs);
  if (sl.h_result) (void)hipHostFree(sl.h_result);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
}

}  // namespace

extern "C" {

fij_status fij_join_server_create(const fij_join_server_config* cfg, fij_join_server** out) {
  if (!cfg || !out || (cfg->type != FIJ_TYPE_I32 && cfg->type != FIJ_TYPE_I64) || cfg->slot_classes == 0 ||
      cfg->slot_classes > FIJ_SERVER_MAX_SLOT_CLASSES)
    return server_fail(FIJ_INVALID_ARGUMENT, "invalid server configuration");
  uint32_t total_slots = 0;
  for (uint32_t c = 0; c < cfg->slot_classes; ++c) {
    if (cfg->slots[c] && cfg->slot_probe_len[c] == 0)
      return server_fail(FIJ_INVALID_ARGUMENT, "slot class %u holds no probe tuples", c);
    total_slots += cfg->slots[c];
  }
  if (total_slots == 0) return server_fail(FIJ_INVALID_ARGUMENT, "a server needs at least one slot");
  int dev = 0;
  SRV_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SRV_HIP(hipGetDeviceProperties(&prop, dev));
  auto s = std::make_unique<fij_join_server>();
  s->cfg = *cfg;
  s->kt = cfg->type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  s->device = dev;
  s->cus = static_cast<uint32_t>(prop.multiProcessorCount);
  s->join_grid = 2 * s->cus;
  int optin = 0;
  s->lds_total = static_cast<uint32_t>(prop.sharedMemPerBlock);
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
    s->lds_total = std::max<uint32_t>(s->lds_total, static_cast<uint32_t>(optin));
  s->slots.resize(total_slots);
  uint32_t k = 0;
  fij_status st = FIJ_OK;
  for (uint32_t c = 0; c < cfg->slot_classes && st == FIJ_OK; ++c)
    for (uint32_t i = 0; i < cfg->slots[c] && st == FIJ_OK; ++i) {
      st = make_slot(s.get(), s->slots[k], cfg->slot_probe_len[c]);
      s->budget_used += s->slots[k++].bytes;
    }
  if (st == FIJ_OK && (hipMalloc(&s->scratch_counters, kJoinCounterWords * 4) != hipSuccess ||
                       hipMalloc(&s->scratch_agg, s->join_grid * sizeof(JoinAggregate)) != hipSuccess ||
                       hipStreamCreateWithFlags(&s->setup, hipStreamNonBlocking) != hipSuccess))
    st = server_fail(FIJ_OUT_OF_MEMORY, "cannot allocate the server's scratch");
  if (st == FIJ_OK && cfg->hbm_budget && s->budget_used > cfg->hbm_budget)
    st = server_fail(FIJ_OUT_OF_MEMORY, "slots need %llu bytes, more than the %llu-byte budget",
                     static_cast<unsigned long long>(s->budget_used),
                     static_cast<unsigned long long>(cfg->hbm_budget));
  if (st == FIJ_OK && hipDeviceSynchronize() != hipSuccess) st = server_fail(FIJ_DEVICE_ERROR, "setup failed");
  if (st != FIJ_OK) {
    for (Slot& sl : s->slots) destroy_slot(sl);
    (void)hipFree(s->scratch_counters);
    (void)hipFree(s->scratch_agg);
    if (s->setup) (void)hipStreamDestroy(s->setup);
    return st;
  }
  fij_join_server* raw = s.release();
  raw->scheduler = std::thread(scheduler_loop, raw);
  *out = raw;
  return FIJ_OK;
}

fij_status fij_join_server_destroy(fij_join_server* s) {
  if (!s) return FIJ_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->scheduler.joinable()) s->scheduler.join();
  for (Slot& sl : s->slots) destroy_slot(sl);
  for (auto& b : s->builds) free_all(b->allocs);
  (void)hipFree(s->scratch_counters);
  (void)hipFree(s->scratch_agg);
  if (s->setup) (void)hipStreamDestroy(s->setup);
  delete s;
  return FIJ_OK;
}

}  // extern "C"

namespace {

// Charges a built side against the HBM budget and makes it resident.
fij_status register_build(fij_join_server* s, std::unique_ptr<ResidentBuild> b, uint32_t* build_id) {
  b->bytes = allocated_bytes(b->allocs);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->cfg.hbm_budget && s->budget_used + b->bytes > s->cfg.hbm_budget) {
    free_all(b->allocs);
    return server_fail(FIJ_OUT_OF_MEMORY, "resident build of %llu bytes exceeds the HBM budget",
                       static_cast<unsigned long long>(b->bytes));
  }
  s->budget_used += b->bytes;
  *build_id = static_cast<uint32_t>(s->builds.size());
  s->builds.push_back(std::move(b));
  return FIJ_OK;
}

// One-pass resident build side: pass 1 only (bits1 bits), kept in the
// pass-1 layout with its offsets and partition lengths.
fij_status add_one_pass_build(fij_join_server* s, std::unique_ptr<ResidentBuild> b, const void* keys,
                              const void* payloads, uint64_t len, uint32_t b1, uint32_t* build_id) {
  b->bits1 = b1;
  b->bits2 = 0;
  b->n = len;
  const uint32_t F1 = 1u << b1;
  hipError_t e = alloc_pipeline_side(b->side, s->kt, len, F1, 1, s->cus, false, 0, b->allocs);
  if (e != hipSuccess) {
    free_all(b->allocs);
    return hip_status(e, "hipMalloc");
  }
  PipelineDevice d = base_device(s, *b);
  d.build.keys = keys;
  d.build.pays = payloads;
  d.build.n = len;
  d.G1 = chunks_for(s->kt, len, s->cus);
  d.sides = 1;
  d.work_counter = s->scratch_counters;
  d.aggregates = s->scratch_agg;
  hipStream_t stm = s->setup;
  if ((e = enqueue_side_prefix_sum(d, d.build, s->kt, 1024, stm)) != hipSuccess ||
      (e = enqueue_side_scatter(d, d.build, s->kt, 1024, s->lds_total, stm)) != hipSuccess ||
      (e = launch_pipeline_plan(d, stm)) != hipSuccess || (e = hipStreamSynchronize(stm)) != hipSuccess) {
    free_all(b->allocs);
    return hip_status(e, "partitioning the build side");
  }
  // the join reads rel1, off1 and plen: release the histogram and scan buffers
  std::vector<void*> keep;
  for (void* p : b->allocs) {
    if (p == b->side.local1 || p == b->side.counts1 || p == b->side.scan1)
      (void)hipFree(p);
    else
      keep.push_back(p);
  }
  b->allocs.swap(keep);
  b->side.local1 = nullptr;
  b->side.counts1 = nullptr;
  b->side.scan1 = nullptr;
  b->side.keys = nullptr;
  b->side.pays = nullptr;
  return register_build(s, std::move(b), build_id);
}

}  // namespace

extern "C" {

fij_status fij_join_server_add_build(fij_join_server* s, const void* keys, const void* payloads, uint64_t len,
                                     uint32_t* build_id) {
  if (!s || !build_id || (len && (!keys || !payloads))) return server_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const size_t kb = key_bytes(s->kt);
  if (reinterpret_cast<uintptr_t>(keys) % kb || reinterpret_cast<uintptr_t>(payloads) % kb)
    return server_fail(FIJ_INVALID_ARGUMENT, "columns must be aligned to their element size");
  if (len + static_cast<uint64_t>(kMaxF1) * pad_tuples(s->kt) * 2 >= 0xFFFFFFFFull)
    return server_fail(FIJ_INTEGER_OVERFLOW, "build side exceeds the 32-bit pass-1 offsets");
  auto b = std::make_unique<ResidentBuild>();
  if (len == 0) {  // joins nothing: requests against it complete at submission
    std::lock_guard<std::mutex> lk(s->mu);
    *build_id = static_cast<uint32_t>(s->builds.size());
    s->builds.push_back(std::move(b));
    return FIJ_OK;
  }
  if (s->cfg.table_build_max && len <= s->cfg.table_build_max) {
    // Small enough for the GPU's cache: one linear-probing table of at least
    // twice its tuples (the no-partitioning join's build,
    // no_partitioning_join.cu:39-132); a request then probes it in one
    // launch, its probe side read once instead of partitioned twice.
    uint32_t lg = 1;
    while ((1ull << lg) < 2 * len) ++lg;
    const uint64_t entries = 1ull << lg;
    void* tab = nullptr;
    hipError_t e = hipMalloc(&tab, entries * tuple_bytes(s->kt));
    if (e != hipSuccess) return hip_status(e, "hipMalloc");
    b->allocs.push_back(tab);
    if ((e = launch_np_table_init(s->kt, tab, entries, true, s->setup)) != hipSuccess ||
        (e = launch_np_build(s->kt, JoinScheme::LinearProbing, false, tab, entries, keys, payloads, len, 4 * s->cus,
                             256, s->setup)) != hipSuccess ||
        (e = hipStreamSynchronize(s->setup)) != hipSuccess) {
      free_all(b->allocs);
      return hip_status(e, "building the resident table");
    }
    b->table = tab;
    b->table_lg = lg;
    b->n = len;
    return register_build(s, std::move(b), build_id);
  }
  // A build side whose pass-1 partitions fit block tables is partitioned in
  // one pass of 9 bits and kept in its pass-1 layout: a request partitions
  // its probe side once and streams its partitions against them.
  uint32_t b1 = 0, b2 = 0;
  fij_status st = fij_join_plan_bits(static_cast<fij_type>(s->cfg.type), len, s->cfg.max_probe_len, &b1, &b2);
  if (st != FIJ_OK) return st;
  if (b2 == 0) return add_one_pass_build(s, std::move(b), keys, payloads, len, b1, build_id);
  // Otherwise two passes of 9 + b2 bits, whose items hold about a third of
  // a block table's build tuples: the build side serves probes of any length
  // up to max_probe_len, and most probe tuples come from long probes, whose
  // items the block join streams over long runs.  Against items of about 512
  // build tuples (the size-class rule) the served stream ran 6.6 % faster
  // (saturated 60.4 -> 64.4 G tuples/s, p50 1.76 -> 1.59 ms).
  b1 = kOnePassBits;
  b2 = probe_heavy_bits2(s->kt, len);
  if (b1 > 9 || b2 == 0 || b2 > 12) return server_fail(FIJ_INVALID_ARGUMENT, "unsupported radix bits %u+%u", b1, b2);
  b->bits1 = b1;
  b->bits2 = b2;
  b->n = len;
  const uint32_t F1 = 1u << b1, F2 = 1u << b2;
  const uint32_t tiles = static_cast<uint32_t>(len / pipeline_tile2(s->kt) + F1);
  hipError_t e = alloc_pipeline_side(b->side, s->kt, len, F1, F2, s->cus, true, tiles, b->allocs);
  if (e != hipSuccess) {
    free_all(b->allocs);
    return hip_status(e, "hipMalloc");
  }
  PipelineDevice d = base_device(s, *b);
  d.build.keys = keys;
  d.build.pays = payloads;
  d.build.n = len;
  d.G1 = chunks_for(s->kt, len, s->cus);
  d.sides = 1;
  d.work_counter = s->scratch_counters;
  d.aggregates = s->scratch_agg;
  hipStream_t stm = s->setup;
  if ((e = enqueue_side_prefix_sum(d, d.build, s->kt, 1024, stm)) != hipSuccess ||
      (e = enqueue_side_scatter(d, d.build, s->kt, 1024, s->lds_total, stm)) != hipSuccess ||
      (e = launch_pipeline_plan(d, stm)) != hipSuccess ||
      (e = launch_pipeline_sort2(d, s->kt, std::max(1u, std::min(s->cus, tiles)), stm)) != hipSuccess ||
      (e = hipStreamSynchronize(stm)) != hipSuccess) {
    free_all(b->allocs);
    return hip_status(e, "partitioning the build side");
  }
  // the join reads the pass-2 output only: release pass 1's buffers
  std::vector<void*> keep;
  for (void* p : b->allocs) {
    if (p == b->side.rel1 || p == b->side.local1 || p == b->side.counts1 || p == b->side.scan1)
      (void)hipFree(p);
    else
      keep.push_back(p);
  }
  b->allocs.swap(keep);
  b->side.rel1 = nullptr;
  b->side.local1 = nullptr;
  b->side.counts1 = nullptr;
  b->side.scan1 = nullptr;
  b->side.keys = nullptr;
  b->side.pays = nullptr;
  return register_build(s, std::move(b), build_id);
}

fij_status fij_join_server_submit(fij_join_server* s, uint32_t build_id, const void* probe_keys,
                                  const void* probe_payloads, uint64_t probe_len, uint64_t* request_id) {
  if (!s || !request_id || (probe_len && (!probe_keys || !probe_payloads)))
    return server_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const size_t kb = key_bytes(s->kt);
  if (reinterpret_cast<uintptr_t>(probe_keys) % kb || reinterpret_cast<uintptr_t>(probe_payloads) % kb)
    return server_fail(FIJ_INVALID_ARGUMENT, "columns must be aligned to their element size");
  const auto now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (build_id >= s->builds.size()) return server_fail(FIJ_INVALID_ARGUMENT, "unknown build side %u", build_id);
    uint64_t cap = 0;
    for (const Slot& sl : s->slots) cap = std::max(cap, sl.probe_cap);
    if (probe_len > cap)
      return server_fail(FIJ_INVALID_ARGUMENT, "probe side of %llu tuples exceeds every slot",
                         static_cast<unsigned long long>(probe_len));
    Request r;
    r.id = s->next_id++;
    r.build = build_id;
    r.keys = probe_keys;
    r.pays = probe_payloads;
    r.len = probe_len;
    r.arrival = now;
    if (probe_len == 0 || s->builds[build_id]->n == 0) {  // no tuple can join
      r.start = r.end = now;
      r.done = true;
      s->requests[r.id] = r;
    } else {
      s->requests[r.id] = r;
      s->queue.push_back(r.id);
    }
    *request_id = r.id;
  }
  s->cv.notify_all();
  return FIJ_OK;
}

fij_status fij_join_server_wait(fij_join_server* s, uint64_t request_id, fij_join_aggregate* result,
                                fij_join_request_times* times) {
  if (!s) return server_fail(FIJ_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return server_fail(FIJ_INVALID_ARGUMENT, "unknown request");
  s->cv.wait(lk, [&] { return it->second.done; });
  const Request r = it->second;
  s->requests.erase(it);
  lk.unlock();
  if (r.status != FIJ_OK) return r.status;
  if (result) *result = r.result;
  if (times) {
    times->queued_ms = std::chrono::duration<double, std::milli>(r.start - r.arrival).count();
    times->latency_ms = std::chrono::duration<double, std::milli>(r.end - r.arrival).count();
  }
  return FIJ_OK;
}

fij_status fij_join_server_memory(const fij_join_server* s, uint64_t* used_bytes, uint32_t* builds) {
  if (!s || !used_bytes) return server_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  // add_build and the scheduler change both under the lock
  std::lock_guard<std::mutex> lock(s->mu);
  *used_bytes = s->budget_used;
  if (builds) *builds = static_cast<uint32_t>(s->builds.size());
  return FIJ_OK;
}

}  // extern "C"
