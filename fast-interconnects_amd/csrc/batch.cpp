// Batched execution of many independent joins (fij_join_batch_*).
//
// The reference joins one pair of relations per run of the radix-join binary
// (radix-join/src/execution_methods/gpu_radix_join.rs:90-565), allocating and
// launching per join.  A batch joins n pairs with one submission:
//   * joins whose build side fits one block's LDS table ("small") go to one
//     persistent launch that joins them without partitioning (batch.hip);
//   * every other join runs the two-pass radix plan of its size class, with
//     radix bits chosen by fij_join_autotune; one pooled plan per class,
//     sized for the largest join seen in that class, is reused by all its
//     joins;
//   * the whole submission (the small-join launch and every plan's kernels)
//     is recorded into one HIP graph and launched once; results are written
//     to device slots and copied back with one copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "../../include/fij.h"
#include "batch.hpp"
#include "internal.hpp"

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

struct ClassPlan {
  fij_join_plan* plan = nullptr;
  uint64_t build_cap = 0, probe_cap = 0;
};

}  // namespace

struct fij_join_batch {
  fij_join_batch_config cfg;
  KeyType kt;
  uint32_t cus;
  SmallJob* d_jobs = nullptr;
  uint32_t* d_ticket = nullptr;
  JoinAggregate* d_results = nullptr;
  SmallJob* h_jobs = nullptr;
  JoinAggregate* h_results = nullptr;
  hipStream_t own = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::map<std::pair<uint32_t, uint32_t>, ClassPlan> plans;
};

extern "C" {

fij_status fij_join_autotune(fij_type type, uint64_t build_len, uint64_t probe_len, uint32_t* small,
                             uint32_t* radix_bits_1, uint32_t* radix_bits_2) {
  if ((type != FIJ_TYPE_I32 && type != FIJ_TYPE_I64) || !small || !radix_bits_1 || !radix_bits_2)
    return batch_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  const KeyType kt = type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  // one block joins a small pair: its build side fits one LDS table fill and
  // its probe side is short enough for one block to stream
  *small = build_len <= small_join_table_tuples(kt) && probe_len <= (1u << 22) ? 1u : 0u;
  // Size classes of the radix plan: at most 512 build tuples per (pass-1,
  // pass-2) partition pair on average, so that pairs fit the join's
  // registers (at most 640), split evenly over the passes (pass 1 at most 9
  // bits: its LDS line buffers), and enough pass-1 partitions that a pair's
  // probe side has at most 64 pass-2 tiles of 8192 tuples.
  uint32_t b = std::max<uint32_t>(2, ceil_log2((build_len + 511) / 512));
  uint32_t b1 = std::min<uint32_t>(9, (b + 1) / 2);
  const uint32_t need = ceil_log2(probe_len) > 19 ? ceil_log2(probe_len) - 19 : 0;
  b1 = std::min<uint32_t>(9, std::max(b1, need));
  b = std::max(b, b1 + 1);
  *radix_bits_1 = b1;
  *radix_bits_2 = std::min<uint32_t>(12, b - b1);
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
  for (auto& kv : b->plans) fij_join_plan_destroy(kv.second.plan);
  (void)hipFree(b->d_jobs);
  (void)hipFree(b->d_ticket);
  (void)hipFree(b->d_results);
  if (b->h_jobs) (void)hipHostFree(b->h_jobs);
  if (b->h_results) (void)hipHostFree(b->h_results);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
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
  const fij_type type = static_cast<fij_type>(b->cfg.type);
  const size_t kb = key_bytes(b->kt);
  // classify
  uint32_t n_small = 0;
  std::vector<std::pair<std::pair<uint32_t, uint32_t>, uint32_t>> large;  // (class, join index)
  std::map<std::pair<uint32_t, uint32_t>, std::pair<uint64_t, uint64_t>> need;
  for (uint32_t i = 0; i < n; ++i) {
    const fij_join_desc& j = joins[i];
    for (const void* c : {j.build_keys, j.build_payloads, j.probe_keys, j.probe_payloads})
      if (c && reinterpret_cast<uintptr_t>(c) % kb)
        return batch_fail(FIJ_INVALID_ARGUMENT, "join %u: columns must be aligned to their element size", i);
    if ((j.build_len && (!j.build_keys || !j.build_payloads)) || (j.probe_len && (!j.probe_keys || !j.probe_payloads)))
      return batch_fail(FIJ_INVALID_ARGUMENT, "join %u: NULL input column", i);
    uint32_t small = 0, b1 = 0, b2 = 0;
    fij_status s = fij_join_autotune(type, j.build_len, j.probe_len, &small, &b1, &b2);
    if (s != FIJ_OK) return s;
    if (small || j.probe_len == 0) {
      SmallJob& sj = b->h_jobs[n_small++];
      sj.build_keys = j.build_keys;
      sj.build_pays = j.build_payloads;
      sj.build_len = j.build_len;
      sj.probe_keys = j.probe_keys;
      sj.probe_pays = j.probe_payloads;
      sj.probe_len = j.probe_len;
      sj.result = b->d_results + i;
    } else {
      large.push_back({{b1, b2}, i});
      auto& m = need[{b1, b2}];
      m.first = std::max(m.first, j.build_len);
      m.second = std::max(m.second, j.probe_len);
    }
  }
  // pooled plans: grow a class's plan when a join exceeds it
  for (auto& kv : need) {
    ClassPlan& cp = b->plans[kv.first];
    if (cp.plan && cp.build_cap >= kv.second.first && cp.probe_cap >= kv.second.second) continue;
    if (cp.plan) {
      BATCH_HIP(hipStreamSynchronize(st));
      fij_join_plan_destroy(cp.plan);
      cp.plan = nullptr;
    }
    fij_join_plan_config pc{};
    pc.type = b->cfg.type;
    pc.radix_bits_1 = kv.first.first;
    pc.radix_bits_2 = kv.first.second;
    pc.build_len = std::max(cp.build_cap, kv.second.first);
    pc.probe_len = std::max(cp.probe_cap, kv.second.second);
    fij_status s = fij_join_plan_create(&pc, &cp.plan);
    if (s != FIJ_OK) return s;
    cp.build_cap = pc.build_len;
    cp.probe_cap = pc.probe_len;
  }
  if (n_small) {
    BATCH_HIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, n_small * sizeof(SmallJob), hipMemcpyHostToDevice, st));
    BATCH_HIP(hipStreamSynchronize(st));
  }
  // record the submission (or launch it directly)
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
  for (size_t k = 0; k < large.size() && s == FIJ_OK; ++k) {
    const fij_join_desc& j = joins[large[k].second];
    s = fij_join_plan_enqueue_sized(b->plans[large[k].first].plan, j.build_keys, j.build_payloads, j.build_len,
                                    j.probe_keys, j.probe_payloads, j.probe_len,
                                    reinterpret_cast<fij_join_aggregate*>(b->d_results + large[k].second), st);
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
    stats->partitioned_joins = static_cast<uint32_t>(large.size());
    stats->size_classes = static_cast<uint32_t>(need.size());
    stats->graph = use_graph ? 1u : 0u;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return FIJ_OK;
}

fij_status fij_join_batch_device_bytes(const fij_join_batch* b, uint64_t* bytes) {
  if (!b || !bytes) return batch_fail(FIJ_INVALID_ARGUMENT, "NULL argument");
  uint64_t total = static_cast<uint64_t>(b->cfg.max_joins) * (sizeof(SmallJob) + sizeof(JoinAggregate)) + 256;
  for (const auto& kv : b->plans) {
    uint64_t x = 0;
    if (fij_join_plan_device_bytes(kv.second.plan, &x) == FIJ_OK) total += x;
  }
  *bytes = total;
  return FIJ_OK;
}

}  // extern "C"
