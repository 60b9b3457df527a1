// Out-of-core radix join (fij_ooc_join): relations in host memory, pass-1
// partitions cached in device memory as far as a budget allows and spilled
// to pinned host memory beyond it, pass 2 and the join streamed over groups
// of pass-1 partitions on two streams.
//
// Follows the reference's Triton join (radix-join/src/execution_methods/
// gpu_triton_join.rs:99-638: pass 1 on the GPU reading the relations over
// the interconnect, its output cached in GPU memory and spilled to CPU
// memory, `max_partitions_cache_bytes` split between the relations in
// proportion to their lengths, then per pass-1 partition a transform,
// pass 2 and the join on two streams) and, with pass 1 on the CPU, its
// CPU-partitioned radix join (cpu_partitioned_radix_join.rs:118-769).
// Differences by design:
//   * the cache boundary is a partition boundary chosen after the prefix
//     sum (the contiguous pass-1 layout makes the cached partitions one
//     range), where the reference splits the allocation at a byte count;
//   * spilled partitions are moved by DMA copies into a staging buffer, a
//     group of consecutive partitions per copy, instead of being read by the
//     transform kernel over the interconnect;
//   * a group of pass-1 partitions goes through the same plan, pass-2 and
//     join kernels as the in-memory execution method, a fixed number of
//     launches per group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/fij.h"
#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {
fij_status set_error(fij_status s, const char* msg);
}

using namespace fij;

namespace {

fij_status ooc_fail(fij_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(s, buf);
}

fij_status ooc_hip(hipError_t e, const char* what) {
  return ooc_fail(e == hipErrorOutOfMemory ? FIJ_OUT_OF_MEMORY : FIJ_DEVICE_ERROR, "%s: %s", what,
                  hipGetErrorString(e));
}

#define OOC_HIP(expr)                                 \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return ooc_hip(e_, #expr);  \
  } while (0)

// Device and pinned host allocations released when the join returns.
struct Arena {
  std::vector<void*> dev, host;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  ~Arena() {
    for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
    for (void* p : dev) (void)hipFree(p);
    for (void* p : host) (void)hipHostFree(p);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
  }
  template <typename T>
  hipError_t dalloc(T** p, uint64_t count) {
    void* v = nullptr;
    const hipError_t e = hipMalloc(&v, (count ? count : 1) * sizeof(T));
    if (e == hipSuccess) {
      dev.push_back(v);
      *p = static_cast<T*>(v);
    }
    return e;
  }
  template <typename T>
  hipError_t halloc(T** p, uint64_t count) {
    void* v = nullptr;
    const hipError_t e = hipHostMalloc(&v, (count ? count : 1) * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) {
      host.push_back(v);
      *p = static_cast<T*>(v);
    }
    return e;
  }
};

// One relation's pass-1 output: tuple indices [0, cut) in device memory
// (`cache`), [cut, padded_len) in pinned host memory (`spill`, element 0
// standing for index cut).
struct Pass1 {
  uint64_t n = 0, padded_len = 0, cut = 0;
  uint32_t cut_part = 0;  // first spilled partition
  std::vector<uint64_t> off;  // host copy of the partition starts
  uint8_t* cache = nullptr;
  uint8_t* spill = nullptr;
  uint64_t start(uint32_t p) const { return off[p] - pad; }
  uint64_t end(uint32_t p) const { return p + 1 < off.size() ? off[p + 1] - pad : padded_len; }
  uint64_t pad = 0;
};

struct Group {
  uint32_t p0, p1;
  uint64_t shift[2], len[2];  // first tuple index and tuple count per relation
  uint64_t off_at;            // position of its local offsets in the pinned offsets array
};

bool device_accessible(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.devicePointer != nullptr;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" fij_status fij_ooc_join(const fij_ooc_config* cfg, const void* build_keys, const void* build_pays,
                                   uint64_t build_len, const void* probe_keys, const void* probe_pays,
                                   uint64_t probe_len, fij_join_aggregate* result, fij_ooc_timing* timing) {
  if (!cfg || !result || (cfg->type != FIJ_TYPE_I32 && cfg->type != FIJ_TYPE_I64))
    return ooc_fail(FIJ_INVALID_ARGUMENT, "invalid argument");
  if (cfg->radix_bits_1 < 1 || cfg->radix_bits_1 > 12 || cfg->radix_bits_2 < 1 || cfg->radix_bits_2 > 12)
    return ooc_fail(FIJ_INVALID_ARGUMENT, "radix bits %u+%u: each pass takes 1 to 12 bits", cfg->radix_bits_1,
                    cfg->radix_bits_2);
  if ((build_len && (!build_keys || !build_pays)) || (probe_len && (!probe_keys || !probe_pays)))
    return ooc_fail(FIJ_INVALID_ARGUMENT, "NULL input column");
  const KeyType kt = cfg->type == FIJ_TYPE_I32 ? KeyType::I32 : KeyType::I64;
  const uint32_t tb = tuple_bytes(kt);
  const bool cpu_pass1 = cfg->pass1_cpu != 0;
  const void* cols[2][2] = {{build_keys, build_pays}, {probe_keys, probe_pays}};
  const uint64_t lens[2] = {build_len, probe_len};
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) {
      if (!lens[r]) continue;
      if (reinterpret_cast<uintptr_t>(cols[r][c]) % key_bytes(kt))
        return ooc_fail(FIJ_INVALID_ARGUMENT, "input columns must be aligned to their element size");
      if (!cpu_pass1 && !device_accessible(cols[r][c]))
        return ooc_fail(FIJ_INVALID_ARGUMENT,
                        "pass 1 on the GPU reads the relations in place: they must be device memory or pinned "
                        "host memory");
    }
  *result = fij_join_aggregate{0, 0, 0};
  if (timing) memset(timing, 0, sizeof(*timing));
  if (build_len == 0 || probe_len == 0) return FIJ_OK;

  int dev = 0;
  OOC_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  OOC_HIP(hipGetDeviceProperties(&prop, dev));
  const uint32_t cus = static_cast<uint32_t>(prop.multiProcessorCount);
  uint32_t lds_total = static_cast<uint32_t>(prop.sharedMemPerBlock);
  int optin = 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && optin > 0)
    lds_total = std::max<uint32_t>(lds_total, static_cast<uint32_t>(optin));
  const uint32_t bits1 = cfg->radix_bits_1, bits2 = cfg->radix_bits_2;
  const uint32_t F1 = 1u << bits1, F2 = 1u << bits2;
  if (pipeline_sort2_lds(kt, F2) > lds_total)
    return ooc_fail(FIJ_UNSUPPORTED, "%u pass-2 bits need more LDS than a block has", bits2);
  const uint64_t pad = pad_tuples(kt);
  const uint32_t join_grid = 2 * cus;

  Arena A;
  hipStream_t main_st;
  OOC_HIP(hipStreamCreateWithFlags(&main_st, hipStreamNonBlocking));
  A.streams.push_back(main_st);
  hipEvent_t ev[4];
  for (hipEvent_t& e : ev) {
    OOC_HIP(hipEventCreate(&e));
    A.events.push_back(e);
  }
  const auto t_start = std::chrono::steady_clock::now();
  Pass1 rel[2];
  double prefix_ms = 0, partition_ms = 0;

  // ---- pass 2 + join over groups of consecutive pass-1 partitions -------
  // A group holds at most `cap` tuples of either relation (a partition
  // longer than that gets a group of its own); groups end at either
  // relation's cache boundary, so a relation's part of a group is wholly
  // cached (read in place) or wholly spilled (copied).
  auto make_groups = [&](const Pass1& R, const Pass1& S, uint64_t cap, std::vector<Group>& groups) {
    groups.clear();
    for (uint32_t p = 0; p < F1;) {
      Group g{};
      g.p0 = p;
      uint32_t q = p + 1;
      while (q < F1 && q != R.cut_part && q != S.cut_part) {
        if (R.end(q) - R.start(p) > cap || S.end(q) - S.start(p) > cap) break;
        ++q;
      }
      g.p1 = q;
      groups.push_back(g);
      p = q;
    }
  };
  // every group's partition starts relative to the group, both relations
  // (2 F1 words in all)
  auto group_offsets = [&](const Pass1& R, const Pass1& S, std::vector<Group>& groups, uint64_t* h_goff) {
    uint64_t at = 0;
    const Pass1* rs[2] = {&R, &S};
    for (Group& g : groups) {
      g.off_at = at;
      for (int r = 0; r < 2; ++r) {
        g.shift[r] = rs[r]->start(g.p0);
        g.len[r] = rs[r]->end(g.p1 - 1) - g.shift[r];
        for (uint32_t p = g.p0; p < g.p1; ++p) h_goff[at++] = rs[r]->off[p] - g.shift[r];
      }
    }
  };
  auto longest = [&](const Pass1& r) {
    uint64_t m = 0;
    for (uint32_t p = 0; p < F1; ++p) m = std::max(m, r.end(p) - r.start(p));
    return m;
  };

  // two streams, each with the buffers of one group
  constexpr int NS = 2;
  const uint32_t tile2 = pipeline_tile2(kt);
  PipelineDevice sd[NS];
  hipStream_t st[NS];
  for (int s = 0; s < NS; ++s) {
    OOC_HIP(hipStreamCreateWithFlags(&st[s], hipStreamNonBlocking));
    A.streams.push_back(st[s]);
  }
  std::vector<void*> stream_allocs;
  struct StreamBuffers {  // freed when the join returns, before the streams
    std::vector<void*>& allocs;
    hipStream_t* st;
    ~StreamBuffers() {
      for (int s = 0; s < NS; ++s) (void)hipStreamSynchronize(st[s]);
      for (void* p : allocs) (void)hipFree(p);
    }
  } stream_buffers{stream_allocs, st};
  uint64_t cap = 0;
  // (re)allocates both streams' buffers for groups of up to `want` tuples
  auto size_streams = [&](uint64_t want) -> fij_status {
    if (want <= cap) return FIJ_OK;
    if (want + F1 * pad >= 0xFFFFFFFFull) return ooc_fail(FIJ_INTEGER_OVERFLOW, "a partition group exceeds 2^32 tuples");
    for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamSynchronize(st[s]));
    for (void* q : stream_allocs) OOC_HIP(hipFree(q));
    stream_allocs.clear();
    cap = want;
    const uint32_t max_tiles = static_cast<uint32_t>(cap / tile2 + F1 + 1);
    auto dal = [&](auto** ptr, uint64_t count) -> hipError_t {
      void* v = nullptr;
      const hipError_t e = hipMalloc(&v, std::max<uint64_t>(count, 1) * sizeof(**ptr));
      if (e == hipSuccess) {
        stream_allocs.push_back(v);
        *ptr = static_cast<std::remove_reference_t<decltype(*ptr)>>(v);
      }
      return e;
    };
    for (int s = 0; s < NS; ++s) {
      PipelineDevice& d = sd[s];
      d = PipelineDevice{};
      d.F2 = F2;
      d.bits1 = bits1;
      d.bits2 = bits2;
      d.pad = static_cast<uint32_t>(pad);
      d.tile2 = tile2;
      d.sides = 3;
      d.agg_len = join_grid;
      for (PipelineSide* ps : {&d.build, &d.probe}) {
        std::vector<void*> allocs;
        const hipError_t e = alloc_pipeline_side(*ps, kt, cap, F1, F2, 1, true, max_tiles, allocs);
        stream_allocs.insert(stream_allocs.end(), allocs.begin(), allocs.end());
        OOC_HIP(e);
      }
      OOC_HIP(dal(&d.work_counter, kJoinCounterWords));
      const uint64_t items = static_cast<uint64_t>(F1) * F2;
      OOC_HIP(dal(&d.overflow, items));
      // the deferred-item flags start at zero; every join leaves them at zero
      OOC_HIP(hipMemsetAsync(d.overflow, 0, items, st[s]));
      d.unit_cap = static_cast<uint32_t>(items + 2 * cap / 65536 + 2);
      OOC_HIP(dal(&d.units, d.unit_cap));
      OOC_HIP(dal(&d.aggregates, join_grid));
    }
    for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamSynchronize(st[s]));
    return FIJ_OK;
  };
  // Enqueues pass 2 and the join of every group, alternating the streams;
  // group gi's result goes to results[gi].
  int rr = 0;
  auto enqueue_groups = [&](const Pass1& R, const Pass1& S, const std::vector<Group>& groups, const uint64_t* h_goff,
                            JoinAggregate* results) -> fij_status {
    const Pass1* rs[2] = {&R, &S};
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const Group& g = groups[gi];
      const int s = rr++ % NS;
      PipelineDevice d = sd[s];
      d.F1 = g.p1 - g.p0;
      uint64_t at_r = g.off_at;
      PipelineSide* sides[2] = {&d.build, &d.probe};
      for (int r = 0; r < 2; ++r) {
        PipelineSide& ps = *sides[r];
        const Pass1& src = *rs[r];
        ps.padded_len1 = g.len[r];
        ps.n = g.len[r];
        if (g.shift[r] + g.len[r] <= src.cut) {
          ps.rel1 = src.cache + g.shift[r] * tb;  // cached: read in place
        } else {
          // spilled: one copy into this stream's staging buffer
          uint8_t* dst = static_cast<uint8_t*>(r == 0 ? sd[s].build.rel1 : sd[s].probe.rel1);
          OOC_HIP(hipMemcpyAsync(dst, src.spill + (g.shift[r] - src.cut) * tb, g.len[r] * tb, hipMemcpyHostToDevice,
                                 st[s]));
          ps.rel1 = dst;
        }
        OOC_HIP(hipMemcpyAsync(ps.off1, h_goff + at_r, 8ull * d.F1, hipMemcpyHostToDevice, st[s]));
        at_r += d.F1;
      }
      OOC_HIP(launch_pipeline_plan(d, st[s]));
      const uint32_t tiles = static_cast<uint32_t>((g.len[0] + g.len[1]) / tile2 + 2 * d.F1);
      OOC_HIP(launch_pipeline_sort2(d, kt, std::max(1u, std::min(cus, tiles)), st[s]));
      OOC_HIP(launch_pipeline_join(d, kt, false, false, join_grid, st[s]));
      OOC_HIP(launch_pipeline_join(d, kt, false, true, join_grid, st[s]));
      OOC_HIP(launch_sum_aggregates(d.aggregates, join_grid, results + gi, st[s]));
    }
    return FIJ_OK;
  };
  auto add_results = [&](const JoinAggregate* d_res, size_t n) -> fij_status {
    std::vector<JoinAggregate> h(n);
    if (n) OOC_HIP(hipMemcpy(h.data(), d_res, sizeof(JoinAggregate) * n, hipMemcpyDeviceToHost));
    for (const JoinAggregate& a : h) {
      result->count += a.count;
      result->sum_build += a.sum_build;
      result->sum_probe += a.sum_probe;
    }
    return FIJ_OK;
  };
  const uint64_t staging = std::max<uint64_t>(cfg->staging_tuples ? cfg->staging_tuples : (1ull << 24), 1ull << 16);
  double join_ms = 0;
  uint64_t n_groups = 0;

  if (cpu_pass1) {
    // Pass 1 on host threads, overlapped with the GPU as in the reference's
    // CPU-partitioned join (cpu_partitioned_radix_join.rs:457-547: the build
    // side's transfer and pass 2 run on the GPU while the CPU partitions the
    // probe side):
    //   1. the build relation R is partitioned into pinned host memory;
    //   2. its first partitions, up to cache_bytes, are copied to device
    //      memory by one DMA copy while
    //   3. the probe relation S is partitioned in slices of its input, each
    //      into its own contiguous pass-1 layout, and every slice is joined
    //      with R on the GPU while the host threads partition the next one
    //      (Join(R, S) is the sum of the Join(R, S_k)).
    const uint32_t threads = cfg->cpu_threads ? cfg->cpu_threads : std::max(1u, std::thread::hardware_concurrency());
    double cpu_ms = 0;
    auto t0 = std::chrono::steady_clock::now();
    Pass1& R = rel[0];
    R.n = build_len;
    R.pad = pad;
    R.padded_len = R.n + F1 * pad;
    R.off.assign(F1, 0);
    uint8_t* r_host = nullptr;
    OOC_HIP(A.halloc(&r_host, R.padded_len * tb));
    cpu_partition_contiguous(kt, build_keys, build_pays, R.n, bits1, 0, threads, r_host, R.off.data());
    cpu_ms += ms_since(t0);
    uint32_t P = 0;
    while (P < F1 && R.end(P) * tb <= cfg->cache_bytes) ++P;
    R.cut_part = P;
    R.cut = P < F1 ? R.start(P) : R.padded_len;
    R.spill = r_host + R.cut * tb;  // stands for tuple index R.cut
    if (R.cut) {
      // one spare tuple: pass 2 reads 8-byte tuples in aligned pairs
      OOC_HIP(A.dalloc(&R.cache, (R.cut + 1) * tb));
      OOC_HIP(hipMemcpyAsync(R.cache, r_host, R.cut * tb, hipMemcpyHostToDevice, main_st));
      OOC_HIP(hipEventRecord(ev[3], main_st));
      for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamWaitEvent(st[s], ev[3], 0));
    }
    // probe slices of about `staging` tuples, at least 1, at most 8
    uint64_t K = cfg->probe_slices ? cfg->probe_slices : std::min<uint64_t>(8, std::max<uint64_t>(1, probe_len / staging));
    K = std::max<uint64_t>(1, std::min<uint64_t>(K, probe_len));
    const uint64_t slice_len = (probe_len + K - 1) / K;
    // two slice buffers: slice k + 1 is partitioned while slice k's copies
    // and kernels run
    Pass1 slice[2];
    uint64_t* h_goff[2] = {nullptr, nullptr};
    hipEvent_t slice_free[2];
    for (int b = 0; b < 2; ++b) {
      OOC_HIP(A.halloc(&slice[b].spill, (slice_len + F1 * pad) * tb));
      OOC_HIP(A.halloc(&h_goff[b], 2ull * F1));
      OOC_HIP(hipEventCreate(&slice_free[b]));
      A.events.push_back(slice_free[b]);
    }
    JoinAggregate* d_res = nullptr;
    OOC_HIP(A.dalloc(&d_res, K * F1));
    std::vector<Group> groups;
    const size_t kb = key_bytes(kt);
    for (uint64_t k = 0; k < K; ++k) {
      const uint64_t b0 = std::min<uint64_t>(probe_len, k * slice_len);
      const uint64_t n = std::min<uint64_t>(probe_len, b0 + slice_len) - b0;
      Pass1& S = slice[k & 1];
      if (k >= 2) OOC_HIP(hipEventSynchronize(slice_free[k & 1]));  // slice k - 2 has left this buffer
      t0 = std::chrono::steady_clock::now();
      S.n = n;
      S.pad = pad;
      S.padded_len = n + F1 * pad;
      S.off.assign(F1, 0);
      S.cut = 0;
      S.cut_part = 0;
      cpu_partition_contiguous(kt, static_cast<const uint8_t*>(probe_keys) + b0 * kb,
                               static_cast<const uint8_t*>(probe_pays) + b0 * kb, n, bits1, 0, threads, S.spill,
                               S.off.data());
      cpu_ms += ms_since(t0);
      fij_status fs = size_streams(std::max({staging, longest(R), longest(S)}));
      if (fs != FIJ_OK) return fs;
      make_groups(R, S, cap, groups);
      group_offsets(R, S, groups, h_goff[k & 1]);
      if (k + 1 == K) t0 = std::chrono::steady_clock::now();
      fs = enqueue_groups(R, S, groups, h_goff[k & 1], d_res + n_groups);
      if (fs != FIJ_OK) return fs;
      n_groups += groups.size();
      // the buffer is free again once both streams have passed this point
      for (int s = 0; s < NS; ++s) OOC_HIP(hipEventRecord(ev[s], st[s]));
      for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamWaitEvent(main_st, ev[s], 0));
      OOC_HIP(hipEventRecord(slice_free[k & 1], main_st));
    }
    for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamSynchronize(st[s]));
    join_ms = ms_since(t0);  // the last slice's transfers and kernels, after the host threads finished
    partition_ms = cpu_ms;
    fij_status fs = add_results(d_res, n_groups);
    if (fs != FIJ_OK) return fs;
  } else {
    // pass 1 on the GPU: histogram and offsets, then the cache split, then
    // the scatter into cache and spill
    PipelineDevice d{};
    d.F1 = F1;
    d.bits1 = bits1;
    d.pad = static_cast<uint32_t>(pad);
    PipelineSide side[2]{};
    uint32_t G1 = cus;
    auto fits = [&](uint64_t n, uint32_t chunks) {
      uint64_t c = 0;
      return fij_input_chunk_size(static_cast<fij_type>(cfg->type), n, chunks, &c) == FIJ_OK && c * (chunks - 1) <= n;
    };
    while (G1 > 1 && !(fits(build_len, G1) && fits(probe_len, G1))) G1 = G1 > 64 ? G1 - 32 : G1 - 1;
    d.G1 = G1;
    for (int r = 0; r < 2; ++r) {
      PipelineSide& s = side[r];
      s.keys = cols[r][0];
      s.pays = cols[r][1];
      s.n = lens[r];
      s.padded_len1 = s.n + F1 * pad;
      s.scan1_len = scan_state_words(static_cast<uint64_t>(F1) * G1);
      OOC_HIP(A.dalloc(&s.off1, F1));
      OOC_HIP(A.dalloc(&s.local1, static_cast<uint64_t>(G1) * F1));
      OOC_HIP(A.dalloc(&s.counts1, static_cast<uint64_t>(F1) * G1));
      OOC_HIP(A.dalloc(&s.scan1, s.scan1_len));
    }
    OOC_HIP(hipEventRecord(ev[0], main_st));
    for (int r = 0; r < 2; ++r) OOC_HIP(enqueue_side_prefix_sum(d, side[r], kt, 1024, main_st));
    OOC_HIP(hipEventRecord(ev[1], main_st));
    uint64_t* h_off = nullptr;
    OOC_HIP(A.halloc(&h_off, 2ull * F1));
    for (int r = 0; r < 2; ++r)
      OOC_HIP(hipMemcpyAsync(h_off + r * F1, side[r].off1, F1 * 8ull, hipMemcpyDeviceToHost, main_st));
    OOC_HIP(hipStreamSynchronize(main_st));
    float f = 0;
    OOC_HIP(hipEventElapsedTime(&f, ev[0], ev[1]));
    prefix_ms = f;
    // the cache budget, split in proportion to the relations' lengths
    const double share = static_cast<double>(build_len) / static_cast<double>(build_len + probe_len);
    const uint64_t cache_tuples[2] = {static_cast<uint64_t>(cfg->cache_bytes * share) / tb,
                                      (cfg->cache_bytes - static_cast<uint64_t>(cfg->cache_bytes * share)) / tb};
    for (int r = 0; r < 2; ++r) {
      Pass1& s = rel[r];
      s.n = lens[r];
      s.pad = pad;
      s.padded_len = s.n + F1 * pad;
      s.off.assign(h_off + r * F1, h_off + (r + 1) * F1);
      uint32_t P = 0;
      while (P < F1 && s.end(P) <= cache_tuples[r]) ++P;
      s.cut_part = P;
      s.cut = P < F1 ? s.start(P) : s.padded_len;
      if (s.cut) OOC_HIP(A.dalloc(&s.cache, (s.cut + 1) * tb));  // + the spare tuple of pass 2
      if (s.cut < s.padded_len) {
        // the spill region keeps the relation's alignment modulo 128 bytes,
        // so the scatter's full lines stay aligned there too
        uint8_t* raw = nullptr;
        OOC_HIP(A.halloc(&raw, (s.padded_len - s.cut) * tb + 128));
        s.spill = raw + (s.cut * tb) % 128;
      }
    }
    for (int r = 0; r < 2; ++r) {
      PartitionLaunch a{};
      a.type = kt;
      a.keys = cols[r][0];
      a.pays = cols[r][1];
      a.n = lens[r];
      a.chunks = G1;
      a.radix_bits = bits1;
      a.ignore_bits = 0;
      a.chunk_offsets = side[r].local1;
      // with everything spilled the relation pointer only anchors the
      // alignment: a 256-byte aligned device address
      a.relation = rel[r].cache ? static_cast<void*>(rel[r].cache) : static_cast<void*>(side[r].off1);
      a.block = 1024;
      a.lds_bytes = lds_total;
      a.buffered = true;
      a.spill = rel[r].spill;
      a.spill_base = rel[r].cut;
      a.spill_from = rel[r].cut_part;
      OOC_HIP(launch_partition(a, main_st));
    }
    OOC_HIP(hipEventRecord(ev[2], main_st));
    OOC_HIP(hipStreamSynchronize(main_st));
    OOC_HIP(hipEventElapsedTime(&f, ev[1], ev[2]));
    partition_ms = f;
    fij_status fs = size_streams(std::max({staging, longest(rel[0]), longest(rel[1])}));
    if (fs != FIJ_OK) return fs;
    std::vector<Group> groups;
    make_groups(rel[0], rel[1], cap, groups);
    uint64_t* h_goff = nullptr;
    OOC_HIP(A.halloc(&h_goff, 2ull * F1));
    group_offsets(rel[0], rel[1], groups, h_goff);
    JoinAggregate* d_res = nullptr;
    OOC_HIP(A.dalloc(&d_res, groups.size()));
    const auto t_join = std::chrono::steady_clock::now();
    fs = enqueue_groups(rel[0], rel[1], groups, h_goff, d_res);
    if (fs != FIJ_OK) return fs;
    for (int s = 0; s < NS; ++s) OOC_HIP(hipStreamSynchronize(st[s]));
    join_ms = ms_since(t_join);
    fs = add_results(d_res, groups.size());
    if (fs != FIJ_OK) return fs;
    n_groups = groups.size();
  }
  if (timing) {
    timing->prefix_sum_ms = static_cast<float>(prefix_ms);
    timing->partition_ms = static_cast<float>(partition_ms);
    timing->join_ms = static_cast<float>(join_ms);
    timing->total_ms = static_cast<float>(ms_since(t_start));
    timing->cached_build_tuples = rel[0].cut;
    timing->cached_probe_tuples = rel[1].cut;
    timing->groups = static_cast<uint32_t>(n_groups);
  }
  return FIJ_OK;
}
