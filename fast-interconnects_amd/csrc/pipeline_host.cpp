// Host helpers shared by the execution-method plan (plan.cpp) and the join
// server (server.cpp): one side's buffers and its pass-1 launches.
#include <hip/hip_runtime.h>

#include <vector>

#include "internal.hpp"
#include "pipeline.hpp"

namespace fij {

namespace {

template <typename T>
hipError_t alloc_n(T** out, uint64_t count, std::vector<void*>& allocs) {
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, (count ? count : 1) * sizeof(T));
  if (e != hipSuccess) return e;
  allocs.push_back(p);
  *out = static_cast<T*>(p);
  return hipSuccess;
}

}  // namespace

hipError_t alloc_pipeline_side(PipelineSide& s, KeyType t, uint64_t n, uint32_t F1, uint32_t F2, uint32_t G1,
                               bool two_pass, uint32_t tiles, std::vector<void*>& allocs) {
  const uint32_t tb = tuple_bytes(t);
  s.n = n;
  s.padded_len1 = n + static_cast<uint64_t>(F1) * pad_tuples(t);
  s.scan1_len = scan_state_words(static_cast<uint64_t>(F1) * G1);
  hipError_t e;
  if ((e = alloc_n(&s.off1, F1, allocs)) != hipSuccess) return e;
  if ((e = alloc_n(&s.local1, static_cast<uint64_t>(G1) * F1, allocs)) != hipSuccess) return e;
  if ((e = alloc_n(&s.counts1, static_cast<uint64_t>(F1) * G1, allocs)) != hipSuccess) return e;
  if ((e = alloc_n(&s.scan1, s.scan1_len, allocs)) != hipSuccess) return e;
  uint8_t* rel1 = nullptr;
  // one spare tuple after the end: pass 2 reads 8-byte tuples in aligned
  // pairs, and the last pair may end one tuple past the relation
  if ((e = alloc_n(&rel1, (s.padded_len1 + 1) * tb, allocs)) != hipSuccess) return e;
  s.rel1 = rel1;
  if ((e = alloc_n(&s.plen, F1, allocs)) != hipSuccess) return e;
  if ((e = alloc_n(&s.pstart, F1 + 1, allocs)) != hipSuccess) return e;
  if ((e = alloc_n(&s.tile_start, F1 + 1, allocs)) != hipSuccess) return e;
  if (two_pass) {
    uint8_t* rel2 = nullptr;
    if ((e = alloc_n(&rel2, n * tb, allocs)) != hipSuccess) return e;
    s.rel2 = rel2;
    if ((e = alloc_n(&s.meta, (static_cast<uint64_t>(tiles) + 1) * F2, allocs)) != hipSuccess) return e;
  } else {
    s.rel2 = nullptr;
    s.meta = nullptr;
  }
  return hipSuccess;
}

hipError_t enqueue_side_prefix_sum(const PipelineDevice& d, PipelineSide& s, KeyType t, uint32_t block,
                                   hipStream_t stream) {
  hipError_t e = launch_hist_counts(t, s.keys, nullptr, nullptr, nullptr, s.n, d.G1, d.bits1, 0, s.counts1, s.scan1,
                                    s.scan1_len, block, stream);
  if (e != hipSuccess) return e;
  return launch_offsets_scan(s.counts1, d.F1, d.G1, d.pad, s.local1, s.off1, s.scan1, stream);
}

hipError_t enqueue_side_scatter(const PipelineDevice& d, PipelineSide& s, KeyType t, uint32_t block,
                                uint32_t lds_total, hipStream_t stream) {
  PartitionLaunch a{};
  a.type = t;
  a.keys = s.keys;
  a.pays = s.pays;
  a.n = s.n;
  a.chunks = d.G1;
  a.radix_bits = d.bits1;
  a.ignore_bits = 0;
  a.chunk_offsets = s.local1;
  a.relation = s.rel1;
  a.block = block;
  a.lds_bytes = lds_total;
  a.buffered = true;
  return launch_partition(a, stream);
}

}  // namespace fij
