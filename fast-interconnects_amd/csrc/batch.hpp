// Batched joins: device descriptors and launchers shared by batch.hip and
// batch.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace fij {

// One small join of a batch (device memory): input columns and the device
// slot its aggregate is written to.
struct SmallJob {
  const void* build_keys;
  const void* build_pays;
  uint64_t build_len;
  const void* probe_keys;
  const void* probe_pays;
  uint64_t probe_len;
  JoinAggregate* result;
};

// Build tuples one small-join table fill holds (a larger build side is
// joined in several fills).
uint64_t small_join_table_tuples(KeyType t);
// One persistent launch joining jobs[0..n_jobs) (`ticket` is cleared by a
// kernel first, so the launch can be recorded into a graph).
hipError_t launch_small_joins(KeyType t, const SmallJob* jobs, uint32_t n_jobs, uint32_t* ticket, uint32_t grid,
                              hipStream_t stream);

}  // namespace fij
