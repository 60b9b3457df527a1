// Seeded, counter-based relation generators (device side).
//
// The reference generates relations with unseeded thread_rng
// (datagen/src/relation.rs:76-290), so runs cannot be replayed.  These
// generators are this library's own definition: every element is a pure
// function of (seed, stream, index), so the GPU produces, in parallel, the
// same integers as a sequential CPU implementation of the same formulas.
//   mix64:      SplitMix64 finaliser
//   rand_u64:   mix64(mix64(seed ^ mix64(stream + G)) + (i + 1) * G),
//               G = 0x9e3779b97f4a7c15
//   permutation: 4-round balanced Feistel network on 2*ceil(bits/2) bits
//               with cycle walking (a bijection on [0, n))
//   uniform:    lo + mulhi64(rand_u64, range)
#include "common.hpp"
#include "internal.hpp"

namespace fij {
namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rand_u64(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t k = mix64(seed ^ mix64(stream + kGolden));
  return mix64(k + (i + 1) * kGolden);
}

__device__ uint64_t feistel_perm(uint64_t i, uint64_t n, uint64_t seed, uint32_t h, uint64_t mask,
                                 const uint64_t* round_keys) {
  uint64_t x = i;
  do {
    uint64_t l = x >> h, r = x & mask;
#pragma unroll
    for (int round = 0; round < 4; ++round) {
      const uint64_t f = mix64(r ^ round_keys[round]) & mask;
      const uint64_t nl = r;
      r = l ^ f;
      l = nl;
    }
    x = (l << h) | r;
  } while (x >= n);
  return x;
}

template <typename K>
__global__ void gen_perm_kernel(K* out, uint64_t n, uint64_t seed, int64_t base) {
  if (n <= 1) {
    if (n == 1 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = static_cast<K>(base);
    return;
  }
  const uint32_t b = 64u - __builtin_clzll(n - 1);
  const uint32_t h = (b + 1u) / 2u;
  const uint64_t mask = (1ull << h) - 1ull;
  uint64_t rk[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rk[r] = mix64(seed + (r + 1) * kGolden);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<K>(base + static_cast<int64_t>(feistel_perm(i, n, seed, h, mask, rk)));
}

// Zipf(s) over ranks 1..N by rejection-inversion (W. Hormann and G.
// Derflinger, "Rejection-inversion to generate variates from monotone
// discrete distributions", ACM TOMACS 6(3), 1996); H is the antiderivative of
// h(x) = x^-s written with log1p / expm1 forms for stability near s = 1.
// Attempt a of element i draws u from rand_u64(seed, stream, 64 i + a).
__device__ __forceinline__ double zh1(double x) {
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__device__ __forceinline__ double zh2(double x) {
  return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__device__ __forceinline__ double zipf_H(double x, double s) {
  const double lx = log(x);
  return zh2((1.0 - s) * lx) * lx;
}
__device__ __forceinline__ double zipf_Hinv(double y, double s) {
  double t = y * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zh1(t) * y);
}
__device__ __forceinline__ double zipf_h(double x, double s) { return exp(-s * log(x)); }

// Keys base + perm(rank - 1): the keys gen_perm_kernel makes for N keys with
// perm_seed, with Zipf frequencies (rank 1 the most frequent).
template <typename K>
__global__ void gen_zipf_kernel(K* out, uint64_t n, uint64_t N, double s, uint64_t seed, uint64_t stream_id,
                                uint64_t perm_seed, int64_t base) {
  const double Nd = static_cast<double>(N);
  const double hx1 = zipf_H(1.5, s) - 1.0;
  const double hN = zipf_H(Nd + 0.5, s);
  const double sc = 2.0 - zipf_Hinv(zipf_H(2.5, s) - zipf_h(2.0, s), s);
  const uint32_t b = N > 1 ? 64u - __builtin_clzll(N - 1) : 1u;
  const uint32_t h = (b + 1u) / 2u;
  const uint64_t mask = (1ull << h) - 1ull;
  uint64_t rk[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) rk[r] = mix64(perm_seed + (r + 1) * kGolden);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t k = 1;
    for (uint64_t a = 0; a < 64; ++a) {
      const double u01 = static_cast<double>(rand_u64(seed, stream_id, i * 64 + a) >> 11) * 0x1.0p-53;
      const double u = hN + u01 * (hx1 - hN);
      const double x = zipf_Hinv(u, s);
      double kd = floor(x + 0.5);
      if (kd < 1.0) kd = 1.0;
      if (kd > Nd) kd = Nd;
      k = static_cast<uint64_t>(kd);
      if (kd - x <= sc || u >= zipf_H(kd + 0.5, s) - zipf_h(kd, s)) break;
    }
    const uint64_t pos = N > 1 ? feistel_perm(k - 1, N, perm_seed, h, mask, rk) : 0;
    out[i] = static_cast<K>(base + static_cast<int64_t>(pos));
  }
}

template <typename K>
__global__ void gen_uniform_kernel(K* out, uint64_t n, uint64_t seed, uint64_t stream_id, int64_t lo,
                                   uint64_t range) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<K>(lo + static_cast<int64_t>(__umul64hi(rand_u64(seed, stream_id, i), range)));
}

template <typename K>
__global__ void gen_iota_kernel(K* out, uint64_t n, int64_t base) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<K>(base + static_cast<int64_t>(i));
}

template <typename U>
__global__ void fill_kernel(U* out, uint64_t n, U value) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = value;
}

// Plain device copy of 16-byte words: each block copies 4 * 256 consecutive
// words, every thread four independent non-temporal loads, then four
// stores.  The bandwidth reference the partition and join kernels are
// compared with, and the known-size kernel that calibrates the HBM byte
// counters.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                   uint64_t n) {
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 1024 + threadIdx.x;
  u32x4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k * 256 < n) v[k] = __builtin_nontemporal_load(src + base + k * 256);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k * 256 < n) __builtin_nontemporal_store(v[k], dst + base + k * 256);
}

// The pass-1 scatter's write pattern without its placement work: block c
// reads chunk c (words / chunks consecutive 16-byte words, 1024 per step,
// eight steps of loads in flight) and writes every step's 128 lines of 128
// bytes round-robin to `partitions` regions, each holding the chunks' pieces
// in chunk order -- the layout and line order of a uniform pass-1 partition.
// The host checks that every region and piece divides evenly.
constexpr uint32_t kCalibThreads = 1024, kCalibDepth = 8;
__global__ __launch_bounds__(kCalibThreads) void line_scatter_kernel(const u32x4* __restrict__ src,
                                                                      u32x4* __restrict__ dst, uint64_t words,
                                                                      uint32_t partitions) {
  const uint64_t chunk = words / gridDim.x;  // words per block
  const uint64_t piece = chunk / partitions;  // words per (partition, chunk)
  const uint64_t region = piece * gridDim.x;
  const uint32_t steps = static_cast<uint32_t>(chunk / kCalibThreads);
  const uint32_t t = threadIdx.x, line = t / 8, word = t % 8;
  const u32x4* in = src + blockIdx.x * chunk;
  for (uint32_t s0 = 0; s0 < steps; s0 += kCalibDepth) {
    u32x4 v[kCalibDepth];
#pragma unroll
    for (uint32_t k = 0; k < kCalibDepth; ++k)
      v[k] = __builtin_nontemporal_load(in + static_cast<uint64_t>(s0 + k) * kCalibThreads + t);
#pragma unroll
    for (uint32_t k = 0; k < kCalibDepth; ++k) {
      // the block's line number l goes to partition l mod P, its (l / P)-th line there
      const uint64_t l = static_cast<uint64_t>(s0 + k) * (kCalibThreads / 8) + line;
      const uint64_t p = l % partitions, m = l / partitions;
      __builtin_nontemporal_store(v[k], dst + p * region + blockIdx.x * piece + m * 8 + word);
    }
  }
}

// The register join's gather shape without its table work: one wave per
// item (p, q) of `items_per_part` = part_words / f2 words, read as runs of
// run_words words, run r of the item at tile r of partition p (a tile is f2
// runs), every run but a tile's last shifted by `shift` words -- runs that
// start inside a cache line, as real pass-2 runs do.  Each lane loads
// ceil(item / 64) words per item (lanes past the end repeat the last one),
// two items in flight; a per-wave xor of the loaded words keeps the loads.
__global__ __launch_bounds__(256) void run_gather_kernel(const u32x4* __restrict__ src, uint64_t words,
                                                         uint32_t part_words, uint32_t f2, uint32_t run_words,
                                                         uint32_t shift, uint32_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t item_words = part_words / f2;
  const uint64_t items = words / part_words * f2;
  const uint32_t loads = (item_words + 63) / 64;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * 4;
  uint32_t acc = 0;
  auto word = [&](uint64_t i, uint32_t k) {
    const uint64_t p = i / f2;
    const uint32_t q = static_cast<uint32_t>(i % f2);
    const uint32_t sh = q + 1 < f2 ? shift : 0u;
    const uint32_t e = min(k * 64 + lane, item_words - 1);
    return p * part_words + static_cast<uint64_t>(e / run_words) * (f2 * run_words) +
           static_cast<uint64_t>(q) * run_words + e % run_words + sh;
  };
  // two items per round, all their loads issued before either is used
  for (uint64_t i = blockIdx.x * 4ull + threadIdx.x / 64; i < items; i += 2 * waves) {
    const bool two = i + waves < items;
    u32x4 v[16], u[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
      if (k < loads) v[k] = src[word(i, k)];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
      if (k < loads && two) u[k] = src[word(i + waves, k)];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
      if (k < loads) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
      if (k < loads && two) acc ^= u[k].x ^ u[k].y ^ u[k].z ^ u[k].w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc ^= static_cast<uint32_t>(__shfl_xor(static_cast<int>(acc), off));
  if (lane == 0) atomicXor(out, acc);
}

uint32_t grid_for(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return g ? static_cast<uint32_t>(g) : 1u;
}

}  // namespace

hipError_t launch_gen_perm_keys(KeyType t, void* out, uint64_t n, uint64_t seed, int64_t base, hipStream_t stream) {
  if (t == KeyType::I32)
    hipLaunchKernelGGL(gen_perm_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int32_t*>(out), n, seed, base);
  else
    hipLaunchKernelGGL(gen_perm_kernel<int64_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int64_t*>(out), n, seed, base);
  return hipGetLastError();
}

hipError_t launch_gen_uniform(KeyType t, void* out, uint64_t n, uint64_t seed, uint64_t stream_id, int64_t lo,
                              uint64_t range, hipStream_t stream) {
  if (t == KeyType::I32)
    hipLaunchKernelGGL(gen_uniform_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int32_t*>(out), n, seed, stream_id, lo, range);
  else
    hipLaunchKernelGGL(gen_uniform_kernel<int64_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int64_t*>(out), n, seed, stream_id, lo, range);
  return hipGetLastError();
}

hipError_t launch_gen_zipf(KeyType t, void* out, uint64_t n, uint64_t num_ranks, double exponent, uint64_t seed,
                           uint64_t stream_id, uint64_t perm_seed, int64_t base, hipStream_t stream) {
  if (t == KeyType::I32)
    hipLaunchKernelGGL(gen_zipf_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream, static_cast<int32_t*>(out),
                       n, num_ranks, exponent, seed, stream_id, perm_seed, base);
  else
    hipLaunchKernelGGL(gen_zipf_kernel<int64_t>, dim3(grid_for(n)), dim3(256), 0, stream, static_cast<int64_t*>(out),
                       n, num_ranks, exponent, seed, stream_id, perm_seed, base);
  return hipGetLastError();
}

hipError_t launch_gen_iota(KeyType t, void* out, uint64_t n, int64_t base, hipStream_t stream) {
  if (t == KeyType::I32)
    hipLaunchKernelGGL(gen_iota_kernel<int32_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int32_t*>(out), n, base);
  else
    hipLaunchKernelGGL(gen_iota_kernel<int64_t>, dim3(grid_for(n)), dim3(256), 0, stream,
                       static_cast<int64_t*>(out), n, base);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* out, uint64_t n, uint64_t value, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint64_t>, dim3(grid_for(n)), dim3(256), 0, stream, out, n, value);
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, uint64_t words16, hipStream_t stream) {
  const uint64_t blocks = (words16 + 1023) / 1024;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(copy_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream,
                     static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), words16);
  return hipGetLastError();
}

hipError_t launch_line_scatter(void* dst, const void* src, uint64_t words16, uint32_t partitions, uint32_t chunks,
                               hipStream_t stream) {
  hipLaunchKernelGGL(line_scatter_kernel, dim3(chunks), dim3(kCalibThreads), 0, stream,
                     static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), words16, partitions);
  return hipGetLastError();
}

hipError_t launch_run_gather(const void* src, uint64_t words16, uint32_t part_words, uint32_t f2, uint32_t run_words,
                             uint32_t shift, uint32_t* out, hipStream_t stream, uint32_t grid) {
  hipLaunchKernelGGL(run_gather_kernel, dim3(grid), dim3(256), 0, stream, static_cast<const u32x4*>(src), words16,
                     part_words, f2, run_words, shift, out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* out, uint64_t n, uint32_t value, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(grid_for(n)), dim3(256), 0, stream, out, n, value);
  return hipGetLastError();
}

}  // namespace fij
