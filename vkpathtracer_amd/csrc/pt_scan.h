// pt_scan.h — exclusive ranks of flagged items over a whole array, in two launches: chunk-local ranks
// (k_chunk_ranks-style kernels call chunk_rank) and the chunks' offsets (k_chunk_offsets). An item's
// rank is rank_local[i] + chunk_off[i / kScanChunk]. Used by the device BVH build (node numbering) and
// the device pool packing (inner-node numbering).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ptg {
namespace {

constexpr uint32_t kScanChunk = 1024;  // items per chunk: 256 threads, four items each

// Chunk-local exclusive ranks of the items i < m for which flag(i) is 1, and the chunk's count.
// Launched with 256 threads per chunk.
template <typename Flag>
__device__ __forceinline__ void chunk_rank(uint32_t m, Flag flag, uint32_t* rank_local, uint32_t* chunk_sum) {
  __shared__ uint32_t wsum[4];
  const uint32_t base = blockIdx.x * kScanChunk + threadIdx.x * 4u;
  uint32_t f[4], t = 0;
  for (uint32_t k = 0; k < 4u; ++k) {
    const uint32_t i = base + k;
    f[k] = i < m ? flag(i) : 0u;
    t += f[k];
  }
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = t;  // inclusive scan over the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= static_cast<uint32_t>(off)) x += y;
  }
  if (lane == 63u) wsum[w] = x;
  __syncthreads();
  uint32_t run = x - t, tot = 0;
  for (uint32_t j = 0; j < 4u; ++j) {
    if (j < w) run += wsum[j];
    tot += wsum[j];
  }
  for (uint32_t k = 0; k < 4u; ++k) {
    const uint32_t i = base + k;
    if (i < m) rank_local[i] = run;
    run += f[k];
  }
  if (threadIdx.x == 0u) chunk_sum[blockIdx.x] = tot;
}

// The chunks' counts become exclusive offsets in place (one workgroup of 1024 threads); the total of
// all flags goes to *total.
__global__ __launch_bounds__(1024) void k_chunk_offsets(uint32_t* chunk_sum, uint32_t chunks, uint32_t* total) {
  __shared__ uint32_t wsum[16];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < chunks; c0 += 1024u) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t t = i < chunks ? chunk_sum[i] : 0u;
    uint32_t x = t;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    if (lane == 63u) wsum[w] = x;
    __syncthreads();
    uint32_t pre = x - t, tot = 0;
    for (uint32_t j = 0; j < 16u; ++j) {
      if (j < w) pre += wsum[j];
      tot += wsum[j];
    }
    if (i < chunks) chunk_sum[i] = carry + pre;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0u) *total = carry;
}

__device__ __forceinline__ uint32_t rank_of(const uint32_t* rank_local, const uint32_t* chunk_off, uint32_t i) {
  return rank_local[i] + chunk_off[i / kScanChunk];
}

}  // namespace
}  // namespace ptg
