// Exclusive prefix sums over u32 counts (cell counts, hit flags), shared by
// the 2D and 3D engines.  4096 elements per block; the block totals are
// scanned by one block (so up to 16M elements), then a downsweep writes the
// exclusive prefix and the grand total at out[n], optionally clearing the
// counts for the next use.  Integer adds in a fixed tree: exact, and the
// result does not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {
// Exclusive prefix of cellCount into cellStart: 4096 cells per block,
// block totals scanned by one block, then a downsweep that also clears the
// counts for the next substep.  Integer adds in a fixed tree: exact.
constexpr uint32_t SCAN_PER_THREAD = 16, SCAN_BLOCK = 256 * SCAN_PER_THREAD;

// Work-list reservation from a 256-thread workgroup: thread i reserves n_i
// consecutive entries of a list and gets the index of its first one; the
// workgroup's total is reserved with one atomic on *counter (every wave's
// appends land on the same few counters, whose atomics are performed one at
// a time at the memory side).  The order of the threads' runs in the list is
// left open.  Every thread of the workgroup must call it.
__device__ __forceinline__ uint32_t wg_reserve(uint32_t n, uint32_t* counter) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t incl = n;
#pragma unroll
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const uint32_t v = (uint32_t)__shfl_up((int)incl, off);
    if (lane >= off) incl += v;
  }
  __shared__ uint32_t wtot[4], wbase[4];
  const uint32_t wv = threadIdx.x >> 6;
  if (lane == 63u) wtot[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0u) {
    const uint32_t s1 = wtot[0], s2 = s1 + wtot[1], s3 = s2 + wtot[2], tot = s3 + wtot[3];
    const uint32_t b = tot ? atomicAdd(counter, tot) : 0u;
    wbase[0] = b;
    wbase[1] = b + s1;
    wbase[2] = b + s2;
    wbase[3] = b + s3;
  }
  __syncthreads();
  const uint32_t base = wbase[wv] + incl - n;
  __syncthreads();   // wtot / wbase are reused by the next call
  return base;
}

// Work-list append: thread i appends the n_i items {key_i, 0..n_i-1}.  Every
// thread of the workgroup must call it.
__device__ __forceinline__ void wg_append_chunks(uint32_t n, uint32_t* counter, uint2* list, uint32_t key) {
  const uint32_t base = wg_reserve(n, counter);
  for (uint32_t k = 0; k < n; ++k) list[base + k] = make_uint2(key, k);
}

// Arrival ranks in the cell histogram, one device atomic per distinct cell
// of a 256-thread workgroup.  A thread whose `cell` is `none` takes no rank
// (returns 0).  The workgroup groups its cells in an LDS hash table (linear
// probing, 512 entries for at most 256 keys: a probe always ends), takes a
// local rank per thread from an LDS counter, reserves each cell's total
// with one atomic on counts[cell], and every thread gets base + local rank.
// The ranks within a cell are distinct, in no particular order.  Every
// thread of the workgroup must call it.
__device__ __forceinline__ uint32_t wg_cell_rank(uint32_t* counts, uint32_t cell, uint32_t none) {
  constexpr uint32_t HT = 512u;
  __shared__ uint32_t hkey[HT], hcnt[HT], hbase[HT];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < HT; i += 256u) {
    hkey[i] = none;
    hcnt[i] = 0u;
  }
  __syncthreads();
  uint32_t h = 0, local = 0;
  if (cell != none) {
    h = (cell * 2654435761u) >> 23;   // 9 bits
    while (true) {
      const uint32_t old = atomicCAS(&hkey[h], none, cell);
      if (old == none || old == cell) break;
      h = (h + 1u) & (HT - 1u);
    }
    local = atomicAdd(&hcnt[h], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < HT; i += 256u)
    if (hkey[i] != none) hbase[i] = atomicAdd(&counts[hkey[i]], hcnt[i]);
  __syncthreads();
  return cell != none ? hbase[h] + local : 0u;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  // 256 threads, Hillis-Steele in LDS
  uint32_t t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    uint32_t add = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[255];
  uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* blockSums) {
  __shared__ uint32_t sh[256];
  uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) s += (base + k < n) ? in[base + k] : 0u;
  uint32_t total;
  block_excl_scan(s, sh, total);
  if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_top(uint32_t* blockSums, uint32_t nb) {
  __shared__ uint32_t sh[256];
  uint32_t base = threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    v[k] = (base + k < nb) ? blockSums[base + k] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(s, sh, total);
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    if (base + k < nb) blockSums[base + k] = run;
    run += v[k];
  }
}

__global__ __launch_bounds__(256) void k_scan_down(uint32_t* inout_count, uint32_t n, const uint32_t* blockSums,
                                                   uint32_t* out, bool clear) {
  __shared__ uint32_t sh[256];
  uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t s = 0;
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    v[k] = (base + k < n) ? inout_count[base + k] : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(s, sh, total) + blockSums[blockIdx.x];
#pragma unroll
  for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
    if (base + k < n) {
      out[base + k] = run;
      if (clear) inout_count[base + k] = 0u;
    }
    run += v[k];
  }
  // the running sum after the last element of the last block is the total
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) out[n] = run;
}


}  // namespace
