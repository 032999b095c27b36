// digestState (src/hash/statehash.js:80-159, the CPU twin of
// src/hash/statehash.wgsl:83-159) over words already resident in HBM.
//
// The engine's own hash runs fused into its last per-particle pass.  This
// file is the stateless form: any caller holding a particle buffer on the
// device (a restored snapshot, another engine's state, the 3D records with
// stride 16) gets the reference's 8-word digest of it without a copy to the
// host.  Four salted MurmurHash3 lanes per particle over its first
// `hashed_words` words, the optional order-free bond-row fold, then u32
// wrapping sums across particles: each workgroup of 256 lanes covers 1024
// particles and writes one 8-word partial, and one workgroup folds the
// partials.  Both sums are u32 additions, so the grouping changes nothing.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/sunaengine.h"
#include "digest.hpp"
#include "fixed.hpp"

using namespace sn;
using sndg::DG_PER_THREAD;
using sndg::DG_SPAN;
using sndg::DG_THREADS;

namespace {

constexpr uint32_t BOND_SALT = 0x5bd1e995u, BOND_SALT2 = 0x2f5b3c17u;

thread_local std::string g_err;

__device__ __forceinline__ uint32_t dg_wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
  return v;
}

// One workgroup: the digest terms of particles [blockIdx.x * DG_SPAN, +DG_SPAN)
// summed into part[blockIdx.x][0..7].
__global__ __launch_bounds__(DG_THREADS) void k_digest_part(const uint32_t* __restrict__ st, uint32_t count,
                                                            uint32_t stride, uint32_t hashed, uint32_t n_fluid,
                                                            const uint32_t* __restrict__ bond, uint32_t bond_cap,
                                                            uint32_t* __restrict__ part) {
  const uint32_t SEED[4] = {0x9e3779b1u, 0x85ebca77u, 0xc2b2ae3du, 0x27d4eb2fu};
  const uint32_t SALT[4] = {0x00000000u, 0x51633e2du, 0x9c8f2a17u, 0xd0e5b943u};
  uint32_t sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t r = 0; r < DG_PER_THREAD; ++r) {
    const uint32_t i = blockIdx.x * DG_SPAN + r * DG_THREADS + threadIdx.x;
    if (i >= count) break;
    const uint32_t* w = st + (size_t)i * stride;
    uint32_t h[4] = {SEED[0], SEED[1], SEED[2], SEED[3]};
    for (uint32_t k = 0; k < hashed; ++k) {
      const uint32_t x = w[k];
#pragma unroll
      for (int l = 0; l < 4; ++l) h[l] = mm_step(h[l], x ^ SALT[l]);
    }
    uint32_t s[4] = {h[0], h[1], h[2], h[3]}, q[4] = {h[0], h[1], h[2], h[3]};
    if (bond != nullptr && i < n_fluid) {
      uint32_t bn = 0, com = 0, scom = 0;
      const uint32_t* row = bond + (size_t)i * bond_cap;
      for (uint32_t k = 0; k < bond_cap; ++k) {
        const uint32_t b = row[k];
        if (b == 0u) break;  // rows are dense: the tail is zero
        com += fmix32(((b >> 17) & 0x1fffu) ^ BOND_SALT);
        scom += fmix32(b ^ BOND_SALT2);
        ++bn;
      }
      if (bn != 0u) {
#pragma unroll
        for (int l = 0; l < 4; ++l) {
          s[l] = mm_step(s[l], bn) ^ com;
          q[l] = mm_step(q[l], bn) ^ scom;
        }
      }
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      sum[l] += fmix32(s[l] ^ hashed);
      sum[4 + l] += fmix32(mm_step(q[l], i) ^ (hashed + 1u));
    }
  }
  __shared__ uint32_t wpart[DG_THREADS / 64][8];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const uint32_t v = dg_wave_sum(sum[l]);
    if (lane == 0u) wpart[wv][l] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8u) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t k = 0; k < DG_THREADS / 64; ++k) t += wpart[k][threadIdx.x];
    part[(size_t)blockIdx.x * 8u + threadIdx.x] = t;
  }
}

// One workgroup: acc[l] = sum over b of part[b][l].
__global__ __launch_bounds__(DG_THREADS) void k_digest_fold(const uint32_t* __restrict__ part, uint32_t nb,
                                                            uint32_t* __restrict__ acc) {
  uint32_t sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < nb; b += DG_THREADS) {
#pragma unroll
    for (int l = 0; l < 8; ++l) sum[l] += part[(size_t)b * 8u + l];
  }
  __shared__ uint32_t wpart[DG_THREADS / 64][8];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const uint32_t v = dg_wave_sum(sum[l]);
    if (lane == 0u) wpart[wv][l] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8u) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t k = 0; k < DG_THREADS / 64; ++k) t += wpart[k][threadIdx.x];
    acc[threadIdx.x] = t;
  }
}

}  // namespace

namespace sndg {

int fail(int code, const char* msg) {
  g_err = msg;
  return code;
}

// The argument checks of digestState (statehash.js:97-108), as SN_E_ARG.
int check(const uint32_t* state, size_t state_words, uint32_t count, uint32_t stride, uint32_t hashed,
          const uint32_t* bond, uint32_t bond_cap, const uint32_t* acc_out) {
  if (!acc_out) return fail(SN_E_ARG, "[aether] digestState: acc_out is null");
  if (count && !state) return fail(SN_E_ARG, "[aether] digestState: state is null");
  if (hashed > stride) return fail(SN_E_ARG, "hashedWords > strideWords");
  if ((uint64_t)count * stride > (uint64_t)state_words) return fail(SN_E_ARG, "state buffer too small");
  if (bond && bond_cap == 0u) return fail(SN_E_ARG, "digestState: p.bond given without p.bondCap");
  return SN_OK;
}

// Enqueue both kernels on `s`, then copy the 8 words to the host and wait.
hipError_t run(const uint32_t* state, uint32_t count, uint32_t stride, uint32_t hashed, uint32_t n_fluid,
               const uint32_t* bond, uint32_t bond_cap, uint32_t* part, uint32_t* acc_dev, uint32_t acc_out[8],
               hipStream_t s) {
  const uint32_t nb = partials(count);
  if (nb) hipLaunchKernelGGL(k_digest_part, dim3(nb), dim3(DG_THREADS), 0, s, state, count, stride, hashed,
                             n_fluid, bond, bond_cap, part);
  hipLaunchKernelGGL(k_digest_fold, dim3(1), dim3(DG_THREADS), 0, s, part, nb, acc_dev);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(acc_out, acc_dev, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(s);
}

}  // namespace sndg

extern "C" {

const char* sn_digest_last_error(void) { return g_err.c_str(); }

int sn_digest_state(const uint32_t* state, size_t state_words, uint32_t particle_count, uint32_t stride_words,
                    uint32_t hashed_words, uint32_t n_fluid, const uint32_t* bond, uint32_t bond_cap,
                    uint32_t acc_out[8], void* stream) {
  int rc = sndg::check(state, state_words, particle_count, stride_words, hashed_words, bond, bond_cap, acc_out);
  if (rc) return rc;
  const uint32_t nb = sndg::partials(particle_count);
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* buf = nullptr;  // nb partials of 8 words, then the 8-word result
  if (hipMalloc((void**)&buf, ((size_t)nb + 1) * 8 * sizeof(uint32_t)) != hipSuccess)
    return sndg::fail(SN_E_DEVICE, "[aether] digestState: hipMalloc failed");
  hipError_t err = sndg::run(state, particle_count, stride_words, hashed_words, n_fluid, bond, bond_cap, buf,
                             buf + (size_t)nb * 8, acc_out, s);
  hipFree(buf);
  if (err != hipSuccess) return sndg::fail(SN_E_DEVICE, hipGetErrorString(err));
  return SN_OK;
}

}  // extern "C"
