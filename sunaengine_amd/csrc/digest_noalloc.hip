// sn_digest_state_noalloc: the state digest of digest.hip for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  It runs the same two kernels,
// with its partial sums and result held in one file-scope device array
// instead of a hipMalloc'd buffer, and calls no allocating routine.
//
// The array is sized for SN_DIGEST_NOALLOC_MAX_PARTICLES particles (one
// 8-word partial per 1024 particles, plus the 8-word result: 256 KiB + 32 B
// per GPU, held until the process ends).  Each GPU holds its own copy at its
// own address; the address is looked up once per device, with that device
// current, on the first call there, and only that address is ever handed to
// kernels or copies.  Every call on one GPU shares the array: the partials
// and the result are written before they are read, so nothing needs
// clearing, but two calls that overlap on one GPU (from different host
// threads or streams) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/sunaengine.h"
#include "digest.hpp"

namespace {

constexpr uint32_t NOALLOC_PARTIALS = SN_DIGEST_NOALLOC_MAX_PARTICLES / sndg::DG_SPAN;
static_assert(NOALLOC_PARTIALS * sndg::DG_SPAN == SN_DIGEST_NOALLOC_MAX_PARTICLES,
              "the particle limit is a whole number of workgroup spans");
constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
uint32_t* g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device

}  // namespace

// NOALLOC_PARTIALS partials of 8 words, then the 8-word result.
static __device__ __align__(256) uint32_t scratch[(NOALLOC_PARTIALS + 1) * 8];

namespace {

// The device address of `scratch` on the current device, looked up once.
int scratch_on_device(uint32_t** out) {
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return sndg::fail(SN_E_DEVICE, hipGetErrorString(err));
  if (dev < 0 || dev >= MAX_DEVICES) return sndg::fail(SN_E_DEVICE, "[aether] digestState: device ordinal too large");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || p == nullptr)
      return sndg::fail(SN_E_DEVICE, "[aether] digestState: hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SN_OK;
}

}  // namespace

extern "C" {

int sn_digest_state_noalloc(const uint32_t* state, size_t state_words, uint32_t particle_count,
                            uint32_t stride_words, uint32_t hashed_words, uint32_t n_fluid, const uint32_t* bond,
                            uint32_t bond_cap, uint32_t acc_out[8], void* stream) {
  int rc = sndg::check(state, state_words, particle_count, stride_words, hashed_words, bond, bond_cap, acc_out);
  if (rc) return rc;
  if (particle_count > SN_DIGEST_NOALLOC_MAX_PARTICLES)
    return sndg::fail(SN_E_CAPACITY,
                      "[aether] digestState (no-alloc): particleCount exceeds SN_DIGEST_NOALLOC_MAX_PARTICLES "
                      "(8388608); use sn_digest_state");
  uint32_t* buf = nullptr;
  rc = scratch_on_device(&buf);
  if (rc) return rc;
  hipError_t err = sndg::run(state, particle_count, stride_words, hashed_words, n_fluid, bond, bond_cap, buf,
                             buf + (size_t)NOALLOC_PARTIALS * 8, acc_out, static_cast<hipStream_t>(stream));
  if (err != hipSuccess) return sndg::fail(SN_E_DEVICE, hipGetErrorString(err));
  return SN_OK;
}

}  // extern "C"
