// Shared pieces of the stateless state digest (digest.hip): the argument
// checks of digestState and the enqueue of its two kernels, used by every
// entry point that digests caller-held device words.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sndg {

constexpr uint32_t DG_THREADS = 256;
constexpr uint32_t DG_PER_THREAD = 4;
constexpr uint32_t DG_SPAN = DG_THREADS * DG_PER_THREAD;  // particles per workgroup

// Workgroups (and 8-word partials) a digest of `count` particles uses.
inline uint32_t partials(uint32_t count) { return (uint32_t)(((uint64_t)count + DG_SPAN - 1) / DG_SPAN); }

// Record the message returned by sn_digest_last_error(); returns `code`.
int fail(int code, const char* msg);

// digestState's argument checks (src/hash/statehash.js:97-108) as SN_E_ARG.
int check(const uint32_t* state, size_t state_words, uint32_t count, uint32_t stride, uint32_t hashed,
          const uint32_t* bond, uint32_t bond_cap, const uint32_t* acc_out);

// Enqueue the partial and fold kernels on `s` with `part` holding
// partials(count) x 8 words and `acc_dev` 8 words, copy the result to
// acc_out and wait for `s`.
hipError_t run(const uint32_t* state, uint32_t count, uint32_t stride, uint32_t hashed, uint32_t n_fluid,
               const uint32_t* bond, uint32_t bond_cap, uint32_t* part, uint32_t* acc_dev, uint32_t acc_out[8],
               hipStream_t s);

}  // namespace sndg
