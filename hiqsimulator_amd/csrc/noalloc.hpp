// Entry points that allocate no device memory (noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace hq {

constexpr int kNoAllocMaxQubits = 35;               // 2^35 complex64 amplitudes = 256 GiB
constexpr uint32_t kNoAllocMaxShots = 1u << 20;

// hiqsim_sample's sampler over the caller's 2^n_qubits amplitudes at d_state
// (current device), outcomes into d_out[0 .. shots), enqueued on `stream`.
void sample_noalloc(const void* d_state, int n_qubits, Prec prec, uint32_t shots, uint64_t seed,
                    uint64_t* d_out, hipStream_t stream);

}  // namespace hq
