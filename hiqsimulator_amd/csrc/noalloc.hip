// Sampling a caller's device state without allocating device memory.
//
// hiqsim_sample draws shots through the engine's own device buffers (block
// probabilities, descriptors, an output array it grows on demand).  Callers
// that cannot allocate (code under HIP graph capture, request loops that
// forbid hipMalloc) get the same two kernels here, launched on their stream
// over their state and output arrays, with every intermediate in one static
// device array per GPU:
//   [0, 2^15) doubles       block probabilities, turned into a running sum
//   [2^15, 2^15 + 32)       the block-sum item descriptor (one BatchItem)
//   [2^15 + 32, 2^15 + 64)  the sampling job (one SampleJob)
// The descriptors are written by a one-thread kernel from its arguments, so
// the call makes no host-to-device copy either.  Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU (two streams, two host
// threads) race on it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "batch.hpp"
#include "noalloc.hpp"

namespace hq {
namespace {

constexpr uint64_t kSumDoubles = uint64_t(1) << 15;  // sample_log_blocks() <= 15
constexpr uint64_t kDescDoubles = 32;                // 256 bytes per descriptor slot
constexpr uint64_t kScratchDoubles = kSumDoubles + 2 * kDescDoubles;
static_assert(sizeof(BatchItem) <= kDescDoubles * sizeof(double), "BatchItem slot");
static_assert(sizeof(SampleJob) <= kDescDoubles * sizeof(double), "SampleJob slot");

// Host code never passes `scratch` itself to a kernel: it is a host-side
// stand-in.  The device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
double* g_ptr[kMaxDevices] = {};
hipError_t g_lookup[kMaxDevices] = {};

__global__ void k_put_sample_desc(BatchItem* __restrict__ item_dst, SampleJob* __restrict__ job_dst,
                                  const BatchItem item, const SampleJob job)
{
     if (threadIdx.x == 0) {
          *item_dst = item;
          *job_dst = job;
     }
}

void check(hipError_t e, const char* what)
{
     if (e != hipSuccess)
          throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
double* scratch_ptr()
{
     int dev = -1;
     check(hipGetDevice(&dev), "hipGetDevice");
     if (dev < 0 || dev >= kMaxDevices)
          throw std::runtime_error("sample_noalloc(): device index out of range");
     std::call_once(g_once[dev], [dev] {
          void* p = nullptr;
          g_lookup[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
          if (g_lookup[dev] == hipSuccess) g_ptr[dev] = static_cast<double*>(p);
     });
     check(g_lookup[dev], "hipGetSymbolAddress(scratch)");
     if (!g_ptr[dev]) throw std::runtime_error("sample_noalloc(): scratch address is null");
     return g_ptr[dev];
}

}  // namespace

void sample_noalloc(const void* d_state, int n_qubits, Prec prec, uint32_t shots, uint64_t seed,
                    uint64_t* d_out, hipStream_t stream)
{
     if (n_qubits < 0 || n_qubits > kNoAllocMaxQubits)
          throw std::runtime_error("sample_noalloc(): n_qubits must lie in [0, " +
                                   std::to_string(kNoAllocMaxQubits) + "], got " +
                                   std::to_string(n_qubits));
     if (shots > kNoAllocMaxShots)
          throw std::runtime_error("sample_noalloc(): at most 2^20 shots, got " + std::to_string(shots));
     if (shots == 0) return;
     if (!d_state) throw std::runtime_error("null pointer: d_state");
     if (!d_out) throw std::runtime_error("null pointer: d_out");
     const uint32_t lb = sample_log_blocks(n_qubits);
     if ((uint64_t(1) << lb) > kSumDoubles)
          throw std::runtime_error("sample_noalloc(): " + std::to_string(uint64_t(1) << lb) +
                                   " probability blocks exceed the scratch area of " +
                                   std::to_string(kSumDoubles));
     double* const base = scratch_ptr();
     double* const sums = base;
     BatchItem* const d_item = reinterpret_cast<BatchItem*>(base + kSumDoubles);
     SampleJob* const d_job = reinterpret_cast<SampleJob*>(base + kSumDoubles + kDescDoubles);

     BatchItem it{};
     it.log_groups = static_cast<uint32_t>(n_qubits) - lb;  // amplitudes per block, log2
     SampleJob job{};
     job.seed = seed;
     job.log_blocks = lb;
     job.log_block_amps = static_cast<uint32_t>(n_qubits) - lb;
     job.shots = shots;
     hipLaunchKernelGGL(k_put_sample_desc, dim3(1), dim3(64), 0, stream, d_item, d_job, it, job);
     check(hipGetLastError(), "k_put_sample_desc");
     // Every block sum the sampler reads is written first, so the area needs
     // no clearing between calls.
     check(launch_block_sums_multi(d_state, prec, d_item, 1, uint64_t(1) << lb, sums, stream),
           "block_sums_multi");
     check(launch_sample(d_state, prec, d_job, 1, sums, d_out, stream), "sample");
}

}  // namespace hq
