// Launch interface of the gfx950 state-vector kernels (kernels.hip).
//
// State layout in HBM: one contiguous array of 2^n interleaved complex
// amplitudes (complex128 = 16 B, complex64 = 8 B); bit p of the array index is
// the qubit at bit position p (the engine's `locals` list maps positions to
// qubit ids, SimulatorMPI.cpp:546-556).  All launches are asynchronous on the
// caller's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hq {

enum class Prec : int { C128 = 0, C64 = 1 };

inline size_t amp_bytes(Prec p) { return p == Prec::C128 ? 16 : 8; }

// Ascending bit positions at which zero bits are inserted into a thread's
// linear index (targets and/or controls).  At most 63 positions.
struct InsertSet {
     int n;
     uint8_t p[63];
};

// Description of one fused dense pass.
struct DenseDesc {
     int k;                 // fused width, 1..5
     uint32_t pos[5];       // matrix bit l <-> index bit pos[l]
     uint64_t ctrlmask;     // all these bits must be 1 (SimulatorMPI.cpp:463)
     InsertSet ins;         // sorted(targets ∪ controls)
};

// psi[I + off_i] = sum_j m[i][j] psi[I + off_j] for every base I whose target
// bits are 0 and control bits are 1.  `mat` is a device array of
// (2^k)^2 complex values in the state's precision, row-major.
hipError_t launch_dense(void* psi, Prec prec, int n_qubits, const void* mat,
                        const DenseDesc& d, hipStream_t s);

// The dense pass takes its direct form (instead of the LDS tile form) when
// every target bit is at or above this position: `tile_high_target` for
// passes of k != 4 (default 64: never), `tile_high_target_wide` for k = 4
// (default 13).  get_tile_high_target_for(k) is the one that applies to a
// k-qubit pass, in single-state, batched and request-loop passes alike.
// Every knob below is process-wide and atomic (read once per launch).
void set_tile_high_target(int pos);
int get_tile_high_target();
int get_tile_high_target_for(int k);
// Low block-index bits reversed in the streaming passes (0 = identity
// order, at most 16): `block_swizzle` for passes of k <= 3 and the other
// streaming kernels (default 8), `block_swizzle_wide` for dense passes of
// k >= 4 (default 6).
void set_block_swizzle(int bits);
int get_block_swizzle();
void set_block_swizzle_wide(int bits);
int get_block_swizzle_wide();
int get_block_swizzle_for(int k);
// Op engine of tile programs: 0 every op on the vector units; 1 ops of 3+
// qubits on the matrix cores (default).
void set_tile_mfma(int mode);
int get_tile_mfma();
// Complex128 5-qubit passes without controls on 12+ qubits: 1 (default) as a
// one-op tile pass on the matrix cores; 0 the dense pass forms.
void set_dense5_tile(int on);
int get_dense5_tile();
void set_direct_block(int bs);
int get_direct_block();
void set_tile_high_target_wide(int pos);
int get_tile_high_target_wide();
// Tiled qubit permutation (n >= 12): low block-index bits reversed (0..16,
// clamped to the grid) and threads per workgroup (256 or 512).
void set_permute_swizzle(int bits);
int get_permute_swizzle();
void set_permute_block(int bs);
int get_permute_block();
// Threads per workgroup of the tile-program kernels with matrix-core ops:
// 256, 512 or 0 (the default: 256 for complex64 and for complex128 programs
// with 5-qubit ops, 512 for the other complex128 programs).
void set_tile_threads(int nt);
int get_tile_threads();
int tile_threads_for(Prec p, int kmax);

// psi[i] *= diag[d(i)] for every i whose control bits are 1; d(i) gathers the
// target bits.  diag holds 2^k complex128 values (k <= 5).
hipError_t launch_diag(void* psi, Prec prec, int n_qubits, const double* diag_re_im,
                       const DenseDesc& d, hipStream_t s);

// psi[i] *= f for every i.
hipError_t launch_scale(void* psi, Prec prec, uint64_t n, double fre, double fim,
                        hipStream_t s);

// psi[i] = v for every i.
hipError_t launch_fill(void* psi, Prec prec, uint64_t n, double vre, double vim,
                       hipStream_t s);

// dst[i] = src[i], n amplitudes (plain streaming copy).
hipError_t launch_copy(void* dst, const void* src, Prec prec, uint64_t n, hipStream_t s);

// (i & msk) == val ? psi[i] * f : 0 (SimulatorMPI.cpp:872-895).
hipError_t launch_normalize(void* psi, Prec prec, uint64_t n, uint64_t msk, uint64_t val,
                            double f, hipStream_t s);

// dst[j] = src[insert_bit(j, pos) | keep << pos], j < n/2 (bit removal,
// SimulatorMPI.cpp:305-321).
hipError_t launch_remove_bit(void* dst, const void* src, Prec prec, uint64_t n_src, int pos,
                             int keep, hipStream_t s);

// dst[j] = src[i(j)] with bit p of j = bit src_pos[p] of i, for p < n_qubits.
hipError_t launch_permute(void* dst, const void* src, Prec prec, int n_qubits,
                          const uint8_t* src_pos, hipStream_t s);

// ---- reductions (FP64 accumulation; results land in device `out`) --------
// Number of FP64 slots of scratch the reductions need.
constexpr int kReduceBlocks = 1024;
constexpr size_t kReduceScratchDoubles = 5 * kReduceBlocks + 16;

// out[0] = sum |psi[i]|^2 over i with (i & msk) == val.
hipError_t launch_masked_norm(const void* psi, Prec prec, int n_qubits, uint64_t msk,
                              uint64_t val, double* scratch, double* out, hipStream_t s);
// out[0] = sum over i with bit pos == 0, out[1] = with bit pos == 1.
hipError_t launch_two_bucket(const void* psi, Prec prec, int n_qubits, int pos,
                             double* scratch, double* out, hipStream_t s);
// out[0] = sum p log2 p over p = |psi[i]|^2 > 0 (SimulatorMPI.cpp:681-694).
hipError_t launch_entropy(const void* psi, Prec prec, uint64_t n, double* scratch,
                          double* out, hipStream_t s);
// out[0..1] = sum_i conj(psi[i ^ x]) * phase(i) * psi[i],
// phase(i) = i^{popcount(x & z)} * (-1)^{popcount(i & z)}.
hipError_t launch_pauli_expectation(const void* psi, Prec prec, uint64_t n, uint64_t xmask,
                                    uint64_t zmask, double* scratch, double* out,
                                    hipStream_t s);
// out[b] = sum_{j < K} |psi[b*K + j]|^2 for b < nb, K = n / nb.
hipError_t launch_block_sums(const void* psi, Prec prec, uint64_t n, uint64_t nb, double* out,
                             hipStream_t s);

// psi[i] = complex Gaussian N(0,1) + i N(0,1) from Philox4x32-10 keyed by
// `seed` with counter i (not normalised).
hipError_t launch_random_state(void* psi, Prec prec, uint64_t n, uint64_t seed, hipStream_t s);
// out = {Re, Im of sum conj(ref) psi, sum |ref|^2, max |psi - ref|, sum |psi|^2}.
hipError_t launch_compare(const void* psi, const void* ref, Prec prec, uint64_t n, double* scratch,
                          double* out, hipStream_t s);

// dst[i ^ x] += c * phase(i) * src[i] for every i (one Pauli term of a sum).
hipError_t launch_pauli_accumulate(void* dst, const void* src, Prec prec, uint64_t n,
                                   uint64_t xmask, uint64_t zmask, double cre, double cim,
                                   hipStream_t s);

}  // namespace hq
