/*
 * hiqsim_amd.h — C ABI of the MI355X (gfx950) state-vector engine.
 *
 * Drop-in boundary for HiQsimulator's full-amplitude simulator.  The
 * reference exposes this path to Python through the pybind11 module
 * `_cppsim_mpi` (_cppsim_mpi.cpp:61-83), class SimulatorMPI
 * (src/simulator-mpi/SimulatorMPI.hpp:43-308).  Each entry point below names
 * the binding and the C++ member it replaces.  Reference paths are relative to
 * the HiQsimulator source root.
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero status on error; the
 *    message (the reference's std::runtime_error text where one exists) is
 *    available from hiqsim_last_error() on the calling thread.
 *  - Matrices are dense, row-major, interleaved complex128:
 *    m[2*(i*dim+j)] = Re m[i][j], m[2*(i*dim+j)+1] = Im m[i][j]; matrix bit l
 *    acts on ids[l] (ids[0] is the least significant bit), as in
 *    SimulatorMPI::ApplyGate.
 *  - Bit strings / outcomes are one byte per qubit (0 or 1).
 *  - State vectors crossing the boundary are interleaved complex128 in host
 *    memory; index bit p <-> the qubit hiqsim_get_qubits_ids() lists at p.
 *  - A handle owns one device state and one HIP stream; it is not
 *    thread-safe (the reference object is not either).
 */
#ifndef HIQSIM_AMD_H
#define HIQSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hiqsim_engine hiqsim_t;

enum hiqsim_precision { HIQSIM_C128 = 0, HIQSIM_C64 = 1 };

/* ABI version of this header. */
int hiqsim_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* hiqsim_last_error(void);
/* Number of visible HIP devices. */
int hiqsim_device_count(int* out);

/* SimulatorMPI(seed, max_local, max_cluster_size)
 * (_cppsim_mpi.cpp:64, SimulatorMPI.cpp:66-104).  precision: HIQSIM_C128 or
 * HIQSIM_C64 (amplitudes only; fusion, norms and probabilities stay FP64). */
int hiqsim_create(uint64_t seed, int64_t max_local, int64_t max_cluster_size, int precision,
                  int device, hiqsim_t** out);
int hiqsim_destroy(hiqsim_t* h);

/* allocate_qubit (_cppsim_mpi.cpp:72, SimulatorMPI.cpp:185-216). */
int hiqsim_allocate_qubit(hiqsim_t* h, int64_t id);
/* allocate_qureg (_cppsim_mpi.cpp:71, SimulatorMPI.cpp:218-251). */
int hiqsim_allocate_qureg(hiqsim_t* h, const int64_t* ids, size_t n, double init_re,
                          double init_im);
/* deallocate_qubit (_cppsim_mpi.cpp:73, SimulatorMPI.cpp:369-426). */
int hiqsim_deallocate_qubit(hiqsim_t* h, int64_t id);

/* apply_controlled_gate (_cppsim_mpi.cpp:75, SimulatorMPI.cpp:710-815):
 * deferred until hiqsim_run; dim must equal 2^nids. */
int hiqsim_apply_controlled_gate(hiqsim_t* h, const double* m, size_t dim, const int64_t* ids,
                                 size_t nids, const int64_t* ctrls, size_t nctrls);
/* run (_cppsim_mpi.cpp:79, SimulatorMPI.cpp:441-541): fuse the pending gates
 * into one cluster and apply it.  Raises the reference's "cannot apply N
 * qubits gate" for a cluster wider than 5 qubits.  How the cluster is
 * applied follows the handle's pass mode (hiqsim_set_pass_mode). */
int hiqsim_run(hiqsim_t* h);
/* Pass mode of hiqsim_run (no reference counterpart):
 *   HIQSIM_PASS_EACH (0): every cluster is one HBM pass at once, as the
 *     reference's Run() (SimulatorMPI.cpp:441-541);
 *   HIQSIM_PASS_DEFERRED (1, the default): on a state of 12 or more qubits
 *     the gates of a cluster of at most 5 qubits with its controls are
 *     queued (unfused); queued gates run as tile passes, several clusters
 *     per HBM pass, planned with the LDS tile as the local set and fused per
 *     tile op (as hiqsim_apply_tiled), once 512 of them wait (down to 256)
 *     and before any call that reads the state, changes the qubit set or the
 *     qubit order, or synchronizes.  Results equal the reference's within
 *     floating-point rounding.
 * get: *mode, and *queued = gates waiting (queued may be NULL). */
enum hiqsim_pass_mode { HIQSIM_PASS_EACH = 0, HIQSIM_PASS_DEFERRED = 1 };
int hiqsim_set_pass_mode(hiqsim_t* h, int mode);
int hiqsim_get_pass_mode(hiqsim_t* h, int* mode, size_t* queued);

/* measure_qubits (_cppsim_mpi.cpp:74, SimulatorMPI.cpp:897-1008). */
int hiqsim_measure_qubits(hiqsim_t* h, const int64_t* ids, size_t n, uint8_t* out);
/* get_probability (_cppsim_mpi.cpp:78, SimulatorMPI.cpp:561-600). */
int hiqsim_get_probability(hiqsim_t* h, const uint8_t* bits, size_t nbits, const int64_t* ids,
                           size_t nids, double* out);
/* get_amplitude (_cppsim_mpi.cpp:77, SimulatorMPI.cpp:602-650); out = {re, im}. */
int hiqsim_get_amplitude(hiqsim_t* h, const uint8_t* bits, size_t nbits, const int64_t* ids,
                         size_t nids, double* out);
/* collapse_wavefunction (_cppsim_mpi.cpp:82, SimulatorMPI.cpp:1010-1058). */
int hiqsim_collapse_wavefunction(hiqsim_t* h, const int64_t* ids, size_t nids,
                                 const uint8_t* values, size_t nvalues);
/* entropy (_cppsim_mpi.cpp:80, SimulatorMPI.cpp:681-694). */
int hiqsim_entropy(hiqsim_t* h, double* out);

/* Number of allocated qubits (SimulatorMPI.hpp:275-297). */
int hiqsim_num_qubits(hiqsim_t* h, size_t* out);
/* get_qubits_ids / get_local_qubits_ids (_cppsim_mpi.cpp:65-66,
 * SimulatorMPI.cpp:652-659): id of the qubit at each bit position.  One
 * device has no global qubits (get_global_qubits_ids is empty). */
int hiqsim_get_qubits_ids(hiqsim_t* h, int64_t* out, size_t cap, size_t* n);
/* set_qubits_perm (_cppsim_mpi.cpp:69, SimulatorMPI.cpp:661-667): relabel only. */
int hiqsim_set_qubits_perm(hiqsim_t* h, const int64_t* ids, size_t n);
/* swap_qubits (_cppsim_mpi.cpp:70, SimulatorMPI.cpp:1060-1138): (global, local)
 * pairs; on one device there are no global qubits, so any pair is an error. */
int hiqsim_swap_qubits(hiqsim_t* h, const int64_t* pairs, size_t n);
/* In-HBM qubit permutation replacing the data movement of the reference's
 * qswap (SwapperMT::doSwap, SwapperMT.cpp:100-126; swapping.hpp:33-67):
 * afterwards ids[p] is the qubit at bit p.  Bit-exact. */
int hiqsim_permute_qubits(hiqsim_t* h, const int64_t* ids, size_t n);

/* Number of amplitudes (2^num_qubits). */
int hiqsim_state_size(hiqsim_t* h, uint64_t* out);
/* cheat_local (_cppsim_mpi.cpp:81, SimulatorMPI.cpp:543-559): copy the state
 * out as interleaved complex128 (n_amps must equal hiqsim_state_size). */
int hiqsim_get_state(hiqsim_t* h, double* out, uint64_t n_amps);
/* Zero-copy access to the state (the reference only copies it out,
 * cheat_local, SimulatorMPI.cpp:543-559): runs pending gates, drains the
 * handle's stream and returns the device buffer of n_amps interleaved
 * amplitudes (complex128 or complex64 per *precision, HIQSIM_C128 /
 * HIQSIM_C64) on HIP device *device.  The pointer stays valid until the
 * next call that moves the state to another buffer (allocation or
 * deallocation of a qubit, hiqsim_permute_qubits, hiqsim_apply_pauli_sum,
 * hiqsim_restore_state) or the handle's destruction; work the caller queues
 * on other streams must be complete before the next call on the handle.
 * precision and device may be NULL. */
int hiqsim_device_state(hiqsim_t* h, void** ptr, uint64_t* n_amps, int* precision, int* device);
/* hiqsim_device_state with a pin (no reference counterpart).  While at least
 * one pin is held, permutations and Pauli sums copy their result back into
 * the pinned buffer (one extra streaming pass each) and no state buffer is
 * freed: the pointer stays allocated until hiqsim_unpin_device_state, and
 * its contents follow the state until the qubit count changes.  Unpinning
 * only decrements the count and may be called from any thread; buffers
 * retired while pinned are freed by a later call on the handle. */
int hiqsim_pin_device_state(hiqsim_t* h, void** ptr, uint64_t* n_amps, int* precision, int* device);
int hiqsim_unpin_device_state(hiqsim_t* h);
/* Save/restore of the state to a file (no reference counterpart; the
 * reference can only copy the state out, cheat_local), streamed through a
 * 64 MiB pinned buffer.  Format: "HQSV", u32 version 1, u32 precision
 * (HIQSIM_C128 / HIQSIM_C64), u32 n, n x int64 qubit ids in bit order, then
 * 2^n interleaved amplitudes in that precision, all little-endian.
 * load_state needs the handle to hold the same set of qubit ids; it adopts
 * the file's bit order (a relabel) and converts the precision. */
int hiqsim_save_state(hiqsim_t* h, const char* path);
int hiqsim_load_state(hiqsim_t* h, const char* path);
/* set_wavefunction (_simulator_mpi.py:281-306; declared but unbound in the
 * reference): replace the amplitudes, qubit order unchanged. */
int hiqsim_set_state(hiqsim_t* h, const double* in, uint64_t n_amps);

/* Device-side checkpoint (no counterpart in the reference, whose only state
 * export is the cheat_local copy-out, SimulatorMPI.cpp:543-559): snapshot
 * copies the state and qubit order into a second HBM buffer with one
 * streaming copy; restore copies them back. */
int hiqsim_snapshot_state(hiqsim_t* h);
int hiqsim_restore_state(hiqsim_t* h);
/* Precision report against the snapshot (no reference counterpart):
 * out = {Re, Im of <snapshot|state>, <snapshot|snapshot>, max |state - snapshot|,
 *        <state|state>} (out holds 5 doubles, accumulated in FP64). */
int hiqsim_compare_snapshot(hiqsim_t* h, double* out);
/* Normalised complex Gaussian random state (Philox4x32-10 keyed by seed, one
 * counter per amplitude); no reference counterpart. */
int hiqsim_init_random_state(hiqsim_t* h, uint64_t seed);
/* `shots` computational-basis samples of the state without collapsing it
 * (pending gates are run first): out[s] = sampled index, bit p = the qubit
 * hiqsim_get_qubits_ids() lists at p.  Shot s draws u from Philox4x32-10
 * (counter s, key seed) and takes the first index whose running
 * probability exceeds u * norm.  No reference counterpart: the reference
 * samples by repeated MeasureQubits + collapse (SimulatorMPI.cpp:897-1008). */
int hiqsim_sample(hiqsim_t* h, uint32_t shots, uint64_t seed, uint64_t* out);
/* hiqsim_sample's sampler over a caller's device state, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, request loops that forbid
 * hipMalloc).  d_state: 2^n_qubits amplitudes of the given precision in
 * device memory of the current HIP device; d_out: `shots` uint64 in device
 * memory, d_out[s] = sampled amplitude index, the same outcomes hiqsim_sample
 * returns for that state and seed.  The work is enqueued on `stream` (a
 * hipStream_t of the current device; NULL = the null stream) and nothing is
 * synchronised.  Limits: n_qubits <= 35, shots <= 2^20 (larger values are
 * rejected with an error).  Intermediates (2^15 block probabilities and two
 * descriptors) live in one static device array of 256.5 KiB per GPU, held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU (several streams or host threads) race.  The call may be
 * captured into a HIP graph once it has run eagerly on that device (its first
 * call on a device looks up the array's address). */
int hiqsim_sample_noalloc(const void* d_state, int n_qubits, int precision, uint32_t shots,
                          uint64_t seed, uint64_t* d_out, void* stream);

/* get_expectation_value (_simulator_mpi.py:148-183; unbound in the
 * reference): sum_t coeff[t] <psi|P_t|psi>.  Bit i of xmask[t]/zmask[t] refers
 * to ids[i]; X = x only, Z = z only, Y = both.  coeff and out are {re, im}. */
int hiqsim_expectation_pauli(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                             const double* coeff, size_t nterms, const int64_t* ids, size_t nids,
                             double* out);
/* apply_qubit_operator (_simulator_mpi.py:185-223; unbound in the reference):
 * psi <- sum_t coeff[t] P_t psi (not renormalised). */
int hiqsim_apply_pauli_sum(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                           const double* coeff, size_t nterms, const int64_t* ids, size_t nids);
/* emulate_time_evolution (called by _simulator_mpi.py:469-475 for ProjectQ's
 * TimeEvolution gate, never defined in the reference's C++): psi <- exp(-i
 * time H) psi on the subspace where every ctrls qubit is 1, H the Pauli sum
 * as above (Hermitian for a unitary result).  Identity terms become a phase;
 * the rest is a Taylor series in ceil(|time| sum|coeff|) steps, each summed
 * until a term's norm is below 1e-12. */
int hiqsim_emulate_time_evolution(hiqsim_t* h, const uint64_t* xmask, const uint64_t* zmask,
                                  const double* coeff, size_t nterms, double time, const int64_t* ids,
                                  size_t nids, const int64_t* ctrls, size_t nctrls);

/* One immediate fused pass at explicit bit positions, bypassing fusion: the
 * reference's kernelK / kernel_core(_diag) call in Run()
 * (SimulatorMPI.cpp:470-514).  k = number of positions (1..5), m is 2^k x 2^k;
 * diag != 0 selects the diagonal kernel when m is diagonal. */
int hiqsim_apply_matrix(hiqsim_t* h, const double* m, int k, const uint32_t* pos,
                        uint64_t ctrlmask, int diag);

/* Wait for all queued device work of this handle. */
int hiqsim_synchronize(hiqsim_t* h);
/* HIP-event timing of every kernel launch on the handle's stream (resets the
 * counters).  Classes: 0-4 dense k=1..5, 5 diagonal, 6 scale, 7 permute,
 * 8 reductions, 9 normalize, 10 bit removal, 11 fill/copy, 12 Pauli. */
int hiqsim_set_timing(hiqsim_t* h, int on);
int hiqsim_kernel_time(hiqsim_t* h, int cls, uint64_t* count, double* total_ms);

/* ---- gate schedulers (host only; no device needed) ----------------------
 * Gate lists are ragged: gate g's target ids are gate_ids[gate_off[g] ..
 * gate_off[g+1]), its control ids ctrl_ids[ctrl_off[g] .. ctrl_off[g+1]);
 * gate_off and ctrl_off hold ngates+1 entries; diag[g] != 0 marks a diagonal
 * gate.  At most 63 distinct qubit ids. */

/* ClusterScheduler(gate, gate_ctrl, gate_diag, locals, globals,
 * cluster_size).ScheduleCluster() (_sched_cpp.cpp:36-39,
 * cluster_scheduler.cpp:42-99): indices of the gates forming the next fused
 * cluster, in gate order.  out must hold ngates entries; *nout is set. */
int hiqsim_schedule_cluster(const int64_t* gate_ids, const size_t* gate_off,
                            const int64_t* ctrl_ids, const size_t* ctrl_off, const uint8_t* diag,
                            size_t ngates, const int64_t* locals, size_t nlocals,
                            const int64_t* globals, size_t nglobals, int cluster_size,
                            int32_t* out, size_t* nout);
/* SwapScheduler(gate, gate_ctrl, gate_diag, num_splits, num_locals,
 * fuse).ScheduleSwap() (_sched_cpp.cpp:31-35, swap_scheduler.cpp:25-170):
 * the qubit ids to hold locally in the next stage, ascending.  out must hold
 * 64 entries; *nout is set. */
int hiqsim_schedule_swap(const int64_t* gate_ids, const size_t* gate_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, const uint8_t* diag, size_t ngates,
                         int64_t num_splits, int num_locals, int fuse, int64_t* out,
                         size_t* nout);

/* Whole-circuit cluster plan: cluster_of[g] = index of gate g's cluster
 * (execution order).  mode 0 runs the reference GreedyScheduler loop over
 * ClusterScheduler (_greedyscheduler.py:119-137); mode 1 the one-scan
 * lookahead packer (no reference counterpart; same commutation rule). */
int hiqsim_plan_clusters(const int64_t* gate_ids, const size_t* gate_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, size_t ngates, int cluster_size, int mode,
                         int32_t* cluster_of);

/* Applies a gate list with tile passes (no reference counterpart; the
 * reference applies one fused cluster per state pass, SimulatorMPI::Run,
 * SimulatorMPI.cpp:441-541).  Gates as in hiqsim_batch_add_circuit but on
 * the handle's qubit ids.  Each HBM pass loads tiles of index bits {0..5}
 * plus 6 planned bits into LDS and applies there every op its plan holds
 * (ops of at most op_width qubits, fused in FP64); the plan is the cluster
 * scheduler's rule with the tile as the local set.  Pending gates are run
 * first.  *npasses = HBM passes used. */
int hiqsim_apply_tiled(hiqsim_t* h, size_t ngates, const double* mats, const int64_t* tgt_ids,
                       const size_t* tgt_off, const int64_t* ctrl_ids, const size_t* ctrl_off,
                       int op_width, int* npasses);
/* The tile plan alone (no device): gates given by bit positions (CSR) on
 * n_bits bits.  pass_of[g], op_of[g] = pass and op (execution order) of
 * gate g; hpos[6 * p .. 6 * p + 6) = the planned bits of pass p;
 * *npasses set (hpos must hold 6 * ngates entries). */
int hiqsim_plan_tile_passes(const int64_t* gate_pos, const size_t* gate_off, const int64_t* ctrl_pos,
                            const size_t* ctrl_off, size_t ngates, int n_bits, int op_width,
                            int32_t* pass_of, int32_t* op_of, int32_t* hpos, size_t* npasses);

/* Process-wide kernel tuning (no reference counterpart):
 *   "tile_high_target":   dense passes whose targets all sit at or above this
 *                         bit take the direct form (default 64: never);
 *   "block_swizzle":      low block-index bits reversed in the streaming
 *                         passes of k <= 3 and the other streaming kernels
 *                         (0 = identity order, default 8, at most 16);
 *   "block_swizzle_wide": the same for dense passes of k >= 4 (default 6);
 *   "tile_mfma":          op engine of tile programs: 1 (default) ops of 3+
 *                         qubits on the matrix cores; 0 every op on the
 *                         vector units;
 *   "dense5_tile":        5-qubit passes without controls on 12+
 *                         qubits: 1 (default) as a one-op tile pass on the
 *                         matrix cores; 0 the dense pass forms.
 *   "tile_high_target_wide": the same threshold as "tile_high_target" for
 *                         4-qubit passes (default 13);
 *   "direct_block":       threads per block of the direct form of 4-qubit
 *                         passes: 256, 512 or 1024 (default);
 *   "tile_threads":       threads per workgroup of the tile-program kernels
 *                         on the matrix cores: 0 (default: 256 for complex64
 *                         and for complex128 programs with 5-qubit ops, 512
 *                         for the other complex128 programs), 256 or 512;
 *   "permute_swizzle":    low block-index bits reversed in the tiled qubit
 *                         permutation (default 8, at most 16);
 *   "permute_block":      its threads per workgroup: 256 (default) or 512.
 * Batches and request loops read them when they build or launch. */
int hiqsim_set_tuning(const char* name, int64_t value);
int hiqsim_get_tuning(const char* name, int64_t* value);

/* ---- batches of independent circuits ------------------------------------
 * No counterpart in the reference, which simulates one circuit per
 * SimulatorMPI object.  Each circuit of a batch starts from |0...0> on its
 * own qubits 0..n-1 and its final state equals what a SimulatorMPI (one
 * rank) holds after apply_controlled_gate for every gate and run()
 * (SimulatorMPI.cpp:407-436): the gates are cut into fused clusters of at
 * most k qubits each (k per qubit count, hiqsim_batch_set_plan) and applied
 * as dense passes.  All circuits of a batch run in one graph launch. */
typedef struct hiqsim_batch hiqsim_batch_t;

int hiqsim_batch_create(int precision, int device, hiqsim_batch_t** out);
int hiqsim_batch_destroy(hiqsim_batch_t* b);
/* Adds a circuit of n_qubits (1..30) qubits and ngates gates; returns its
 * index in *index.  Gate g acts on targets tgt_ids[tgt_off[g]..tgt_off[g+1])
 * with controls ctrl_ids[ctrl_off[g]..ctrl_off[g+1]) (all must be 1); its
 * 2^t x 2^t matrix follows the previous gate's in mats (interleaved
 * complex128, row-major).  A gate with its controls spans at most 5 qubits. */
int hiqsim_batch_add_circuit(hiqsim_batch_t* b, int n_qubits, size_t ngates, const double* mats,
                             const int64_t* tgt_ids, const size_t* tgt_off,
                             const int64_t* ctrl_ids, const size_t* ctrl_off, int64_t* index);
/* Appends `count` copies of circuit `index`; *first = the first copy's index. */
int hiqsim_batch_add_copies(hiqsim_batch_t* b, int64_t index, int64_t count, int64_t* first);
/* k_by_n[n] = fusion width (1..5, 0 = default 4) for circuits of n qubits;
 * circuits of more than wg_max_qubits (0..24) qubits are advanced by
 * layered many-workgroup launches, the others each by one workgroup of a
 * persistent launch.  Layered passes of circuits of at least
 * tile_min_qubits qubits may take the wave-tile form (31 = never); layered
 * circuits of at least prog_min_qubits (and 12) qubits run as tile
 * programs, several ops of at most k_by_n[n] qubits per HBM pass (31 =
 * never). */
int hiqsim_batch_set_plan(hiqsim_batch_t* b, const int32_t* k_by_n, size_t len, int wg_max_qubits,
                          int tile_min_qubits, int prog_min_qubits);
/* Plans and fuses every circuit and checks every launch's blocks and
 * addresses against its buffers on the host; needs no device (stats are
 * filled as for a build on a 256-CU device). */
int hiqsim_batch_prepare(hiqsim_batch_t* b);
/* Plans, fuses, checks, uploads and captures the launch graph. */
int hiqsim_batch_build(hiqsim_batch_t* b);
/* Runs every circuit from |0...0> (graph replay); *device_ms = device time. */
int hiqsim_batch_run(hiqsim_batch_t* b, double* device_ms);
int hiqsim_batch_num_circuits(hiqsim_batch_t* b, int64_t* n);
int hiqsim_batch_num_qubits(hiqsim_batch_t* b, int64_t index, int* n_qubits);
/* Final state of circuit `index`: 2^n interleaved complex128 values. */
int hiqsim_batch_get_state(hiqsim_batch_t* b, int64_t index, double* out);
/* out[0..13] = circuits, HBM passes, workgroup-owned circuits, LDS-resident
 * circuits, layered launches, arena bytes, matrix bytes, amplitude-passes
 * (sum over HBM passes of 2^n), persistent workgroups, widest fusion width,
 * layered launches whose 2-D grid carries padding blocks, tile-program
 * passes, ops inside them, algorithmic flops of those ops of 3+ qubits
 * (8 * 2^k * 2^n each: the matrix-core ops of the default op engine). */
int hiqsim_batch_stats(hiqsim_batch_t* b, uint64_t* out, size_t len);

/* ---- request loop (continuous batching) ---------------------------------
 * No counterpart in the reference.  Requests are circuits from |0...0> (as
 * in a batch) with an output: kind 0 = `shots` samples (Philox seed, as
 * hiqsim_sample), kind 1 = expectation values <P_t> of Pauli strings
 * P_t = prod X^x_t Z^z_t with Y on bits set in both masks (bit i = qubit i),
 * kind 2 = the final state (2^n amplitudes, index bit i = qubit i).
 * A request's state (2^n amplitudes) and fused matrices live in two device
 * pools sized at creation; requests are admitted in arrival order as the
 * pools allow and advance together, passes_per_step fused passes per step. */
typedef struct hiqsim_server hiqsim_server_t;

int hiqsim_server_create(int precision, int device, uint64_t state_bytes, uint64_t matrix_bytes,
                         int passes_per_step, int max_active, hiqsim_server_t** out);
int hiqsim_server_destroy(hiqsim_server_t* s);
/* k_by_n[n] = fusion width (1..5, 0 = default 4) for requests of n qubits. */
int hiqsim_server_set_widths(hiqsim_server_t* s, const int32_t* k_by_n, size_t len);
/* Requests of at least q (and 12) qubits run as tile programs, several fused
 * ops per HBM pass (default 12; 31 = never); for later submissions. */
int hiqsim_server_set_prog_min_qubits(hiqsim_server_t* s, int q);
/* Requests of at least q qubits ("heavy") are admitted one at a time: a
 * heavy request waits while another heavy one is active, smaller requests
 * keep being admitted beside it (default 24; 31 = no limit).  A heavy state
 * fills the GPU on its own, so running heavy requests in arrival order
 * instead of side by side finishes each one sooner, and the next one is
 * planned while it waits. */
int hiqsim_server_set_heavy_qubits(hiqsim_server_t* s, int q);
/* Gates as in hiqsim_batch_add_circuit; masks hold nterms entries (kind 1). */
int hiqsim_server_submit(hiqsim_server_t* s, int n_qubits, size_t ngates, const double* mats,
                         const int64_t* tgt_ids, const size_t* tgt_off, const int64_t* ctrl_ids,
                         const size_t* ctrl_off, int kind, uint32_t shots, uint64_t seed,
                         const uint64_t* xmask, const uint64_t* zmask, size_t nterms,
                         int64_t* id);
/* Collects finished work, admits and launches one step; with block != 0 it
 * waits for in-flight work when nothing new can be launched.  The ids of
 * the requests completed during the call go to completed[0..cap) and their
 * number to *ncompleted (cap >= max_active never truncates). */
int hiqsim_server_step(hiqsim_server_t* s, int block, int64_t* completed, size_t cap,
                       size_t* ncompleted);
int hiqsim_server_idle(hiqsim_server_t* s, int* idle);
int hiqsim_server_done(hiqsim_server_t* s, int64_t id, int* done);
/* Results of a completed request: samples (kind 0) or 2 * nterms doubles
 * (Re, Im of each <P_t>, kind 1); *n = number of samples / terms. */
int hiqsim_server_samples(hiqsim_server_t* s, int64_t id, uint64_t* out, size_t cap, size_t* n);
int hiqsim_server_expectations(hiqsim_server_t* s, int64_t id, double* out, size_t cap_terms,
                               size_t* n);
/* Final state of a completed kind-2 request as interleaved complex128
 * (complex64 states are widened); *n_amps = 2^n. */
int hiqsim_server_state(hiqsim_server_t* s, int64_t id, double* out, uint64_t cap_amps, uint64_t* n_amps);
/* out[0..2] = steady-clock seconds of submission, admission, completion. */
int hiqsim_server_times(hiqsim_server_t* s, int64_t id, double* out);
/* Drops a completed request's results. */
int hiqsim_server_forget(hiqsim_server_t* s, int64_t id);
/* out[0..5] = submitted, completed, steps, launches, peak active requests,
 * peak state-pool bytes in use. */
int hiqsim_server_stats(hiqsim_server_t* s, uint64_t* out, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* HIQSIM_AMD_H */
