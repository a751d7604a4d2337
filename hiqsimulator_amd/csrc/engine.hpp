// Single-GPU state-vector engine: the host side of the hot path.
//
// Mirrors the behaviour of the reference's SimulatorMPI at world size 1
// (src/simulator-mpi/SimulatorMPI.{hpp,cpp}): qubit bookkeeping, gate
// canonicalisation with control folding, fusion, one fused HBM pass per
// run(), measurement with the reference's block-CDF sampling and
// std::mt19937 stream, probabilities, amplitudes, collapse, deallocation.
// Errors are thrown as std::runtime_error with the reference's messages.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <complex>
#include <cstdint>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "fusion.hpp"
#include "kernels.hpp"

namespace hq {

// Timed kernel classes (for the HIP-event timing mode).
enum KernelClass : int {
     kDense1 = 0,
     kDense2,
     kDense3,
     kDense4,
     kDense5,
     kDiag,
     kScale,
     kPermute,
     kReduce,
     kNormalize,
     kRemoveBit,
     kFillCopy,
     kPauli,
     kTile,
     kNumKernelClasses
};

class Engine {
public:
     Engine(uint64_t seed, int64_t max_local, int64_t max_cluster, Prec prec, int device);
     ~Engine();
     Engine(const Engine&) = delete;
     Engine& operator=(const Engine&) = delete;

     // SimulatorMPI.cpp:168-251.
     void allocate_qubit(int64_t id);
     void allocate_qureg(const std::vector<int64_t>& ids, cplx init);
     // SimulatorMPI.cpp:278-322, 369-426.
     void deallocate_qubit(int64_t id);
     // SimulatorMPI.cpp:710-815.
     void apply_gate(Mat m, std::vector<int64_t> ids, const std::vector<int64_t>& ctrls);
     // SimulatorMPI.cpp:441-541.  Fuses the pending gates into one cluster.
     // In kPassEach mode the cluster is applied at once as one HBM pass, as
     // the reference does.  In kPassDeferred mode (the default) the gates of
     // a cluster of at most 5 qubits with its controls, on a state of 12 or
     // more qubits, are queued instead (unfused, on bit positions); queued
     // gates run as tile passes (several clusters per HBM pass, planned with
     // the tile as the local set and fused per tile op) once kQueueHigh of
     // them wait, and all of them before any call that reads the state,
     // changes the qubit set or order, or synchronizes.  The width check (at
     // most 5 qubits) still raises here.
     void run();
     enum PassMode : int { kPassEach = 0, kPassDeferred = 1 };
     void set_pass_mode(int mode);
     int pass_mode() const { return pass_mode_; }
     size_t queued_gates() const { return queue_.size(); }
     // SimulatorMPI.cpp:897-1008.
     std::vector<uint8_t> measure(const std::vector<int64_t>& ids);
     // SimulatorMPI.cpp:561-600.
     double get_probability(const std::vector<uint8_t>& bits, const std::vector<int64_t>& ids);
     // SimulatorMPI.cpp:602-650.
     cplx get_amplitude(const std::vector<uint8_t>& bits, const std::vector<int64_t>& ids);
     // SimulatorMPI.cpp:1010-1058.
     void collapse(const std::vector<int64_t>& ids, const std::vector<uint8_t>& values);
     // SimulatorMPI.cpp:681-694.
     double entropy();

     // SimulatorMPI.cpp:652-667 (relabel only, no data movement).
     const std::vector<int64_t>& qubit_ids() const { return locals_; }
     void set_qubits_perm(const std::vector<int64_t>& ids);
     // SimulatorMPI.cpp:1060-1138: at one device there are no global qubits.
     void swap_qubits(const std::vector<int64_t>& pairs);
     // Moves data so that ids[p] becomes the qubit at bit p (the in-HBM
     // replacement of the reference's local<->global qubit exchange).
     void permute_qubits(const std::vector<int64_t>& ids);

     // State transfer (cheat_local, SimulatorMPI.cpp:543-559; set_wavefunction,
     // _simulator_mpi.py:281-306).
     uint64_t state_size() const { return n_amps_; }
     Prec precision() const { return prec_; }
     void get_state_c128(std::complex<double>* out);
     // The state's device buffer after pending gates have run and the
     // engine's stream has drained.  Valid until the next call that moves the
     // state to another buffer: allocation, deallocation, permute_qubits,
     // apply_pauli_sum, restore.
     void* device_state();
     // device_state() with a pin: while at least one pin is held, results of
     // permute_qubits and apply_pauli_sum are copied back into the same
     // buffer (one extra streaming pass each) and no state buffer is freed,
     // so the pointer stays allocated until the pin is released and follows
     // the state until the qubit count changes.  unpin_state() only
     // decrements the pin count (it may be called from any thread); buffers
     // retired meanwhile are freed by a later call on the engine.
     void* pin_state();
     void unpin_state();
     int view_pins() const { return view_pins_.load(); }
     void set_state_c128(const std::complex<double>* in, uint64_t n);
     // Save/restore of the interleaved state to a file, streamed through a
     // pinned staging buffer (no host copy of the whole state).  Format:
     // "HQSV", u32 version 1, u32 precision (0 complex128, 1 complex64),
     // u32 n, n x i64 qubit ids (bit order), 2^n interleaved amplitudes.
     // load_state needs the same set of qubit ids; it adopts the file's bit
     // order (a relabel, as set_qubits_perm) and converts the precision.
     void save_state(const std::string& path);
     void load_state(const std::string& path);

     // Device-side checkpoint: snapshot() copies the state (and the qubit
     // order) into a second HBM buffer with one streaming copy; restore()
     // copies it back.  No counterpart in the reference, which can only copy
     // the state out (cheat_local).
     void snapshot();
     void restore();
     // Compares the state with the snapshot: {Re, Im of <snapshot|state>,
     // <snapshot|snapshot>, max |state - snapshot|}, accumulated in FP64.
     void compare_snapshot(double out[5]);
     // Fills the state with a normalised complex Gaussian random vector
     // (Philox4x32-10 keyed by `seed`, one counter per amplitude).
     void init_random_state(uint64_t seed);
     // `shots` computational-basis samples of the current state (index bit p
     // = the qubit at position p), shot s drawn from Philox4x32-10 with
     // counter s and key `seed`; the state is not changed.
     void sample(uint32_t shots, uint64_t seed, uint64_t* out);

     // Pauli sums; mask bit i refers to ids[i].  Y is encoded as X and Z.
     cplx expectation(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                      const std::vector<cplx>& coeff, const std::vector<int64_t>& ids);
     void apply_pauli_sum(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                          const std::vector<cplx>& coeff, const std::vector<int64_t>& ids);
     // exp(-i t H) on the subspace where every control qubit is 1, H the
     // Pauli sum (masks over ids, as above).  The reference's engine calls
     // this for ProjectQ's TimeEvolution gate (_simulator_mpi.py:469-475)
     // but its C++ never defines it.  Identity terms become a phase; the
     // rest is a Taylor series in s = ceil(|t| sum|c|) steps, each summed
     // until a term's norm drops below 1e-12 (at most 64 terms).
     void time_evolution(const std::vector<uint64_t>& xm, const std::vector<uint64_t>& zm,
                         const std::vector<cplx>& coeff, double t, const std::vector<int64_t>& ids,
                         const std::vector<int64_t>& ctrls);

     // Applies a gate list (ids as in apply_gate; matrices 2^t x 2^t with t
     // targets) with tile passes: each HBM pass loads tiles of index bits
     // {0..5} + 6 planned bits into LDS and applies every op of its plan
     // there (ops of at most op_width qubits, fused in FP64 on the host).
     // Pending gates are run first.  Below 12 qubits the ops run as plain
     // fused passes.  Returns the number of HBM passes.
     struct Gate {
          Mat m;
          std::vector<int64_t> targets, ctrls;
     };
     int apply_tiled(const std::vector<Gate>& gates, int op_width);

     // One immediate pass at explicit bit positions (bypasses fusion).
     void apply_matrix(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask,
                       bool diag);

     // Runs the queued clusters and waits for all device work.
     void synchronize();
     void set_timing(bool on);
     void kernel_time(int cls, uint64_t& count, double& total_ms);
     hipStream_t stream() const { return stream_; }
     int device() const { return device_; }

private:
     struct TimedLaunch {
          int cls;
          hipEvent_t start, stop;
     };

     // A gate on bit positions: targets t (matrix bit l <-> t[l]) and
     // controls c.
     struct PosGate {
          Mat m;
          std::vector<int64_t> t, c;
     };
     // Runs every queued cluster (tile passes).
     void settle() { drain_queue(0); }
     // Applies queued clusters as tile passes until at most `keep` remain.
     void drain_queue(size_t keep);
     // Plans `items` into tile passes and launches them while more than
     // `keep` items are left unplanned; the launched items are removed from
     // `items` (the rest keep program order).  Needs n >= kTileBits.
     // Returns the number of HBM passes launched.
     int tile_items(std::vector<PosGate>& items, int op_width, size_t keep);
     void release_buffer(void*& p);
     void collect_retired();
     // A same-size out-of-place result in d_tmp_ becomes the state: the
     // buffers swap, or (while a pin is held) the result is copied back.
     void commit_tmp();
     size_t find(int64_t id) const;
     size_t find_sure(int64_t id) const;
     void ensure_state(uint64_t n_amps);
     void ensure_tmp(uint64_t n_amps);
     void swap_tmp();
     const void* upload_matrix(const Mat& m, const void* desc = nullptr, size_t desc_bytes = 0);
     void dense_or_diag(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask,
                        bool diag);
     bool dense5_as_tile(const Mat& m, const std::vector<uint32_t>& pos, uint64_t ctrlmask);
     double masked_norm(uint64_t msk, uint64_t val);
     void mask_of(const std::vector<int64_t>& ids, const std::vector<uint8_t>& bits,
                  uint64_t& msk, uint64_t& val) const;
     void normalize(double norm, uint64_t msk, uint64_t val);
     int n_qubits() const { return static_cast<int>(locals_.size()); }
     void begin_timed(int cls);
     void end_timed();
     void read_out(double* host, size_t count);

     Prec prec_;
     int device_;
     size_t max_local_, max_cluster_, min_local_;
     std::vector<int64_t> locals_;
     Fusion fusion_;
     std::mt19937 rng_;
     std::uniform_real_distribution<double> dist_{0.0, 1.0};

     hipStream_t stream_ = nullptr;
     void* d_state_ = nullptr;
     uint64_t cap_state_ = 0;  // amplitudes
     void* d_tmp_ = nullptr;
     uint64_t cap_tmp_ = 0;
     void* d_snap_ = nullptr;
     uint64_t cap_snap_ = 0, n_snap_ = 0;
     std::vector<int64_t> snap_locals_;
     bool has_snap_ = false;
     uint64_t n_amps_ = 1;

     // Fused-matrix upload ring: pinned host slots copied into device slots.
     // A slot holds a 32 x 32 complex128 matrix and, after it, the one-op
     // tile descriptor of a 5-qubit pass run as a tile pass.
     static constexpr int kRing = 64;
     static constexpr size_t kMatBytes = 32 * 32 * 16;
     static constexpr size_t kSlotBytes = kMatBytes + 256;
     char* h_ring_ = nullptr;
     char* d_ring_ = nullptr;
     hipEvent_t ring_ev_[kRing] = {};
     bool ring_used_[kRing] = {};
     int ring_next_ = 0;

     double* d_scratch_ = nullptr;  // reductions + outputs
     double* h_out_ = nullptr;      // pinned, 16 doubles
     double* d_blocks_ = nullptr;   // measurement block sums (2^15)
     double* h_blocks_ = nullptr;   // pinned
     char* h_stage_ = nullptr;      // pinned staging for block scans
     size_t stage_bytes_ = 0;
     uint64_t* d_samples_ = nullptr;  // sample outcomes
     size_t samples_cap_ = 0;
     // Tile-program staging ring: pinned host image and device copy of
     // each launch group's ops and matrices.
     struct ProgSlot {
          char* h = nullptr;
          char* d = nullptr;
          size_t cap = 0;
          hipEvent_t ev = nullptr;
          bool busy = false;
     };
     static constexpr int kProgSlots = 3;
     ProgSlot prog_[kProgSlots];
     int prog_next_ = 0;

     // Deferred gates (kPassDeferred): matrices on bit positions.  Planning
     // over 256+ gates at a time gives nearly the passes of planning the
     // whole list (56 against 56 for the 30-qubit, 1780-gate circuit through
     // the GreedyScheduler loop; 101 when whole clusters are queued).
     static constexpr size_t kQueueHigh = 512, kQueueLow = 256;
     int pass_mode_ = kPassDeferred;
     std::vector<PosGate> queue_;

     std::atomic<int> view_pins_{0};
     std::vector<void*> retired_;

     bool timing_ = false;
     std::vector<TimedLaunch> pending_;
     std::vector<hipEvent_t> free_events_;
     uint64_t t_count_[kNumKernelClasses] = {};
     double t_ms_[kNumKernelClasses] = {};
};

}  // namespace hq
