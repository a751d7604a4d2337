// Request loop with continuous batching over device-resident states.
//
// Each request is one circuit from |0...0> on qubits 0..n-1 (planned and
// fused exactly as in a Batch) plus an output: `shots` computational-basis
// samples drawn with Philox4x32-10 under the request's seed, or the
// expectation values of a list of Pauli strings.  The loop:
//  * admits waiting requests in arrival order while their state (2^n
//    amplitudes) and fused matrices fit in the device pools (first fit, a
//    request that does not fit lets later, smaller ones go first); requests
//    of at least heavy_qubits qubits, each of which fills the GPU alone, are
//    admitted one at a time in arrival order, smaller ones beside them;
//  * launches one step at a time: every admitted request advances by up to
//    `passes_per_step` HBM passes, pass l of all of them in one launch per
//    form, so requests admitted while others run share the launches with
//    them.  A request of at least prog_min_qubits (and 12) qubits runs as
//    tile programs (several fused ops per HBM pass, applied in LDS); smaller
//    ones as fused passes;
//  * plans tile-program requests on planner threads, one tile pass at a
//    time: submit() returns at once, the request is admitted as soon as its
//    first pass is planned and each step runs the passes planned so far, so
//    planning overlaps the GPU work of the same and of other requests;
//  * when a request's last pass is in a step, its outputs are computed in
//    the same step and copied back; its memory returns to the pools at
//    once (the stream orders later use after the outputs);
//  * keeps at most kServeSlots steps in flight and collects finished steps
//    by their events.
// A request's result does not depend on what else is in flight: its passes,
// their arithmetic and its sampling stream are fixed by the request alone.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "range_alloc.hpp"

namespace hq {

// Output of a request: computational-basis samples, expectation values of
// Pauli strings, or the final state itself (2^n amplitudes).
enum class ServeKind : int { Samples = 0, Expectation = 1, State = 2 };

struct ServeStats {
     uint64_t submitted = 0, completed = 0, steps = 0, launches = 0;
     uint64_t peak_active = 0, peak_state_bytes = 0;
};

class Server {
public:
     static constexpr int kServeSlots = 2;

     Server(Prec prec, int device, uint64_t state_bytes, uint64_t matrix_bytes, int passes_per_step,
            int max_active);
     ~Server();
     Server(const Server&) = delete;
     Server& operator=(const Server&) = delete;

     void set_widths(const std::vector<int>& k_by_n);
     // Requests of at least this many qubits (and 12) run as tile programs
     // (31 = never); applies to requests submitted afterwards.
     void set_prog_min_qubits(int q);
     // Requests of at least q qubits are admitted one at a time (31 = no
     // limit; default 24).
     void set_heavy_qubits(int q);
     int64_t submit(int n, std::vector<BatchGate> gates, ServeKind kind, uint32_t shots,
                    uint64_t seed, std::vector<uint64_t> xmask, std::vector<uint64_t> zmask);
     // Collects finished steps (waiting for the oldest one if `block` and
     // nothing else can be done), admits, launches one step.  Returns the
     // ids completed during the call.
     std::vector<int64_t> step(bool block);
     bool idle() const { return waiting_.empty() && active_.empty() && inflight_.empty(); }
     bool done(int64_t id) const;
     const std::vector<uint64_t>& samples(int64_t id) const;
     const std::vector<std::complex<double>>& expectations(int64_t id) const;
     // Final state of a State request, widened to complex128.
     const std::vector<std::complex<double>>& state(int64_t id) const;
     // steady-clock seconds (CLOCK_MONOTONIC) of submit, admission, completion
     void times(int64_t id, double* t_submit, double* t_admit, double* t_done) const;
     void forget(int64_t id);
     const ServeStats& stats() const { return stats_; }

private:
     struct Pass {
          int k;
          uint8_t pos[5];
     };
     struct TPass {
          uint8_t hpos[kTileHigh];
          uint32_t op_first;  // into PlanState::tops
          uint32_t nops;
     };
     // Tile-program plan of one request, filled pass by pass by a planner
     // thread; every field but `gates`, `n`, `width` and `cancel` is guarded
     // by `m`.
     struct PlanState {
          std::vector<BatchGate> gates;
          int n = 0, width = 4;
          std::mutex m;
          std::vector<TPass> tpasses;  // planned passes
          std::vector<TileOp> tops;    // their ops, mat_off into mats (entries)
          std::vector<char> mats;      // their matrices in the state precision
          uint64_t mat_entries = 0;
          int kmax_op = 1;
          bool done = false;
          std::string error;
          std::atomic<bool> cancel{false};
     };
     struct Req {
          int n = 0;
          ServeKind kind = ServeKind::Samples;
          uint32_t shots = 0;
          uint64_t seed = 0;
          std::vector<uint64_t> xmask, zmask;
          std::vector<Pass> passes;
          std::vector<uint64_t> pass_mat;  // first matrix entry of each pass (relative)
          std::shared_ptr<PlanState> plan;  // tile programs (instead of passes)
          size_t adv = 0;                   // passes launched in the current step
          std::vector<char> mats;  // fused matrices in the state precision
          uint64_t mat_entries = 0;
          size_t next = 0;
          uint64_t state_off = 0, mat_off = 0;
          int state = 0;  // 0 waiting, 1 active, 2 finishing, 3 done
          double t_submit = 0, t_admit = 0, t_done = 0;
          std::vector<uint64_t> res_samples;
          std::vector<std::complex<double>> res_exp;
          std::vector<std::complex<double>> res_state;
     };
     struct Slot {
          char* h_desc = nullptr;
          char* d_desc = nullptr;
          size_t desc_cap = 0;
          char* h_out = nullptr;
          char* d_out = nullptr;
          size_t out_cap = 0;
          hipEvent_t ev = nullptr;
          bool busy = false;
          std::vector<std::pair<int64_t, size_t>> finishing;  // id, output offset
     };

     void harvest(bool wait_oldest, std::vector<int64_t>& completed);
     void planner_loop();
     void ensure_slot(Slot& s, size_t desc_bytes, size_t out_bytes);
     int fusion_width(int n) const;

     Prec prec_;
     int device_;
     int pps_;
     size_t max_active_;
     hipStream_t stream_ = nullptr;
     void* d_states_ = nullptr;
     void* d_mats_ = nullptr;
     double* d_sums_ = nullptr;     // block probabilities of finishing sample jobs
     double* d_scratch_ = nullptr;  // reduction partials
     size_t sums_cap_ = 0;
     RangeAlloc state_pool_, mat_pool_;
     uint64_t state_cap_amps_ = 0;
     uint64_t mat_cap_entries_ = 0;
     std::vector<int> k_by_n_;
     int prog_min_qubits_ = 12;
     int heavy_qubits_ = 24;
     std::map<int64_t, Req> reqs_;
     std::deque<int64_t> waiting_;
     std::vector<int64_t> active_;
     Slot slots_[kServeSlots];
     static constexpr int kPlanners = 2;
     std::vector<std::thread> planners_;
     std::mutex qmu_;
     std::condition_variable qcv_;
     std::deque<std::shared_ptr<PlanState>> plan_queue_;
     bool stop_ = false;
     std::deque<int> inflight_;
     int next_slot_ = 0;
     int64_t next_id_ = 0;
     ServeStats stats_;
};

}  // namespace hq
