// Gate schedulers (host, integer bitmask work).
//
// ClusterScheduler picks the next fused cluster; SwapScheduler picks the next
// stage's local qubit set.  Their selection rules are those of the
// reference's src/scheduler (cluster_scheduler.{h,cpp}, swap_scheduler.{h,cpp},
// convertors.cpp), kept identical because the Python GreedyScheduler drives
// them through the same interface (_sched_cpp.cpp:28-40,
// _greedyscheduler.py:119-193) and the order of the returned gate indices is
// observable:
//   * qubit ids are renumbered densely in ascending order (< 64 qubits);
//   * a gate can join a cluster if it touches no qubit of an earlier skipped
//     gate and its local qubits lie in the cluster; a non-diagonal gate must
//     be fully local;
//   * the best cluster maximises the number of takeable gates, ties go to
//     fewer qubits, then to the first candidate in enumeration order.
// The candidate enumeration (supersets of each gate's local set, adding bits
// in ascending order, each (cluster, next bit) pair expanded once) is kept;
// the memo is an open-addressing table instead of std::map and a candidate's
// scan stops once every cluster qubit is blocked and no later gate could
// still be taken, which leaves every result unchanged.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hq {

using msk_t = uint64_t;

struct GateList {
     std::vector<std::vector<int64_t>> gate;
     std::vector<std::vector<int64_t>> ctrl;
     std::vector<bool> diag;
};

// cluster_scheduler.cpp:42-78.  Returns indices into the gate list.
std::vector<int> schedule_cluster(const GateList& g, const std::vector<int64_t>& locals,
                                  const std::vector<int64_t>& globals, int cluster_size);

// Whole-circuit cluster plans.  Returns, for every gate, the index of the
// cluster it belongs to; clusters are numbered in execution order and a
// cluster's gates run in program order.
//  * kPlanReference: the GreedyScheduler loop (_greedyscheduler.py:119-137)
//    calling schedule_cluster on the remaining gates until none are left;
//  * kPlanLookahead: one scan per cluster: starting from the first remaining
//    gate, take every later gate that fits in k qubits and touches no qubit
//    of a gate skipped so far (the same commutation rule), stop when the
//    cluster is full and blocked.  O(gates) per cluster.
enum PlanMode : int { kPlanReference = 0, kPlanLookahead = 1 };
std::vector<int> plan_clusters(const GateList& g, int cluster_size, int mode);

// Tile-capacity planning (the cluster scheduler retargeted from MPI rank
// locality to the LDS tile of one workgroup): ids are index bit positions.
// Each pass holds the gates whose bits all lie in {0..5} + hpos, where hpos
// are `tile_high` positions >= 6 chosen to take as many gates as possible
// from the front of the remaining list (same commutation rule as
// ClusterScheduler: a gate is taken if it touches no bit of a gate skipped
// before it).  The taken gates are cut into ops of at most op_width bits
// (lookahead packing); every op becomes one fused matrix applied in LDS.
struct TilePlan {
     std::vector<int> hpos;              // tile_high ascending positions in [6, n)
     std::vector<std::vector<int>> ops;  // gate indices of each op, program order
};
std::vector<TilePlan> plan_tile_passes(const GateList& g, int n, int tile_high, int op_width);

// The same plan one tile pass at a time (plan_tile_passes is its loop), so
// that a caller can start on the first passes while later ones are planned.
class TilePassPlanner {
public:
     TilePassPlanner(const GateList& g, int n, int tile_high, int op_width);
     // The next tile pass into `out`; false when every gate is planned.
     bool next(TilePlan& out);
     bool done() const { return remaining_.empty(); }
     // Indices of the gates not planned yet, in program order.
     const std::vector<size_t>& remaining() const { return remaining_; }

private:
     int score(uint64_t S) const;
     GateList g_;
     int n_, tile_high_, op_width_;
     std::vector<uint64_t> q_;
     std::vector<size_t> remaining_;
};

// swap_scheduler.cpp:70-88.  Returns the qubit ids to be local.
std::vector<int64_t> schedule_swap(const GateList& g, int64_t num_splits, int num_locals,
                                   bool fuse);

}  // namespace hq
