// Host-only run of the schedulers and the batch planner, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (g++, no device code of
// its own; the batch's kernel launchers are linked but never called):
//  * ClusterScheduler / SwapScheduler / whole-circuit cluster plans on
//    random gate lists with controls and diagonal flags;
//  * tile plans, whole and one pass at a time with the remaining list read
//    back (the deferred-cluster route), on random lists up to 30 bits;
//  * Batch::prepare() (plan, fuse, lay out and bounds-check every launch)
//    on random mixed workloads under several plans, padded 2-D grids
//    included;
//  * error paths: gates beyond the register, too wide, repeated qubits;
//  * the tile ops' matrix layout (tile_mat_slot) is a bijection.
// Every plan is checked: each gate exactly once, widths within bounds, and
// every pair of gates that share a qubit kept in order.
#include <algorithm>
#include <complex>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "../../hiqsimulator_amd/csrc/batch.hpp"
#include "../../hiqsimulator_amd/csrc/sched.hpp"

#define CHECK(c)                                                                   \
     do {                                                                          \
          if (!(c)) {                                                              \
               std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c);   \
               std::exit(1);                                                       \
          }                                                                        \
     } while (0)

namespace {

using hq::GateList;

GateList random_gates(std::mt19937& rng, int nq, int ng, int maxw)
{
     GateList g;
     for (int i = 0; i < ng; ++i) {
          std::vector<int64_t> q(nq);
          for (int j = 0; j < nq; ++j) q[j] = j;
          std::shuffle(q.begin(), q.end(), rng);
          const int w = 1 + static_cast<int>(rng() % maxw);
          const int t = 1 + static_cast<int>(rng() % w);
          g.gate.emplace_back(q.begin(), q.begin() + t);
          g.ctrl.emplace_back(q.begin() + t, q.begin() + w);
          g.diag.push_back(rng() % 4 == 0);
     }
     return g;
}

uint64_t support(const GateList& g, size_t i)
{
     uint64_t m = 0;
     for (auto p: g.gate[i]) m |= uint64_t(1) << p;
     for (auto p: g.ctrl[i]) m |= uint64_t(1) << p;
     return m;
}

// order[i] = execution rank of gate i; gates sharing a qubit keep their order.
void check_order(const GateList& g, const std::vector<long>& order)
{
     for (size_t i = 0; i < g.gate.size(); ++i)
          for (size_t j = i + 1; j < g.gate.size(); ++j)
               if (support(g, i) & support(g, j)) CHECK(order[i] < order[j]);
}

void check_cluster_plan(const GateList& g, const std::vector<int>& cl, int k)
{
     CHECK(cl.size() == g.gate.size());
     const int nc = cl.empty() ? 0 : *std::max_element(cl.begin(), cl.end()) + 1;
     std::vector<uint64_t> sup(nc, 0);
     std::vector<long> order(g.gate.size());
     for (size_t i = 0; i < cl.size(); ++i) {
          CHECK(cl[i] >= 0 && cl[i] < nc);
          sup[cl[i]] |= support(g, i);
          order[i] = static_cast<long>(cl[i]) * 1000000 + static_cast<long>(i);
     }
     for (auto s: sup) CHECK(__builtin_popcountll(s) <= k);
     check_order(g, order);
}

void check_tile_plans(const GateList& g, const std::vector<hq::TilePlan>& plans, int n, int op_width)
{
     std::vector<long> order(g.gate.size(), -1);
     long rank = 0;
     for (const auto& tp: plans) {
          CHECK(static_cast<int>(tp.hpos.size()) == hq::kTileHigh);
          uint64_t tile = 63;
          for (int p: tp.hpos) {
               CHECK(p >= 6 && p < n);
               tile |= uint64_t(1) << p;
          }
          for (const auto& op: tp.ops) {
               uint64_t s = 0;
               for (int i: op) {
                    CHECK(i >= 0 && static_cast<size_t>(i) < g.gate.size() && order[i] < 0);
                    order[i] = rank++;
                    s |= support(g, i);
               }
               CHECK((s & ~tile) == 0);
               CHECK(__builtin_popcountll(s) <= op_width);
          }
     }
     for (auto o: order) CHECK(o >= 0);
     check_order(g, order);
}

void schedulers(std::mt19937& rng)
{
     for (int rep = 0; rep < 60; ++rep) {
          const int nq = 2 + static_cast<int>(rng() % 12);
          const int k = 2 + static_cast<int>(rng() % 4);
          GateList g = random_gates(rng, nq, 1 + static_cast<int>(rng() % 80), std::min(nq, k));
          std::vector<int64_t> locals(nq);
          for (int j = 0; j < nq; ++j) locals[j] = j;
          const std::vector<int> first = hq::schedule_cluster(g, locals, {}, k);
          CHECK(!first.empty());
          for (size_t j = 1; j < first.size(); ++j) CHECK(first[j - 1] < first[j]);
          for (int mode: {hq::kPlanReference, hq::kPlanLookahead})
               check_cluster_plan(g, hq::plan_clusters(g, k, mode), k);
          const int nl = std::max(k, nq - static_cast<int>(rng() % 3));
          const std::vector<int64_t> sw = hq::schedule_swap(g, 1000, nl, rng() % 2 == 0);
          CHECK(static_cast<int>(sw.size()) <= nl);
     }
}

void tile_plans(std::mt19937& rng)
{
     for (int rep = 0; rep < 40; ++rep) {
          const int n = 12 + static_cast<int>(rng() % 19);
          const int w = 2 + static_cast<int>(rng() % 4);
          GateList g = random_gates(rng, n, 1 + static_cast<int>(rng() % 300), w);
          check_tile_plans(g, hq::plan_tile_passes(g, n, hq::kTileHigh, w), n, w);
          // one pass at a time, stopping with `keep` gates left, then a new
          // planner over the remaining gates (as the deferred clusters do)
          const size_t keep = rng() % 40;
          hq::TilePassPlanner pl(g, n, hq::kTileHigh, w);
          std::vector<hq::TilePlan> plans;
          hq::TilePlan tp;
          while (pl.remaining().size() > keep && pl.next(tp)) plans.push_back(tp);
          CHECK(pl.remaining().size() <= std::max<size_t>(keep, 0) || pl.done());
          for (size_t j = 1; j < pl.remaining().size(); ++j) CHECK(pl.remaining()[j - 1] < pl.remaining()[j]);
          GateList rest;
          for (size_t i: pl.remaining()) {
               rest.gate.push_back(g.gate[i]);
               rest.ctrl.push_back(g.ctrl[i]);
               rest.diag.push_back(false);
          }
          if (!rest.gate.empty()) {
               const auto more = hq::plan_tile_passes(rest, n, hq::kTileHigh, w);
               for (auto p: more) {
                    for (auto& op: p.ops)
                         for (int& i: op) i = static_cast<int>(pl.remaining()[i]);
                    plans.push_back(p);
               }
          }
          check_tile_plans(g, plans, n, w);
     }
}

std::vector<hq::BatchGate> random_circuit(std::mt19937& rng, int n, int ngates)
{
     std::vector<hq::BatchGate> out;
     std::normal_distribution<double> nd;
     for (int i = 0; i < ngates; ++i) {
          const int w = 1 + static_cast<int>(rng() % std::min(n, 3));
          std::vector<int64_t> q(n);
          for (int j = 0; j < n; ++j) q[j] = j;
          std::shuffle(q.begin(), q.end(), rng);
          const int t = 1 + static_cast<int>(rng() % w);
          hq::BatchGate g;
          g.targets.assign(q.begin(), q.begin() + t);
          g.ctrls.assign(q.begin() + t, q.begin() + w);
          g.m.resize(size_t(1) << (2 * t));
          for (auto& a: g.m) a = std::complex<double>(nd(rng), nd(rng));
          out.push_back(std::move(g));
     }
     return out;
}

void batches(std::mt19937& rng)
{
     for (int rep = 0; rep < 6; ++rep) {
          hq::Batch b(rep % 2 ? hq::Prec::C64 : hq::Prec::C128, 0);
          const int nc = 20 + static_cast<int>(rng() % 60);
          for (int c = 0; c < nc; ++c) {
               const int n = 1 + static_cast<int>(rng() % 18);
               b.add_circuit(n, random_circuit(rng, n, static_cast<int>(rng() % 60)));
          }
          b.add_copies(0, 3);
          std::vector<int> k_by_n(19);
          for (auto& k: k_by_n) k = 1 + static_cast<int>(rng() % 5);
          const int plans[4][3] = {{16, 0, 31}, {8, 31, 31}, {11, 0, 12}, {0, 31, 12}};
          for (const auto& p: plans) {
               b.set_plan(k_by_n, p[0], p[1], p[2]);
               b.prepare();
               CHECK(b.stats().circuits == static_cast<uint64_t>(nc + 3));
          }
     }
}

template <class F>
bool throws(F&& f)
{
     try {
          f();
     } catch (const std::exception&) {
          return true;
     }
     return false;
}

// The tile ops' matrix layout: a bijection of the 4^k entries for every
// width, row major below 3 qubits, and each 64-entry operand block of the
// 4-5-qubit layout holding one column quartet of 16 consecutive rows.
void matrix_layouts()
{
     for (int k = 1; k <= 5; ++k) {
          const uint32_t n = 1u << (2 * k), D = 1u << k;
          std::vector<int> seen(n, 0);
          for (uint32_t i = 0; i < n; ++i) {
               const uint32_t s = hq::tile_mat_slot(k, i);
               CHECK(s < n);
               ++seen[s];
               if (k <= 2) CHECK(s == i);
               if (k >= 4) {
                    const uint32_t o = i / D, j = i % D;
                    CHECK(s / 64 == (o / 16) * (D / 4) + j / 4);
                    CHECK(s % 64 == (j % 4) * 16 + o % 16);
               }
          }
          for (uint32_t s = 0; s < n; ++s) CHECK(seen[s] == 1);
     }
}

void errors()
{
     hq::Batch b(hq::Prec::C128, 0);
     hq::BatchGate g;
     g.m.assign(4, 0.0);
     g.targets = {5};
     CHECK(throws([&] { b.add_circuit(3, {g}); }));  // beyond the register
     g.targets = {0};
     g.ctrls = {0};
     CHECK(throws([&] { b.add_circuit(3, {g}); }));  // repeated qubit
     g.ctrls = {1, 2, 3, 4, 5};
     CHECK(throws([&] { b.add_circuit(8, {g}); }));  // six qubits
     GateList gl;
     gl.gate.push_back({0, 40});
     gl.ctrl.push_back({});
     gl.diag.push_back(false);
     CHECK(throws([&] { hq::plan_tile_passes(gl, 30, hq::kTileHigh, 4); }));  // position out of range
}

}  // namespace

// The batch's kernel launchers (batch_kernels.hip) and knob readers
// (kernels.hip): prepare() never launches, so these stand-ins only satisfy
// the linker of this host-only program and fail loudly if reached.
namespace hq {
#define NEVER(...)                                           \
     {                                                       \
          std::fprintf(stderr, "device launch in a host test\n"); \
          std::abort();                                      \
     }
hipError_t launch_batch_tprog(void*, Prec, const void*, const BatchTileProg*, int, const TileOp*, int, uint64_t,
                              hipStream_t) NEVER()
hipError_t launch_batch_wg(void*, Prec, const void*, const BatchPass*, const BatchCirc*, const uint32_t*,
                           const uint32_t*, int, int, hipStream_t) NEVER()
hipError_t launch_batch_layer(void*, Prec, const void*, const BatchItem*, int, int, uint64_t, hipStream_t) NEVER()
hipError_t launch_batch_tile(void*, Prec, const void*, const BatchTileItem*, int, int, uint64_t,
                             hipStream_t) NEVER()
hipError_t launch_set_ones(void*, Prec, const uint64_t*, int, hipStream_t) NEVER()
hipError_t launch_fill(void*, Prec, uint64_t, double, double, hipStream_t) NEVER()
int batch_wg_threads(Prec prec, int kmax) { return prec == Prec::C128 && kmax >= 5 ? 256 : 512; }
int get_tile_high_target_for(int k) { return k == 4 ? 13 : 64; }
int get_block_swizzle() { return 8; }
int get_block_swizzle_for(int k) { return k >= 4 ? 6 : 8; }
}  // namespace hq

int main()
{
     std::mt19937 rng(2024);
     schedulers(rng);
     tile_plans(rng);
     batches(rng);
     errors();
     matrix_layouts();
     std::printf("host plans ok\n");
     return 0;
}
